// frank_wolfe.hip -- Frank-Wolfe step whose LP direction is solved by
// Chambolle-Pock (PDHG), reusing the MFMA convolution / correlation tiles.
//
// Reference: src/kernels.cuh:382-532 (kFrankWolfe), launch src/main.cu:170-178.
// For a (sum 1) and M = max a*a the LP  max sum g  s.t. (a*g)_k <= M, g >= 0
// is solved as the saddle point max_g min_{lambda>=0} 1'g - lambda'(A g - M)
// with tau = 0.9/n, sigma = 0.9 n (|A|_2 <= sum a = 1):
//   lambda <- [lambda + sigma (a * gbar - M)]_+      (Toeplitz tiles, fused)
//   g_new  <- [g + tau (1 - A' lambda)]_+            (Hankel tiles, fused)
//   gbar   <- 2 g_new - g
// then g is normalised and a grid line search picks t = tMax*s/nT.
//
// LDS (n = 8400: 144 KB): a and gbar in the padded segment layout, lambda in
// the normal layout.  g lives in registers: the wave that owns a correlation
// tile owns those 256 heights for the whole solve, so the primal update,
// the extrapolation and the write of gbar are all in the tile epilogue.  The
// dual update is likewise fused into the Toeplitz-tile epilogue; Toeplitz
// tiles (trapezoidal costs) are dealt to waves longest-first.  lambda is zero
// on every slack constraint: the dual epilogue flags its non-zero 16-entry
// segments (one ballot and one LDS atomicOr per tile) and the primal
// correlation skips the Hankel steps that would read only zeros; two flag
// arrays alternate between iterations, so one is cleared while the other is
// in use.
#include <mutex>

#include "common.hpp"
#include "runtime.hpp"

namespace sidon {
namespace {

constexpr int kMaxConvTiles = 80;   // n <= 10240

// Matrix block steps (four MFMAs, 8192 FLOP a wave) the Frank-Wolfe kernel has
// run since the last sidon_fw_block_steps(reset): Toeplitz steps of the dual
// updates, the M and line-search autoconvolutions, and the correlation steps
// actually executed (steps over all-zero multipliers are skipped).
__device__ unsigned long long g_fw_block_steps = 0ull;
__device__ unsigned long long g_fw_block_steps_taken = 0ull;

// Reads the counter (and with `reset` zeroes it) in one device atomic.
__global__ void k_take_fw_block_steps(int reset) {
  g_fw_block_steps_taken =
      reset ? atomicExch(&g_fw_block_steps, 0ull) : atomicAdd(&g_fw_block_steps, 0ull);
}

template <int BD, int SS = kSegStride>
__host__ __device__ inline long long fw_lds_bytes(int n) {
  const long long f = 2LL * pad_len<SS>(n) + cbuf_len(n) + 2 * (BD / kWave);
  return f * 4 + 4LL * kMaxConvTiles + 4LL * 2 * segmask_words(n);
}

struct FwArgs {
  float* pop;
  const int* skip;
  int n, rounds, cpIters, nT;
  float tMax;
};

// The correlation tile a wave owns in its slot u (-1: none).  A tile's
// Hankel work grows with its index at the multipliers this solve produces
// (measured: from 82 to 175 executed steps across n = 4200's tiles), so the
// tiles are dealt from the last to the first in boustrophedon rounds over the
// waves: a wave's tiles pair a costly one with a cheap one, and waves w,
// w + 4, ... (one SIMD's) end up with similar totals.
__device__ __forceinline__ int fw_owned_tile(int wave, int u, int nw, int ntg) {
  const int k = u * nw + ((u & 1) ? nw - 1 - wave : wave);
  return k < ntg ? ntg - 1 - k : -1;
}

// GT: correlation tiles owned per wave (ceil(corr_tiles / waves)).
template <int BD, int GT, int SS>
__global__ __launch_bounds__(BD) void k_frank_wolfe(FwArgs A) {
  constexpr int NW = BD / kWave;
  extern __shared__ __align__(16) unsigned char smem[];
  const int cand = blockIdx.x;
  if (A.skip && A.skip[cand]) return;
  const int n = A.n;
  float* a = A.pop + (size_t)cand * n;
  float* sA = reinterpret_cast<float*>(smem);
  float* sH = sA + pad_len<SS>(n);     // gbar during PDHG, h during the line search
  float* sL = sH + pad_len<SS>(n);     // lambda / autoconvolution
  float* slots = sL + cbuf_len(n);
  int* owner = reinterpret_cast<int*>(slots + 2 * NW);  // Toeplitz tile -> wave
  const int SW = segmask_words(n);
  unsigned* lmask = reinterpret_cast<unsigned*>(owner + kMaxConvTiles);  // 2 x SW: non-zero lambda
  const int tid = threadIdx.x, wave = wave_id();
  const int lane = tid & (kWave - 1), t = lane & 15, kk = lane >> 4;
  const int LA = pad_len<SS>(n), LC = cbuf_len(n);
  const int N16 = nseg16(n), NTc = conv_tiles(n), NTg = corr_tiles(n);
  const int nb = 2 * n - 1;
  int U = 0;  // block steps of one full autoconvolution
  for (int j = 0; j < NTc; ++j) U += conv_cost(j, N16);
  unsigned long long ran = 0ull;  // block steps this wave ran (wave 0 also counts conv_mfma)

  for (int i = tid; i < LA; i += BD) { sA[i] = 0.0f; sH[i] = 0.0f; }
  for (int i = tid; i < LC; i += BD) sL[i] = 0.0f;
  if (tid == 0) {
    // longest-processing-time deal of Toeplitz tiles: each to the SIMD with
    // the least work (waves w, w + 4, w + 8, ... of a workgroup share SIMD
    // w % 4 and its matrix core), then to that SIMD's least-loaded wave
    int load[NW], sload[4] = {0, 0, 0, 0};
    for (int w = 0; w < NW; ++w) load[w] = 0;
    for (int j = 0; j < NTc; ++j) owner[j] = -1;
    for (int r = 0; r < NTc; ++r) {
      int best = -1, bc = -1;
      for (int j = 0; j < NTc; ++j)
        if (owner[j] < 0 && conv_cost(j, N16) > bc) { bc = conv_cost(j, N16); best = j; }
      int smin = 0;
      for (int q = 1; q < 4; ++q) if (sload[q] < sload[smin]) smin = q;
      int wmin = smin;
      for (int w = smin + 4; w < NW; w += 4) if (load[w] < load[wmin]) wmin = w;
      owner[best] = wmin;
      load[wmin] += bc;
      sload[smin] += bc;
    }
  }
  __syncthreads();
  float sl = 0.0f;
  for (int i = tid; i < n; i += BD) { const float x = a[i]; pad_put<SS>(sA, i, x); sl += x; }
  const float s = block_sum<float, NW>(sl, slots);
  if (!(s > 0.0f)) return;
  {
    const float inv = 1.0f / s;
    for (int i = tid; i < n; i += BD) pad_put<SS>(sA, i, sA[pad_addr<SS>(i)] * inv);
  }
  __syncthreads();

  const float tau = 0.9f / (float)n, sig = 0.9f * (float)n;
  float g[GT][4];

  for (int r = 0; r < A.rounds; ++r) {
    conv_mfma<float, SS>(sA, sA, sL, n, wave, NW);
    if (wave == 0) ran += U;
    __syncthreads();
    float loc = 0.0f;
    for (int k = tid; k < nb; k += BD) loc = sL[k] > loc ? sL[k] : loc;
    const float M = block_max<float, NW>(loc, slots + NW);
    if (!(M > 0.0f)) break;

    // g = gbar = a, lambda = 0
#pragma unroll
    for (int u = 0; u < GT; ++u) {
      const int j = fw_owned_tile(wave, u, NW, NTg);
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int i = 16 * (16 * j + t) + 4 * kk + q;
        g[u][q] = (j >= 0 && i < n) ? sA[pad_addr<SS>(i)] : 0.0f;
      }
    }
    for (int i = tid; i < n; i += BD) pad_put<SS>(sH, i, sA[pad_addr<SS>(i)]);
    for (int k = tid; k < LC; k += BD) sL[k] = 0.0f;
    for (int w = tid; w < 2 * SW; w += BD) lmask[w] = 0u;
    __syncthreads();

    for (int it = 0; it < A.cpIters; ++it) {
      unsigned* mcur = lmask + (it & 1) * SW;
      unsigned* mnext = lmask + ((it + 1) & 1) * SW;
      for (int w = tid; w < SW; w += BD) mnext[w] = 0u;  // last read by iteration it-1
      // dual: lambda_k <- [lambda_k + sigma ((a * gbar)_k - M)]_+
      for (int j = 0; j < NTc; ++j) {
        if (owner[j] != wave) continue;
        int dlo, dhi;
        conv_drange(j, N16, dlo, dhi);
        const f32x4 acc = conv_tile<float, Pipe<float>::kDepth, SS>(sA, sH, 16 * j, dlo, dhi + 1);
        ran += dhi + 1 - dlo;
        const int k0 = 16 * (16 * j + t) + 4 * kk;
        bool nz = false;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const int k = k0 + q;
          if (k < nb) {
            float l = sL[k] + sig * (acc[q] - M);
            l = l > 0.0f ? l : 0.0f;
            sL[k] = l;
            nz = nz || l != 0.0f;
          }
        }
        // segment 16j + t is non-zero if any of its lanes (t, t+16, t+32, t+48) saw one
        const unsigned long long bal = __ballot(nz);
        const unsigned seg = (unsigned)((bal | (bal >> 16) | (bal >> 32) | (bal >> 48)) & 0xFFFFull);
        if (seg && lane == 0) atomicOr(mcur + (j >> 1), seg << (16 * (j & 1)));
      }
      __syncthreads();
      // primal: g_new <- [g + tau (1 - A' lambda)]_+, gbar <- 2 g_new - g
#pragma unroll
      for (int u = 0; u < GT; ++u) {
        const int j = fw_owned_tile(wave, u, NW, NTg);
        if (j < 0) continue;
        const int I0 = 16 * j;
        int steps = 0;
        const f32x4 c = corr_tile_sparse<float, Pipe<float>::kDepth, SS>(sL, sA, I0, I0, I0 + N16 + 15, mcur, steps);
        ran += steps;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const int i = 16 * (I0 + t) + 4 * kk + q;
          if (i < n) {
            float gn = g[u][q] + tau * (1.0f - c[q]);
            gn = gn > 0.0f ? gn : 0.0f;
            pad_put<SS>(sH, i, 2.0f * gn - g[u][q]);
            g[u][q] = gn;
          }
        }
      }
      __syncthreads();
    }

    // normalise g
    float sgl = 0.0f;
#pragma unroll
    for (int u = 0; u < GT; ++u)
#pragma unroll
      for (int q = 0; q < 4; ++q) sgl += g[u][q];
    const float sg = block_sum<float, NW>(sgl, slots);
    if (!(sg > 0.0f)) break;
    const float ig = 1.0f / sg;
#pragma unroll
    for (int u = 0; u < GT; ++u)
#pragma unroll
      for (int q = 0; q < 4; ++q) g[u][q] *= ig;

    // line search over t = tMax * s / nT (sum h = 1 for every t)
    float bestT = 0.0f, bestM = M;
    for (int st = 1; st <= A.nT; ++st) {
      const float tt = A.tMax * (float)st / (float)A.nT;
#pragma unroll
      for (int u = 0; u < GT; ++u) {
        const int j = fw_owned_tile(wave, u, NW, NTg);
        if (j < 0) continue;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const int i = 16 * (16 * j + t) + 4 * kk + q;
          if (i < n) pad_put<SS>(sH, i, (1.0f - tt) * sA[pad_addr<SS>(i)] + tt * g[u][q]);
        }
      }
      for (int k = tid; k < LC; k += BD) sL[k] = 0.0f;
      __syncthreads();
      conv_mfma<float, SS>(sH, sH, sL, n, wave, NW);
      if (wave == 0) ran += U;
      __syncthreads();
      float lc = 0.0f;
      for (int k = tid; k < nb; k += BD) lc = sL[k] > lc ? sL[k] : lc;
      const float mm = block_max<float, NW>(lc, (st & 1) ? slots + NW : slots);
      if (mm < bestM) { bestM = mm; bestT = tt; }
    }
    if (!(bestT > 0.0f)) break;
#pragma unroll
    for (int u = 0; u < GT; ++u) {
      const int j = fw_owned_tile(wave, u, NW, NTg);
      if (j < 0) continue;
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int i = 16 * (16 * j + t) + 4 * kk + q;
        if (i < n) pad_put<SS>(sA, i, (1.0f - bestT) * sA[pad_addr<SS>(i)] + bestT * g[u][q]);
      }
    }
    for (int k = tid; k < LC; k += BD) sL[k] = 0.0f;
    __syncthreads();
  }
  __syncthreads();
  for (int i = tid; i < n; i += BD) a[i] = sA[pad_addr<SS>(i)];
  if (lane == 0 && ran) atomicAdd(&g_fw_block_steps, ran);
}

template <int BD, int GT, int SS>
int launch_fw_layout(const FwArgs& A, int P, hipStream_t stream) {
  int rc = SIDON_OK;
  DeviceState* ds = device_state(&rc);
  if (!ds) return rc;
  const long long lds = fw_lds_bytes<BD, SS>(A.n);
  if (lds > ds->lds_max) {  // first: a length that does not fit is SIDON_ERR_LDS
    set_last_error_msg("frank_wolfe: n needs more LDS than a workgroup may use");
    return SIDON_ERR_LDS;
  }
  if (corr_tiles(A.n) > GT * (BD / kWave) || conv_tiles(A.n) > kMaxConvTiles)
    return SIDON_ERR_UNSUPPORTED;
  auto kfn = k_frank_wolfe<BD, GT, SS>;
  SIDON_HIP_TRY(hipFuncSetAttribute(reinterpret_cast<const void*>(kfn),
                                    hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  hipLaunchKernelGGL(kfn, dim3(P), dim3(BD), (size_t)lds, stream, A);
  SIDON_HIP_TRY(hipGetLastError());
  return SIDON_OK;
}

// The wide (bank-conflict-free) segment layout where both padded vectors fit
// the LDS with it (n up to about 7900), else the narrow one (to 8720), else
// the plain one (to 9856).
template <int BD, int GT>
int launch_fw(const FwArgs& A, int P, hipStream_t stream) {
  int rc = SIDON_OK;
  DeviceState* ds = device_state(&rc);
  if (!ds) return rc;
  if (fw_lds_bytes<BD, kSegStrideWide>(A.n) <= ds->lds_max)
    return launch_fw_layout<BD, GT, kSegStrideWide>(A, P, stream);
  if (fw_lds_bytes<BD, kSegStride>(A.n) <= ds->lds_max)
    return launch_fw_layout<BD, GT, kSegStride>(A, P, stream);
  // the plain layout: same time where both fit (n = 1050-8400, bit-identical
  // results), and it holds the two padded vectors to n = 9900
  return launch_fw_layout<BD, GT, kSeg>(A, P, stream);
}

}  // namespace
}  // namespace sidon

using namespace sidon;

// the layout a 160 KiB workgroup takes
template <int BD>
long long fw_lds_choice(int n) {
  const long long wide = fw_lds_bytes<BD, kSegStrideWide>(n);
  if (wide <= 160LL * 1024) return wide;
  const long long narrow = fw_lds_bytes<BD>(n);
  return narrow <= 160LL * 1024 ? narrow : fw_lds_bytes<BD, kSeg>(n);
}

extern "C" long long sidon_fw_lds_bytes_impl(int n, int block) {
  if (n < 1) return -1;
  switch (block) {
    case 256: return fw_lds_choice<256>(n);
    case 512: return fw_lds_choice<512>(n);
    case 1024: return fw_lds_choice<1024>(n);
    default: return -1;
  }
}

extern "C" int sidon_fw_block_steps(unsigned long long* out, int reset) {
  if (!out) return SIDON_ERR_ARG;
  int rc = SIDON_OK;
  if (!device_state(&rc)) return rc;
  static std::mutex mu;  // one reader of g_fw_block_steps_taken at a time
  std::lock_guard<std::mutex> lk(mu);
  hipLaunchKernelGGL(k_take_fw_block_steps, dim3(1), dim3(1), 0, 0, reset);
  SIDON_HIP_TRY(hipGetLastError());
  SIDON_HIP_TRY(hipMemcpyFromSymbol(out, HIP_SYMBOL(g_fw_block_steps_taken), sizeof(*out)));
  return SIDON_OK;
}

extern "C" int sidon_frank_wolfe(float* pop, const int* skip, int n, int P, int rounds,
                                 int cpIters, int nT, float tMax, hipStream_t stream) {
  if (!pop || n < 1 || P < 0 || rounds < 0 || cpIters < 0 || nT < 1 || !(tMax > 0.0f))
    return SIDON_ERR_ARG;
  if (P == 0 || rounds == 0) return SIDON_OK;
  FwArgs A{pop, skip, n, rounds, cpIters, nT, tMax};
  if (n >= 2048) return launch_fw<1024, 3>(A, P, stream);
  if (n >= 512) return launch_fw<512, 2>(A, P, stream);
  return launch_fw<256, 2>(A, P, stream);
}
