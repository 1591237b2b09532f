// polish.hip -- fp64 local search of a few candidates at any length, spread
// over the whole device: the fp64 polish of the best vectors.
//
// Reference: src/main.cu:798-846 (fp64 polish: kLocalSearch<double> over the
// elite, passes of rising exponent) and src/kernels.cuh:177-289 (the step).
// The fused fp64 local search keeps a, b/c and G in one workgroup's LDS and so
// stops at n = 4560 (3n doubles do not fit 160 KB at n = 8400).  Here one
// step is six kernels over global memory (L2-resident: ~3n doubles per
// candidate), and both products are cut into work units that every CU can take:
//
//   unit (candidate, tile q of 16 outputs, chunk cc of 16 blocks of 16 i)
//     = one lane, 16 trips of 16 x 16 fp64 FMAs, partial sums to `part`;
//
// a workgroup holds 256 consecutive tiles of one chunk, so every lane runs the
// same 16 trips (products outside a tile's range read zero padding), the
// chunk's 256 heights are broadcast and the other operand's window sits in
// LDS with one unused slot per 16 (lanes one tile apart read 17 doubles apart).
// The next kernel adds a tile's chunk partials in chunk order, so the sums are
// deterministic.  Step: conv units -> conv reduce (b, M) -> weights (c, R
// partials) -> corr units -> step (G, R, Adam or mirror step, clamp >= 0,
// S partials) -> renormalise, the same arithmetic as the fused kernel's step
// in fp64.
#include <cmath>

#include "common.hpp"
#include "runtime.hpp"

namespace sidon {
namespace {

constexpr int PU = 256;     // tiles (lanes) per unit workgroup
constexpr int CH = 16;      // 16-blocks per chunk
constexpr int PADA = 576;   // zero heights each side of the padded copy of a
constexpr int RB = 256;     // threads per reduce / weights workgroup

struct Shape {
  int n, NJ, NTc, NCc, NTg, NCg, LA, LB;
};

__host__ __device__ inline Shape shape_of(int n) {
  Shape s;
  s.n = n;
  s.NJ = (n + 15) / 16;                  // 16-blocks of a
  s.NTc = (2 * n - 1 + 15) / 16;         // output tiles of b
  s.NCc = (s.NJ + CH - 1) / CH;          // chunks of the conv reduction
  s.NTg = s.NJ;                          // output tiles of G
  s.NCg = s.NCc;                         // chunks of the corr reduction
  s.LA = n + 2 * PADA;                   // padded copy of a
  s.LB = 16 * (s.NTg + CH * s.NCg) + 64; // b / c, zero beyond 2n - 1
  return s;
}

// workspace layout (doubles unless noted), per candidate:
//   apad[LA], bc[LB], part[NCc * NTc * 16], gpart[NCg * NTg * 16], mom[n], vel[n],
//   rpart[ceil(LB / RB)], mpart[ceil(LB / RB)], spart[ceil(n / RB)], bias[2] (beta1^t,
//   beta2^t), then K u64 max bits, K ints: step counter, stopped.
struct Ws {
  double *apad, *bc, *part, *gpart, *mom, *vel, *rpart, *mpart, *spart, *bias;
  unsigned long long* mbits;
  int *step, *stopped;
  long long sa, sb, sp, sg, sn, sr, ss;  // per-candidate strides
};

__host__ __device__ inline int rblocks(const Shape& s) { return (s.LB + RB - 1) / RB; }
__host__ __device__ inline int sblocks(const Shape& s) { return (s.n + RB - 1) / RB; }

__host__ __device__ inline size_t ws_bytes(int n, int K) {
  const Shape s = shape_of(n);
  const size_t per = (size_t)s.LA + s.LB + (size_t)s.NCc * s.NTc * 16 + (size_t)s.NCg * s.NTg * 16 +
                     2 * (size_t)n + 2 * rblocks(s) + sblocks(s) + 2;
  return per * K * sizeof(double) + K * (sizeof(unsigned long long) + 2 * sizeof(int)) + 256;
}

__host__ __device__ inline Ws ws_of(void* base, int n, int K) {
  const Shape s = shape_of(n);
  Ws w;
  w.sa = s.LA;
  w.sb = s.LB;
  w.sp = (long long)s.NCc * s.NTc * 16;
  w.sg = (long long)s.NCg * s.NTg * 16;
  w.sn = n;
  w.sr = rblocks(s);
  w.ss = sblocks(s);
  double* d = reinterpret_cast<double*>(base);
  w.apad = d; d += w.sa * K;
  w.bc = d; d += w.sb * K;
  w.part = d; d += w.sp * K;
  w.gpart = d; d += w.sg * K;
  w.mom = d; d += w.sn * K;
  w.vel = d; d += w.sn * K;
  w.rpart = d; d += w.sr * K;
  w.mpart = d; d += w.sr * K;
  w.spart = d; d += w.ss * K;
  w.bias = d; d += 2 * K;
  w.mbits = reinterpret_cast<unsigned long long*>(d);
  w.step = reinterpret_cast<int*>(w.mbits + K);
  w.stopped = w.step + K;
  return w;
}

struct Sched {
  int iters;
  double pStart, lnPr, lrStart, lnLr, beta1, beta2;
};

// p and lr of step `it`: the fused kernel's geometric anneal
__device__ __forceinline__ void sched_at(const Sched& S, int it, double& p, double& lr) {
  const double frac = S.iters > 1 ? (double)it / (double)(S.iters - 1) : 0.0;
  p = S.pStart * exp(S.lnPr * frac);
  lr = S.lrStart * exp(S.lnLr * frac);
}

// LDS slot of window element e (one unused slot after every 16)
__device__ __forceinline__ int wslot(int e) { return e + (e >> 4); }

// sum of p[c * stride] for c = c0 .. c1 in ascending c, loaded eight at a time
// so the loads are in flight together (absent terms add an exact zero)
__device__ __forceinline__ double chunk_sum(const double* p, long long stride, int c0, int c1) {
  double r = 0.0;
  for (int c = c0; c <= c1; c += 8) {
    double v[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) v[u] = c + u <= c1 ? p[(long long)(c + u) * stride] : 0.0;
#pragma unroll
    for (int u = 0; u < 8; ++u) r += v[u];
  }
  return r;
}

// ---------------------------------------------------------------- setup --
__global__ void k_polish_setup(const double* __restrict__ pop, int n, Ws w) {
  const int cand = blockIdx.y;
  double* ap = w.apad + w.sa * cand;
  const int LA = n + 2 * PADA;
  for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < LA; e += gridDim.x * blockDim.x) {
    const int i = e - PADA;
    ap[e] = (i >= 0 && i < n) ? pop[(size_t)cand * n + i] : 0.0;
  }
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    w.mom[w.sn * cand + i] = 0.0;
    w.vel[w.sn * cand + i] = 0.0;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    w.mbits[cand] = 0ull;
    w.step[cand] = 0;
    w.stopped[cand] = 0;
    w.bias[2 * cand] = 1.0;
    w.bias[2 * cand + 1] = 1.0;
  }
}

// --------------------------------------------------------- conv units ----
// part[cand][cc][q][t] = sum over i in chunk cc of a_i a_{16q+t-i}.  Tiles of
// chunk cc whose range of i meets the chunk are [qlo, qhi]; workgroup
// blockIdx.x takes 256 of them.
__device__ __forceinline__ void conv_tile_range(int n, int q, int& J0, int& J1) {
  const int k0 = 16 * q;
  J0 = (k0 - (n - 1) > 0 ? k0 - (n - 1) : 0) >> 4;
  J1 = (k0 + 15 < n - 1 ? k0 + 15 : n - 1) >> 4;
}

__device__ void conv_chunk_tiles(const Shape& s, int cc, int& qlo, int& qhi) {
  // J1(q) and J0(q) do not decrease with q: binary searches
  int lo = 0, hi = s.NTc;  // first q with J1(q) >= CH cc
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    int a, b;
    conv_tile_range(s.n, mid, a, b);
    if (b >= CH * cc) hi = mid; else lo = mid + 1;
  }
  qlo = lo;
  lo = 0;
  hi = s.NTc;  // first q with J0(q) > CH cc + CH - 1
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    int a, b;
    conv_tile_range(s.n, mid, a, b);
    if (a > CH * cc + CH - 1) hi = mid; else lo = mid + 1;
  }
  qhi = lo - 1;
}

__global__ __launch_bounds__(PU) void k_polish_conv_units(Shape s, Ws w) {
  __shared__ double sx[CH * 16 + CH];             // the chunk's 16 CH heights (padded)
  __shared__ double sw[(16 * (PU + CH) + 32) * 17 / 16 + 16];  // a window (padded)
  const int cand = blockIdx.z, cc = blockIdx.y;
  int qlo, qhi;
  conv_chunk_tiles(s, cc, qlo, qhi);
  const int qa = qlo + PU * blockIdx.x;
  if (qa > qhi) return;
  const int qb = qa + PU - 1 < qhi ? qa + PU - 1 : qhi;
  const double* ap = w.apad + w.sa * cand + PADA;  // ap[i] = a_i, zero for -PADA <= i < n + PADA
  const int J0c = CH * cc;
  for (int e = threadIdx.x; e < CH * 16; e += PU) sx[wslot(e)] = ap[16 * J0c + e];
  // window: a_x for x in [16 (qa - J0c - CH + 1) - 15, 16 (qb - J0c) + 15]
  const int wlo = 16 * (qa - J0c - (CH - 1)) - 15;
  const int wlen = 16 * (qb - qa + CH) + 15;
  for (int e = threadIdx.x; e < wlen; e += PU) sw[wslot(e)] = ap[wlo + e];
  __syncthreads();
  const int q = qa + (int)threadIdx.x;
  if (q > qb) return;
  double acc[16];
#pragma unroll
  for (int t = 0; t < 16; ++t) acc[t] = 0.0;
  for (int jj = 0; jj < CH; ++jj) {
    double x[16], v[31];
#pragma unroll
    for (int u = 0; u < 16; ++u) x[u] = sx[wslot(16 * jj + u)];
    const int base = 16 * (q - J0c - jj) - 15 - wlo;  // window offset of a_{16(q-J)-15}
#pragma unroll
    for (int m = 0; m < 31; ++m) v[m] = sw[wslot(base + m)];
#pragma unroll
    for (int u = 0; u < 16; ++u)
#pragma unroll
      for (int t = 0; t < 16; ++t) acc[t] = fma(x[u], v[t - u + 15], acc[t]);
  }
  double* o = w.part + w.sp * cand + ((size_t)cc * s.NTc + q) * 16;
#pragma unroll
  for (int t = 0; t < 16; ++t) o[t] = acc[t];
}

// ---------------------------------------------------------- conv reduce --
// b_k = sum of the chunk partials of tile k/16 in chunk order; each
// workgroup's maximum goes to mpart (the weights take the max of those).
__global__ __launch_bounds__(RB) void k_polish_conv_reduce(Shape s, Ws w) {
  constexpr int NW = RB / kWave;
  __shared__ double slots[NW];
  const int cand = blockIdx.y;
  if (w.stopped[cand]) return;
  const int k = blockIdx.x * RB + threadIdx.x;
  double b = 0.0;
  if (k < 16 * s.NTc) {
    const int q = k >> 4, t = k & 15;
    int J0, J1;
    conv_tile_range(s.n, q, J0, J1);
    b = chunk_sum(w.part + w.sp * cand + (size_t)q * 16 + t, (long long)s.NTc * 16, J0 / CH,
                  J1 / CH);
    if (k >= 2 * s.n - 1) b = 0.0;
  }
  if (k < s.LB) w.bc[w.sb * cand + k] = b;
  const double m = block_max<double, NW>(b, slots);
  if (threadIdx.x == 0) w.mpart[w.sr * cand + blockIdx.x] = m;
}

// -------------------------------------------------------------- weights --
// c_k = exp((p-1) log1p((b_k - M) / M)) in place of b_k; R partial sums of
// c (1 + d) per workgroup (added in workgroup order by the update).
__global__ __launch_bounds__(RB) void k_polish_weights(Shape s, Ws w, Sched S) {
  constexpr int NW = RB / kWave;
  __shared__ double slots[NW];
  __shared__ double mslots[NW];
  const int cand = blockIdx.y;
  if (w.stopped[cand]) return;
  // M = the max of the reduce workgroups' maxima (every workgroup forms it)
  double mloc = 0.0;
  for (int b = threadIdx.x; b < w.sr; b += RB) {
    const double v = w.mpart[w.sr * cand + b];
    mloc = v > mloc ? v : mloc;
  }
  const double M = block_max<double, NW>(mloc, mslots);
  if (blockIdx.x == 0 && threadIdx.x == 0)  // for the step kernel
    w.mbits[cand] = (unsigned long long)__double_as_longlong(M);
  if (!(M > 0.0)) {
    if (blockIdx.x == 0 && threadIdx.x == 0) w.stopped[cand] = 1;  // the fused kernel's break
    return;
  }
  double p, lr;
  sched_at(S, w.step[cand], p, lr);
  const double pm1 = p - 1.0;
  const int k = blockIdx.x * RB + threadIdx.x;
  double r = 0.0;
  if (k < s.LB) {
    double* bc = w.bc + w.sb * cand;
    double c = 0.0;
    if (k < 2 * s.n - 1) {
      const double d = (bc[k] - M) / M;
      c = lp_weight_fast(d, pm1);
      r = c * (1.0 + d);
    }
    bc[k] = c;
  }
  const double R = block_sum<double, NW>(r, slots);
  if (threadIdx.x == 0) w.rpart[w.sr * cand + blockIdx.x] = R;
}

// --------------------------------------------------------- corr units ----
// gpart[cand][cc][q][t] = sum over j in chunk cc of c_{16q+t+j} a_j.
__global__ __launch_bounds__(PU) void k_polish_corr_units(Shape s, Ws w) {
  __shared__ double sx[CH * 16 + CH];
  __shared__ double sw[(16 * (PU + CH) + 32) * 17 / 16 + 16];
  const int cand = blockIdx.z, cc = blockIdx.y;
  if (w.stopped[cand]) return;
  const int qa = PU * blockIdx.x;
  if (qa >= s.NTg) return;
  const int qb = qa + PU - 1 < s.NTg - 1 ? qa + PU - 1 : s.NTg - 1;
  const double* ap = w.apad + w.sa * cand + PADA;
  const double* cp = w.bc + w.sb * cand;
  const int J0c = CH * cc;
  for (int e = threadIdx.x; e < CH * 16; e += PU) sx[wslot(e)] = ap[16 * J0c + e];
  // window: c_x for x in [16 (qa + J0c), 16 (qb + J0c + CH - 1) + 30]
  const int wlo = 16 * (qa + J0c);
  const int wlen = 16 * (qb - qa + CH) + 15;
  for (int e = threadIdx.x; e < wlen; e += PU) sw[wslot(e)] = cp[wlo + e];
  __syncthreads();
  const int q = qa + (int)threadIdx.x;
  if (q > qb) return;
  double acc[16];
#pragma unroll
  for (int t = 0; t < 16; ++t) acc[t] = 0.0;
  for (int jj = 0; jj < CH; ++jj) {
    double x[16], v[31];
#pragma unroll
    for (int u = 0; u < 16; ++u) x[u] = sx[wslot(16 * jj + u)];
    const int base = 16 * (q + J0c + jj) - wlo;  // window offset of c_{16(q+J)}
#pragma unroll
    for (int m = 0; m < 31; ++m) v[m] = sw[wslot(base + m)];
#pragma unroll
    for (int u = 0; u < 16; ++u)
#pragma unroll
      for (int t = 0; t < 16; ++t) acc[t] = fma(x[u], v[t + u], acc[t]);
  }
  double* o = w.gpart + w.sg * cand + ((size_t)cc * s.NTg + q) * 16;
#pragma unroll
  for (int t = 0; t < 16; ++t) o[t] = acc[t];
}

// --------------------------------------------------------------- update --
// Many workgroups per candidate.  First: R (rpart in order), G_i (chunks in
// order), g = G/R - M, the Adam step (lr/n) mhat / (sqrt(vhat) + 1e-30) and
// clamp >= 0 (mode 0) or the mirror step a exp(clamp(-lr g / M, +-4)) (mode
// 1) into `out`, partial sums of the new heights per workgroup.  The bias
// corrections use beta^t kept in the workspace, multiplied once a step as the
// fused kernel does.
__global__ __launch_bounds__(RB) void k_polish_step(Shape s, Ws w, Sched S, int mode,
                                                    double* __restrict__ pop) {
  constexpr int NW = RB / kWave;
  __shared__ double slots[NW];
  const int cand = blockIdx.y;
  if (w.stopped[cand]) return;
  const int n = s.n;
  const double M = __longlong_as_double((long long)w.mbits[cand]);
  double R = 0.0;
  for (int b = 0; b < w.sr; ++b) R += w.rpart[w.sr * cand + b];
  double p, lr;
  sched_at(S, w.step[cand], p, lr);
  const double b1t = w.bias[2 * cand] * S.beta1, b2t = w.bias[2 * cand + 1] * S.beta2;
  const double bc1 = 1.0 / (1.0 - b1t), bc2 = 1.0 / (1.0 - b2t);
  const int i = blockIdx.x * RB + threadIdx.x;
  double xx = 0.0;
  if (i < n) {
    const double G = chunk_sum(w.gpart + w.sg * cand + (size_t)(i >> 4) * 16 + (i & 15),
                               (long long)s.NTg * 16, 0, s.NCg - 1);
    const double gi = G / R - M;
    const double ai = w.apad[w.sa * cand + PADA + i];
    if (mode == 0) {
      double* mo = w.mom + w.sn * cand;
      double* ve = w.vel + w.sn * cand;
      const double mm = S.beta1 * mo[i] + (1.0 - S.beta1) * gi;
      const double vv = S.beta2 * ve[i] + (1.0 - S.beta2) * gi * gi;
      mo[i] = mm;
      ve[i] = vv;
      const double step = (lr / (double)n) * (mm * bc1) / (sqrt(vv * bc2) + 1e-30);
      xx = ai - step;
      if (xx < 0.0) xx = 0.0;
    } else {
      double e = -lr * (gi / M);
      if (e > 4.0) e = 4.0;
      if (e < -4.0) e = -4.0;
      xx = ai * exp(e);
    }
    pop[(size_t)cand * n + i] = xx;
  }
  const double part = block_sum<double, NW>(xx, slots);
  if (threadIdx.x == 0) w.spart[w.ss * cand + blockIdx.x] = part;
}

// Second: S (spart in order), renormalise sum a = 1 (uniform 1/n if S is not
// positive) into pop and the padded copy; the last workgroup's thread 0
// resets M and advances the step and the bias products.
__global__ __launch_bounds__(RB) void k_polish_renorm(Shape s, Ws w, Sched S,
                                                      double* __restrict__ pop) {
  const int cand = blockIdx.y;
  if (w.stopped[cand]) return;
  const int n = s.n;
  double Ssum = 0.0;
  for (int b = 0; b < w.ss; ++b) Ssum += w.spart[w.ss * cand + b];
  const bool reset = !(Ssum > 0.0);
  const double inv = reset ? 1.0 : 1.0 / Ssum;
  const int i = blockIdx.x * RB + threadIdx.x;
  if (i < n) {
    double* out = pop + (size_t)cand * n;
    const double x = reset ? 1.0 / (double)n : out[i] * inv;
    out[i] = x;
    w.apad[w.sa * cand + PADA + i] = x;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    // every workgroup of this candidate has read M and the step in the
    // previous kernel; this one reads neither
    w.mbits[cand] = 0ull;
    w.step[cand] += 1;
    w.bias[2 * cand] *= S.beta1;
    w.bias[2 * cand + 1] *= S.beta2;
  }
}

__global__ void k_polish_moments(const Ws w, int n, int K, double* mom, double* vel) {
  const long long tot = (long long)n * K;
  for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < tot;
       e += (long long)gridDim.x * blockDim.x) {
    if (mom) mom[e] = w.mom[e];
    if (vel) vel[e] = w.vel[e];
  }
}

}  // namespace
}  // namespace sidon

using namespace sidon;

extern "C" size_t sidon_polish_f64_workspace(int n, int K) {
  if (n < 1 || K < 1) return 0;
  return ws_bytes(n, K);
}

extern "C" int sidon_polish_f64(double* pop, double* mom, double* vel, int n, int K, int iters,
                                double pStart, double pEnd, double lrStart, double lrEnd,
                                double beta1, double beta2, int mode, void* workspace,
                                size_t workspace_bytes, hipStream_t stream) {
  if (!pop || n < 1 || K < 0 || iters < 0 || !workspace) return SIDON_ERR_ARG;
  if (!(pStart >= 1.0) || !(pEnd >= 1.0) || !(lrStart > 0.0) || !(lrEnd > 0.0) ||
      (mode != 0 && mode != 1))
    return SIDON_ERR_ARG;
  if (K == 0 || iters == 0) return SIDON_OK;
  if (workspace_bytes < ws_bytes(n, K)) {
    set_last_error_msg("polish: workspace smaller than sidon_polish_f64_workspace(n, K)");
    return SIDON_ERR_ARG;
  }
  const Shape s = shape_of(n);
  const Ws w = ws_of(workspace, n, K);
  Sched S;
  S.iters = iters;
  S.pStart = pStart;
  S.lnPr = log(pEnd / pStart);
  S.lrStart = lrStart;
  S.lnLr = log(lrEnd / lrStart);
  S.beta1 = beta1;
  S.beta2 = beta2;
  const int setup_x = (s.LA + 255) / 256 < 64 ? (s.LA + 255) / 256 : 64;
  hipLaunchKernelGGL(k_polish_setup, dim3(setup_x, K), dim3(256), 0, stream, pop, n, w);
  SIDON_HIP_TRY(hipGetLastError());
  // conv units: tiles meeting a chunk number at most NJ + 35
  const dim3 gconv((s.NJ + 35 + PU - 1) / PU, s.NCc, K);
  const dim3 gcorr((s.NTg + PU - 1) / PU, s.NCg, K);
  const dim3 gred(rblocks(s), K);
  const dim3 gupd(sblocks(s), K);
  for (int it = 0; it < iters; ++it) {
    hipLaunchKernelGGL(k_polish_conv_units, gconv, dim3(PU), 0, stream, s, w);
    hipLaunchKernelGGL(k_polish_conv_reduce, gred, dim3(RB), 0, stream, s, w);
    hipLaunchKernelGGL(k_polish_weights, gred, dim3(RB), 0, stream, s, w, S);
    hipLaunchKernelGGL(k_polish_corr_units, gcorr, dim3(PU), 0, stream, s, w);
    hipLaunchKernelGGL(k_polish_step, gupd, dim3(RB), 0, stream, s, w, S, mode, pop);
    hipLaunchKernelGGL(k_polish_renorm, gupd, dim3(RB), 0, stream, s, w, S, pop);
    SIDON_HIP_TRY(hipGetLastError());
  }
  if (mom || vel) {
    hipLaunchKernelGGL(k_polish_moments, dim3(64), dim3(256), 0, stream, w, n, K, mom, vel);
    SIDON_HIP_TRY(hipGetLastError());
  }
  return SIDON_OK;
}
