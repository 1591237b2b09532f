// evolve.hip -- cell split, Gaussian multiplicative mutation, island
// selection and the small device-side helpers that let a whole generation run
// without the host (and so be captured as one hipGraph).
//
// Reference: src/kernels.cuh:534-547 (kRefine), :582-801 (kVariation, op 2 and
// the clone branch), src/main.cu:650-735 (host selection / migration).
// The reference draws mutation noise from curand Philox4_32_10 and selection
// from std::mt19937_64 on the host; here both come from the build's own
// Philox4x32-10 (Salmon et al., SC'11) keyed by (seed, generation, slot), so a
// generation is a pure function of device state.
#include <cfloat>
#include <cmath>

#include "common.hpp"
#include "runtime.hpp"

namespace sidon {
namespace {

constexpr uint32_t kStreamVariation = 0x56415249u;
constexpr uint32_t kStreamSelect = 0x53454C45u;

__device__ __forceinline__ void philox4x32_10(uint32_t c[4], uint32_t k0, uint32_t k1) {
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    if (r) { k0 += 0x9E3779B9u; k1 += 0xBB67AE85u; }
    const uint64_t p0 = (uint64_t)0xD2511F53u * c[0];
    const uint64_t p1 = (uint64_t)0xCD9E8D57u * c[2];
    const uint32_t o0 = (uint32_t)(p1 >> 32) ^ c[1] ^ k0;
    const uint32_t o2 = (uint32_t)(p0 >> 32) ^ c[3] ^ k1;
    c[0] = o0; c[1] = (uint32_t)p1; c[2] = o2; c[3] = (uint32_t)p0;
  }
}

// Standard normal for element i of offspring `slot` (Box-Muller in fp64 on
// u = (2*(x>>9)+1) 2^-24, exact in (0,1)).
__device__ __forceinline__ double normal_at(uint64_t seed, uint64_t gen, int slot, int i) {
  uint32_t c[4] = {(uint32_t)(i >> 2), (uint32_t)slot, (uint32_t)gen,
                   (uint32_t)(gen >> 32) ^ kStreamVariation};
  philox4x32_10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
  const int pair = (i & 3) >> 1;
  const uint32_t x0 = c[2 * pair], x1 = c[2 * pair + 1];
  const double u1 = (double)(2u * (x0 >> 9) + 1u) * (1.0 / 16777216.0);
  const double u2 = (double)(2u * (x1 >> 9) + 1u) * (1.0 / 16777216.0);
  const double r = sqrt(-2.0 * log(u1));
  const double th = 6.283185307179586476925286766559 * u2;
  return (i & 1) ? r * sin(th) : r * cos(th);
}

#pragma clang fp contract(off)
__device__ __forceinline__ void two_sum(double a, double b, double& s, double& e) {
  s = a + b;
  const double bb = s - a;
  e = (a - (s - bb)) + (b - bb);
}
#pragma clang fp contract(on)

// ------------------------------------------------------------ refine -----
__global__ void k_refine(const float* __restrict__ src, int n, float* __restrict__ dst) {
  const float* s = src + (size_t)blockIdx.y * n;
  float2* d = reinterpret_cast<float2*>(dst + (size_t)blockIdx.y * 2 * n);
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const float x = s[i];
    d[i] = make_float2(x, x);
  }
}

// --------------------------------------------------------- variation -----
// One block per offspring.  op 2: O = A exp(sigma z) (product in fp64,
// rounded once), clamp >= 0; other ops clone.  Then O /= S with S the
// double-double (practically correctly rounded) sum of the fp32 heights.
template <int BD>
__global__ __launch_bounds__(BD) void k_variation(const float* __restrict__ pop,
                                                  const int* __restrict__ ns, float* __restrict__ out,
                                                  int* __restrict__ ns_out, long long stride, int n,
                                                  int nOff, const int* __restrict__ p1,
                                                  const int* __restrict__ op,
                                                  const float* __restrict__ strength,
                                                  uint64_t seed, uint64_t gen_host,
                                                  const unsigned long long* gen_dev) {
  constexpr int NW = BD / kWave;
  __shared__ double s_hi[NW], s_lo[NW];
  const int j = blockIdx.x;
  if (j >= nOff) return;
  const uint64_t gen = gen_dev ? *gen_dev : gen_host;
  const int par = p1[j];
  const int len = ns ? ns[par] : n;
  const float* A = pop + (size_t)par * stride;
  float* O = out + (size_t)j * stride;
  const int o = op[j];
  const double sg = (double)strength[j];
  double hi = 0.0, lo = 0.0;
  for (int i = threadIdx.x; i < len; i += BD) {
    float x;
    if (o == 2) {
      const double z = normal_at(seed, gen, j, i);
      x = (float)((double)A[i] * exp(sg * z));
      x = x > 0.0f ? x : 0.0f;
    } else {
      x = A[i];
    }
    O[i] = x;
    double s, e;
    two_sum(hi, (double)x, s, e);
    hi = s;
    lo += e;
  }
  // deterministic double-double reduction: lanes by butterfly, waves in order
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    const double oh = __shfl_xor(hi, off, kWave), ol = __shfl_xor(lo, off, kWave);
    // combine symmetric in (self, other) so every lane gets the same pair
    const double a = hi < oh ? hi : oh, b = hi < oh ? oh : hi;
    double s, e;
    two_sum(b, a, s, e);
    hi = s;
    lo = (lo + ol) + e;
  }
  const int w = threadIdx.x / kWave;
  if ((threadIdx.x & (kWave - 1)) == 0) { s_hi[w] = hi; s_lo[w] = lo; }
  __syncthreads();
  double H = s_hi[0], L = s_lo[0];
  for (int q = 1; q < NW; ++q) {
    double s, e;
    two_sum(H, s_hi[q], s, e);
    H = s;
    L = (L + s_lo[q]) + e;
  }
  const double S = H + L;
  if (S > 0.0) {
    const double inv = 1.0 / S;
    for (int i = threadIdx.x; i < len; i += BD) O[i] = (float)((double)O[i] * inv);
  } else {
    for (int i = threadIdx.x; i < len; i += BD) O[i] = 1.0f / (float)len;
  }
  if (ns_out && threadIdx.x == 0) ns_out[j] = len;
}

// --------------------------------------------------------- selection -----
// One block; thread-strided over slots.  fit/ord are staged in LDS.
constexpr int kSelMax = 4096;

__global__ __launch_bounds__(1024) void k_select(double* fit, float* pop, int* ns, long long stride,
                                                 int P, int NI, int elite, int migrate_every,
                                                 float sigma, uint64_t seed,
                                                 const unsigned long long* gen_dev, int* p1,
                                                 int* p2, int* op, float* strength) {
  __shared__ double sf[kSelMax];
  __shared__ int ord[kSelMax];
  const int PI = P / NI;
  const uint64_t gen = *gen_dev;
  for (int x = threadIdx.x; x < P; x += blockDim.x) sf[x] = fit[x];
  __syncthreads();
  for (int x = threadIdx.x; x < P; x += blockDim.x) {
    const int isl = x / PI;
    int rk = 0;
    for (int q = 0; q < PI; ++q) {
      const int y = isl * PI + q;
      rk += (sf[y] < sf[x] || (sf[y] == sf[x] && y < x)) ? 1 : 0;
    }
    ord[isl * PI + rk] = x;
  }
  __syncthreads();
  if (migrate_every > 0 && gen > 0 && gen % (uint64_t)migrate_every == 0 && NI > 1) {
    for (int i = 0; i < NI; ++i) {  // sequential ring, src/main.cu:685-691
      const int src = ord[i * PI];
      const int dst = ord[((i + 1) % NI) * PI + PI - 1];
      const int len = ns ? ns[src] : (int)stride;
      const float* s = pop + (size_t)src * stride;
      float* d = pop + (size_t)dst * stride;
      for (int k = threadIdx.x; k < len; k += blockDim.x) d[k] = s[k];
      __syncthreads();
      if (threadIdx.x == 0) {
        sf[dst] = sf[src];
        fit[dst] = sf[src];
        if (ns) ns[dst] = len;
      }
      __syncthreads();
    }
  }
  const uint32_t k0 = (uint32_t)seed, k1 = (uint32_t)(seed >> 32);
  for (int slot = threadIdx.x; slot < P; slot += blockDim.x) {
    const int i = slot / PI, r = slot % PI;
    if (r < elite) {
      p1[slot] = ord[i * PI + r];
      p2[slot] = ord[i * PI + r];
      op[slot] = 20;
      strength[slot] = 0.0f;
      continue;
    }
    uint32_t c[4] = {(uint32_t)slot, 0u, (uint32_t)gen, (uint32_t)(gen >> 32) ^ kStreamSelect};
    philox4x32_10(c, k0, k1);
    int pick[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) pick[q] = ord[i * PI + (int)(((uint64_t)c[q] * (uint64_t)PI) >> 32)];
    p1[slot] = sf[pick[0]] < sf[pick[1]] ? pick[0] : pick[1];
    p2[slot] = sf[pick[2]] < sf[pick[3]] ? pick[2] : pick[3];
    op[slot] = 2;
    strength[slot] = sigma;
  }
}

__global__ void k_advance(unsigned long long* g, unsigned long long delta) { *g += delta; }

__global__ __launch_bounds__(1024) void k_order(const int* ns, int P, int* order) {
  for (int p = threadIdx.x; p < P; p += blockDim.x) {
    const int np = ns[p];
    int rk = 0;
    for (int q = 0; q < P; ++q) {
      const int nq = ns[q];
      rk += (nq > np || (nq == np && q < p)) ? 1 : 0;
    }
    order[rk] = p;
  }
}

}  // namespace
}  // namespace sidon

using namespace sidon;

extern "C" int sidon_refine(const float* src, int n, int P, float* dst, hipStream_t stream) {
  if (!src || !dst || n < 1 || P < 0) return SIDON_ERR_ARG;
  if (P == 0) return SIDON_OK;
  const int bx = (n + 255) / 256 < 64 ? (n + 255) / 256 : 64;
  hipLaunchKernelGGL(k_refine, dim3(bx, P), dim3(256), 0, stream, src, n, dst);
  SIDON_HIP_TRY(hipGetLastError());
  return SIDON_OK;
}

extern "C" int sidon_variation(const float* pop, float* out, int n, int nOff, const int* p1,
                               const int* p2, const int* op, const float* strength,
                               unsigned long long seed, unsigned long long gen,
                               hipStream_t stream) {
  (void)p2;  // second parent: used only by the crossover operators (not built)
  if (!pop || !out || !p1 || !op || !strength || n < 1 || nOff < 0) return SIDON_ERR_ARG;
  if (nOff == 0) return SIDON_OK;
  hipLaunchKernelGGL(k_variation<256>, dim3(nOff), dim3(256), 0, stream, pop, (const int*)nullptr,
                     out, (int*)nullptr, (long long)n, n, nOff, p1, op, strength, (uint64_t)seed,
                     (uint64_t)gen, (const unsigned long long*)nullptr);
  SIDON_HIP_TRY(hipGetLastError());
  return SIDON_OK;
}

extern "C" int sidon_variation_ragged(const float* pop, const int* ns, float* out, int* ns_out,
                                      int stride, int nOff, const int* p1, const int* p2,
                                      const int* op, const float* strength,
                                      unsigned long long seed,
                                      const unsigned long long* gen_dev, hipStream_t stream) {
  (void)p2;
  if (!pop || !ns || !out || !p1 || !op || !strength || !gen_dev || stride < 1 || nOff < 0)
    return SIDON_ERR_ARG;
  if (nOff == 0) return SIDON_OK;
  hipLaunchKernelGGL(k_variation<256>, dim3(nOff), dim3(256), 0, stream, pop, ns, out, ns_out,
                     (long long)stride, stride, nOff, p1, op, strength, (uint64_t)seed,
                     (uint64_t)0, gen_dev);
  SIDON_HIP_TRY(hipGetLastError());
  return SIDON_OK;
}

extern "C" int sidon_select(double* fit, float* pop, int* ns, int stride, int P, int NI, int elite,
                            int migrate_every, float sigma, unsigned long long seed,
                            const unsigned long long* gen_dev, int* p1, int* p2, int* op,
                            float* strength, hipStream_t stream) {
  if (!fit || !gen_dev || !p1 || !p2 || !op || !strength || P < 1 || NI < 1 || P % NI != 0 ||
      P > kSelMax || elite < 0 || elite > P / NI)
    return SIDON_ERR_ARG;
  if (migrate_every > 0 && (!pop || stride < 1)) return SIDON_ERR_ARG;
  hipLaunchKernelGGL(k_select, dim3(1), dim3(1024), 0, stream, fit, pop, ns, (long long)stride, P,
                     NI, elite, migrate_every, sigma, (uint64_t)seed, gen_dev, p1, p2, op,
                     strength);
  SIDON_HIP_TRY(hipGetLastError());
  return SIDON_OK;
}

extern "C" int sidon_advance_gen(unsigned long long* gen_dev, unsigned long long delta,
                                 hipStream_t stream) {
  if (!gen_dev) return SIDON_ERR_ARG;
  hipLaunchKernelGGL(k_advance, dim3(1), dim3(1), 0, stream, gen_dev, delta);
  SIDON_HIP_TRY(hipGetLastError());
  return SIDON_OK;
}

extern "C" int sidon_order_by_length(const int* ns, int P, int* order, hipStream_t stream) {
  if (!ns || !order || P < 0) return SIDON_ERR_ARG;
  if (P == 0) return SIDON_OK;
  hipLaunchKernelGGL(k_order, dim3(1), dim3(1024), 0, stream, ns, P, order);
  SIDON_HIP_TRY(hipGetLastError());
  return SIDON_OK;
}
