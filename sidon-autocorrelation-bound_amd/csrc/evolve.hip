// evolve.hip -- cell split, Gaussian multiplicative mutation, island
// selection and the small device-side helpers that let a whole generation run
// without the host (and so be captured as one hipGraph).
//
// Reference: src/kernels.cuh:534-547 (kRefine), :582-801 (kVariation, its
// fourteen operators and the clone branch), :803-849 (kSupportSig),
// src/main.cu:650-735 (host selection / migration).
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
constexpr uint32_t kStreamParams = 0x50524D53u;

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
// The 64 operator parameters of offspring `slot`, in (0, 1] (the role of the
// reference's sPar): Philox4x32-10 counters {q, slot, gen, gen>>32 ^ tag},
// q = 0..15, four words each, u = ((x >> 8) + 1) 2^-24, exact in fp32.
__device__ __forceinline__ void op_params(uint64_t seed, uint64_t gen, int slot, int q,
                                          float* prm) {
  uint32_t c[4] = {(uint32_t)q, (uint32_t)slot, (uint32_t)gen,
                   (uint32_t)(gen >> 32) ^ kStreamParams};
  philox4x32_10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
#pragma unroll
  for (int w = 0; w < 4; ++w) prm[4 * q + w] = (float)((c[w] >> 8) + 1u) * (1.0f / 16777216.0f);
}

__device__ __forceinline__ bool op_uses_mean(int op) {
  return op == 4 || op == 5 || op == 10 || op == 12 || op == 13;
}

// One height of operator `op` (kernels.cuh:614-787) in fp64, rounded once to
// fp32 by the caller: A the first parent, B the second (crossover ops 0, 1,
// 8), u the parameters, mean = sum(A) / n, sg the strength.  No contraction:
// every product and sum rounds on its own, so the host oracle reproduces the
// value operation for operation (library cos / exp / log aside).
#pragma clang fp contract(off)
__device__ double vary(int op, int i, int n, const float* A, const float* B, const float* prm,
                       double mean, double sg, uint64_t seed, uint64_t gen, int slot) {
  const double a = A[i];
  const double twopi = 6.283185307179586476925286766559;
  auto u = [prm](int k) { return (double)prm[k]; };
  switch (op) {
    case 0: {  // blend crossover, lambda in [-0.3, 1.3]
      const double lam = -0.3 + 1.6 * u(0);
      const double x = lam * a + (1.0 - lam) * (double)B[i];
      return x > 0.0 ? x : 0.0;
    }
    case 1: {  // two-point segment crossover
      int c1 = (int)(u(1) * n), c2 = (int)(u(2) * n);
      if (c1 > c2) { const int t = c1; c1 = c2; c2 = t; }
      return (i >= c1 && i < c2) ? (double)B[i] : a;
    }
    case 2: {  // log-normal multiplicative noise
      const double x = a * exp(sg * normal_at(seed, gen, slot, i));
      return x > 0.0 ? x : 0.0;
    }
    case 3: {  // smooth low-frequency perturbation
      const int nf = 1 + (int)(u(3) * 5.0);
      double m = 0.0;
      for (int f = 0; f < nf; ++f) {
        const double freq = 1.0 + u(4 + f) * 12.0;
        m += cos(twopi * freq * (double)i / (double)n + twopi * u(12 + f)) / (double)nf;
      }
      const double x = a * (1.0 + sg * 3.0 * m);
      return x > 0.0 ? x : 0.0;
    }
    case 4: {  // additive Gaussian noise scaled by the mean height
      const double x = a + sg * mean * normal_at(seed, gen, slot, i);
      return x > 0.0 ? x : 0.0;
    }
    case 5: {  // sparsify below a fraction of the mean
      const double thr = mean * (0.05 + 0.9 * u(20));
      return a < thr ? 0.0 : a;
    }
    case 6: {  // symmetrise, or hard reflect
      const double r = A[n - 1 - i];
      return u(21) < 0.5 ? r : 0.5 * (a + r);
    }
    case 7: {  // shift by up to n/8 either way, zero fill
      const int sh = (int)((u(22) - 0.5) * (double)n * 0.25);
      const int k = i - sh;
      return (k >= 0 && k < n) ? (double)A[k] : 0.0;
    }
    case 8: {  // splice a random block of B into A
      const int len = 1 + (int)(u(23) * 0.5 * (double)n);
      const int st0 = (int)(u(24) * (double)(n - len > 0 ? n - len : 1));
      return (i >= st0 && i < st0 + len) ? (double)B[i] : a;
    }
    case 9: {  // random restart: smooth positive random function
      const int nf = 2 + (int)(u(25) * 6.0);
      double m = 0.0;
      for (int f = 0; f < nf; ++f) {
        const double freq = 1.0 + u(4 + (f % 8)) * 10.0;
        m += cos(twopi * freq * (double)i / (double)n + twopi * u(12 + (f % 8)));
      }
      const double v = 0.3 + 0.7 * u(26) + m / (double)nf + 0.4 * normal_at(seed, gen, slot, i);
      return v > 0.0 ? v : 0.0;
    }
    case 10: {  // spike of width 2w+1 at a random location
      const int loc = (int)(u(27) * (double)n);
      const int wdt = 1 + (int)(u(28) * (double)n * 0.03);
      const double amp = mean * (1.0 + 6.0 * u(29));
      return (i >= loc - wdt && i <= loc + wdt) ? a + amp : a;
    }
    case 11: {  // structured seed: plateau, depression, hump, band, gap, spike, atom
      const bool flip = u(30) < 0.5;
      const double atom = 30.0 + 250.0 * u(31);
      const double pl1 = 0.28 + 0.16 * u(32);
      const double pl2 = pl1 + 0.10 + 0.14 * u(33);
      const double hump = pl2 + 0.02 + 0.08 * u(34);
      const double gap0 = 0.70 + 0.16 * u(35);
      const double gap1 = gap0 + 0.05 + 0.15 * u(36);
      const double tail = 3.0 + 14.0 * u(37);
      const int nf = 3 + (int)(u(38) * 6.0);
      const int d = flip ? n - 1 - i : i;
      const double x = ((double)d + 0.5) / (double)n;
      double rough = 0.0;
      for (int f = 0; f < nf; ++f)
        rough += cos(twopi * (2.0 + u(40 + f) * 40.0) * x + twopi * u(48 + f));
      rough /= (double)nf;
      double v;
      if (x < pl1) v = 4.2 * (1.0 + 0.45 * rough);
      else if (x < pl2) v = 0.7 * (1.0 + 0.80 * rough);
      else if (x < hump) v = 5.2 * (1.0 + 0.30 * rough);
      else if (x < gap0) v = 2.0 * (1.0 + 0.70 * rough);
      else if (x < gap1) v = 0.0;
      else v = tail * (1.0 + 0.20 * rough);
      if (d == 0) v += atom;
      else if (d < 4) v += atom * 0.10 / (double)d;
      return v > 0.0 ? v : 0.0;
    }
    case 12: {  // carve out (or fill in) a hard support gap
      const int len = 1 + (int)((0.015 + 0.13 * u(30)) * (double)n);
      const int st0 = (int)(u(31) * (double)(n > len ? n - len : 1));
      if (i < st0 || i >= st0 + len) return a;
      return u(32) < 0.7 ? 0.0 : a + mean * (0.2 + 1.8 * u(33));
    }
    case 13: {  // grow or shrink the endpoint atom
      const int w = 1 + (int)(u(31) * 0.012 * (double)n);
      const int d = u(30) < 0.5 ? i : n - 1 - i;
      if (d >= w) return a;
      const double x = a + mean * (-20.0 + 260.0 * u(32)) / (double)(1 + d);
      return x > 0.0 ? x : 0.0;
    }
    default:  // clone
      return a;
  }
}

// Deterministic double-double block sum of per-thread (hi, lo): lanes by
// butterfly (symmetric in self and partner, so every lane holds the same
// pair), waves in order.  s_hi / s_lo are free again on return.
template <int NW>
__device__ double block_dd_sum(double hi, double lo, double* s_hi, double* s_lo) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    const double oh = __shfl_xor(hi, off, kWave), ol = __shfl_xor(lo, off, kWave);
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
  __syncthreads();
  return H + L;
}
#pragma clang fp contract(on)

// One block per offspring: O = operator op[j] on parents p1[j] (and p2[j]),
// each height in fp64 rounded once to fp32, then O /= S with S the
// double-double (practically correctly rounded) sum of the fp32 heights.  In
// the uneven-length form the offspring keeps p1's length, and a second parent
// of another length is replaced by the first.
template <int BD>
__global__ __launch_bounds__(BD) void k_variation(const float* __restrict__ pop,
                                                  const int* __restrict__ ns, float* __restrict__ out,
                                                  int* __restrict__ ns_out, long long stride, int n,
                                                  int nOff, const int* __restrict__ p1,
                                                  const int* __restrict__ p2,
                                                  const int* __restrict__ op,
                                                  const float* __restrict__ strength,
                                                  uint64_t seed, uint64_t gen_host,
                                                  const unsigned long long* gen_dev) {
  constexpr int NW = BD / kWave;
  __shared__ double s_hi[NW], s_lo[NW];
  __shared__ float prm[64];
  const int j = blockIdx.x;
  if (j >= nOff) return;
  const uint64_t gen = gen_dev ? *gen_dev : gen_host;
  const int par = p1[j];
  const int len = ns ? ns[par] : n;
  const float* A = pop + (size_t)par * stride;
  const int pb = p2 ? p2[j] : par;
  const float* B = (ns && ns[pb] != len) ? A : pop + (size_t)pb * stride;
  float* O = out + (size_t)j * stride;
  const int o = op[j];
  const double sg = (double)strength[j];
  if (o >= 0 && o <= 13) {
    if (threadIdx.x < 16) op_params(seed, gen, j, threadIdx.x, prm);
    __syncthreads();
  }
  double mean = 0.0;
  if (op_uses_mean(o)) {
    double hi = 0.0, lo = 0.0;
    for (int i = threadIdx.x; i < len; i += BD) {
      double s, e;
      two_sum(hi, (double)A[i], s, e);
      hi = s;
      lo += e;
    }
    mean = block_dd_sum<NW>(hi, lo, s_hi, s_lo) / (double)len;
  }
  double hi = 0.0, lo = 0.0;
  for (int i = threadIdx.x; i < len; i += BD) {
    const float x = (float)vary(o, i, len, A, B, prm, mean, sg, seed, gen, j);
    O[i] = x;
    double s, e;
    two_sum(hi, (double)x, s, e);
    hi = s;
    lo += e;
  }
  const double S = block_dd_sum<NW>(hi, lo, s_hi, s_lo);
  if (S > 0.0) {
    const double inv = 1.0 / S;
    for (int i = threadIdx.x; i < len; i += BD) O[i] = (float)((double)O[i] * inv);
  } else {
    for (int i = threadIdx.x; i < len; i += BD) O[i] = 1.0f / (float)len;
  }
  if (ns_out && threadIdx.x == 0) ns_out[j] = len;
}

// ------------------------------------------------- support signature -----
// kSupportSig (kernels.cuh:803-849) with its block of 256 threads and 32-lane
// warps (kernels.cuh:30, :80-90): thread t sums a[t], a[t+256], ... in fp32;
// each 32-lane half of a wave reduces by the shuffle tree v += v[lane + 16],
// +8, +4, +2, +1 (the reference's warp), and thread 0 adds the 8 warp sums in
// order, so the fp32 mean, the threshold mean * tol and every a_i > thr
// decision are the reference's.  The signature is the XOR of mix(i) over the
// support (the per-thread FNV start values cancel in pairs), nnz its size.
__device__ __forceinline__ unsigned long long sig_mix(int i) {
  unsigned long long x = (unsigned long long)i * 0x9E3779B97F4A7C15ull;
  x ^= x >> 29;
  x *= 0xBF58476D1CE4E5B9ull;
  x ^= x >> 32;
  return x;
}

__global__ __launch_bounds__(256) void k_support_sig(const float* __restrict__ pop, int n,
                                                     float tol, unsigned long long* __restrict__ sig,
                                                     int* __restrict__ nnz) {
  __shared__ float s_warp[8];
  __shared__ float s_thr;
  __shared__ unsigned long long s_h[4];
  __shared__ int s_c[4];
  const float* a = pop + (size_t)blockIdx.x * n;
  const int t = threadIdx.x;
  float v = 0.0f;
  for (int i = t; i < n; i += 256) v += a[i];
#pragma unroll
  for (int off = 16; off > 0; off >>= 1) v += __shfl_down(v, off, 32);
  if ((t & 31) == 0) s_warp[t >> 5] = v;
  __syncthreads();
  if (t == 0) {
    float tot = 0.0f;
    for (int w = 0; w < 8; ++w) tot += s_warp[w];
    s_thr = (tot / (float)n) * tol;
  }
  __syncthreads();
  const float thr = s_thr;
  unsigned long long h = 0;
  int c = 0;
  for (int i = t; i < n; i += 256)
    if (a[i] > thr) {
      ++c;
      h ^= sig_mix(i);
    }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    h ^= __shfl_xor(h, off, kWave);
    c += __shfl_xor(c, off, kWave);
  }
  if ((t & (kWave - 1)) == 0) { s_h[t / kWave] = h; s_c[t / kWave] = c; }
  __syncthreads();
  if (t == 0) {
    sig[blockIdx.x] = s_h[0] ^ s_h[1] ^ s_h[2] ^ s_h[3];
    nnz[blockIdx.x] = s_c[0] + s_c[1] + s_c[2] + s_c[3];
  }
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
  if (!pop || !out || !p1 || !op || !strength || n < 1 || nOff < 0) return SIDON_ERR_ARG;
  if (nOff == 0) return SIDON_OK;
  hipLaunchKernelGGL(k_variation<256>, dim3(nOff), dim3(256), 0, stream, pop, (const int*)nullptr,
                     out, (int*)nullptr, (long long)n, n, nOff, p1, p2, op, strength,
                     (uint64_t)seed, (uint64_t)gen, (const unsigned long long*)nullptr);
  SIDON_HIP_TRY(hipGetLastError());
  return SIDON_OK;
}

extern "C" int sidon_variation_ragged(const float* pop, const int* ns, float* out, int* ns_out,
                                      int stride, int nOff, const int* p1, const int* p2,
                                      const int* op, const float* strength,
                                      unsigned long long seed,
                                      const unsigned long long* gen_dev, hipStream_t stream) {
  if (!pop || !ns || !out || !p1 || !op || !strength || !gen_dev || stride < 1 || nOff < 0)
    return SIDON_ERR_ARG;
  if (nOff == 0) return SIDON_OK;
  hipLaunchKernelGGL(k_variation<256>, dim3(nOff), dim3(256), 0, stream, pop, ns, out, ns_out,
                     (long long)stride, stride, nOff, p1, p2, op, strength, (uint64_t)seed,
                     (uint64_t)0, gen_dev);
  SIDON_HIP_TRY(hipGetLastError());
  return SIDON_OK;
}

extern "C" int sidon_support_sig(const float* pop, int n, int P, float tol,
                                 unsigned long long* sig, int* nnz, hipStream_t stream) {
  if (!pop || !sig || !nnz || n < 1 || P < 0) return SIDON_ERR_ARG;
  if (P == 0) return SIDON_OK;
  hipLaunchKernelGGL(k_support_sig, dim3(P), dim3(256), 0, stream, pop, n, tol, sig, nnz);
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
