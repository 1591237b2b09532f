// eval.hip -- FP64 objective: F, near-active peak count, and a correctly
// rounded F for the cell-split check.
//
// Reference: src/kernels.cuh:291-380 (kEval<T,Acc>), launch src/main.cu:161-166.
// One workgroup per candidate, a (fp64) in LDS with 32 zero cells on each
// side so no bounds test sits in the inner loop.  Each thread evaluates tiles
// of 16 consecutive outputs k over aligned 16-blocks of i (16 x 16 fp64 FMAs
// per 47 LDS reads); tiles q and NT-1-q are paired so every thread gets the
// same triangle area.  b is never stored: per-tile maxima are kept, and the
// active-peak pass recomputes only tiles whose maximum reaches the threshold.
//
// For fp32 heights every product is exact in fp64 and each b_k accumulates in
// ascending i, and S is summed sequentially, exactly as hostF
// (src/main.cu:94-105) does, so F matches hostF bit for bit.
#include <cfloat>
#include <cmath>

#include "common.hpp"
#include "runtime.hpp"

namespace sidon {
namespace {

constexpr int EPAD = 32;

__host__ __device__ inline int eval_tiles(int n) { return (2 * n - 1 + 15) / 16; }

__host__ __device__ inline long long eval_lds_elems(int n, int block) {
  return (long long)n + 2 * EPAD + eval_tiles(n) + 2 * (block / kWave) + 2;
}

// acc[t] = sum_i a_i a_{k0+t-i} over the 16-blocks of i the tile needs.
__device__ __forceinline__ void eval_tile(const double* sA, int n, int k0, double acc[16]) {
#pragma unroll
  for (int t = 0; t < 16; ++t) acc[t] = 0.0;
  const int ilo = k0 - (n - 1) > 0 ? k0 - (n - 1) : 0;
  const int ihi = k0 + 15 < n - 1 ? k0 + 15 : n - 1;
  for (int J = ilo >> 4; J <= (ihi >> 4); ++J) {
    double x[16], w[31];
#pragma unroll
    for (int s = 0; s < 16; ++s) x[s] = sA[16 * J + s];
    const int wb = k0 - 16 * J - 15;
#pragma unroll
    for (int m = 0; m < 31; ++m) w[m] = sA[wb + m];
#pragma unroll
    for (int s = 0; s < 16; ++s)
#pragma unroll
      for (int t = 0; t < 16; ++t) acc[t] += x[s] * w[t - s + 15];
  }
}

template <typename Tin, int BD>
__global__ __launch_bounds__(BD) void k_eval(const Tin* __restrict__ pop, long long stride,
                                             const int* __restrict__ ns, int n_uniform,
                                             double* F, int* nActive, double tol) {
  constexpr int NW = BD / kWave;
  const int n = ns ? ns[blockIdx.x] : n_uniform;
  extern __shared__ __align__(16) unsigned char smem[];
  double* base = reinterpret_cast<double*>(smem);
  double* sA = base + EPAD;
  const int NT = eval_tiles(n);
  double* tmax = base + n + 2 * EPAD;
  double* slots = tmax + NT;
  double* sS = slots + 2 * NW;
  const int tid = threadIdx.x;
  const Tin* a = pop + (size_t)blockIdx.x * stride;
  for (int i = tid; i < n + 2 * EPAD; i += BD) base[i] = 0.0;
  __syncthreads();
  for (int i = tid; i < n; i += BD) sA[i] = (double)a[i];
  __syncthreads();
  if (tid == 0) {  // sequential S, hostF order
    double S = 0.0;
    for (int i = 0; i < n; ++i) S += sA[i];
    sS[0] = S;
  }
  const int nb = 2 * n - 1;
  double loc = 0.0;
  for (int q = tid; q < (NT + 1) / 2; q += BD) {
    for (int side = 0; side < 2; ++side) {
      const int tq = side ? NT - 1 - q : q;
      if (side && tq == q) break;
      double acc[16];
      eval_tile(sA, n, 16 * tq, acc);
      double mx = 0.0;
#pragma unroll
      for (int t = 0; t < 16; ++t)
        if (16 * tq + t < nb) mx = acc[t] > mx ? acc[t] : mx;
      tmax[tq] = mx;
      loc = mx > loc ? mx : loc;
    }
  }
  const double M = block_max<double, NW>(loc, slots);
  if (nActive) {
    const double thr = M * (1.0 - tol);
    double cnt = 0.0;
    for (int q = tid; q < NT; q += BD) {
      if (!(tmax[q] >= thr)) continue;
      double acc[16];
      eval_tile(sA, n, 16 * q, acc);
#pragma unroll
      for (int t = 0; t < 16; ++t)
        if (16 * q + t < nb && acc[t] >= thr) cnt += 1.0;
    }
    const double c = block_sum<double, NW>(cnt, slots + NW);
    if (tid == 0) nActive[blockIdx.x] = (int)(c + 0.5);
  }
  if (tid == 0) {
    const double S = sS[0];
    F[blockIdx.x] = S > 0.0 ? 2.0 * (double)n * M / (S * S) : 1e30;
  }
}

// ------------------------------------------------------------ exact F ----
#pragma clang fp contract(off)
__device__ __forceinline__ void two_sum(double a, double b, double& s, double& e) {
  s = a + b;
  const double bb = s - a;
  e = (a - (s - bb)) + (b - bb);
}

// Triple-length accumulator: sum = hi + lo + lo2, where hi collects the
// terms, lo the exact TwoSum errors of hi and lo2 the exact TwoSum errors of
// lo; only lo2 is rounded.  `exact` stays true while every error reaching lo2
// is zero, i.e. while hi + lo is the exact sum.
struct Acc3 {
  double hi, lo, lo2;
  bool exact;
};
__device__ __forceinline__ void acc3_add(Acc3& a, double p) {
  double s, e, l, e2;
  two_sum(a.hi, p, s, e);
  a.hi = s;
  two_sum(a.lo, e, l, e2);
  a.lo = l;
  if (e2 != 0.0) { a.lo2 += e2; a.exact = false; }
}

// Rounds the accumulated sum once.  Exact accumulation: fl(hi + lo) is the
// correctly rounded sum (ties to even included), certified.  Otherwise the
// rounding is certified unless the remainder lies within the error bound of
// half an ulp (rel: bound on the error of lo2 relative to the sum).
__device__ __forceinline__ double round_certified(const Acc3& a, double rel, int& amb) {
  if (a.exact) {
    amb = 0;
    return a.hi + a.lo;
  }
  double s1, r1, s, r2;
  two_sum(a.lo, a.lo2, s1, r1);
  two_sum(a.hi, s1, s, r2);
  const double err = r2 + r1;
  const double delta = rel * fabs(s) + 0x1p-52 * fabs(err);
  const double up = nextafter(s, DBL_MAX) - s;
  const double dn = s - nextafter(s, 0.0);
  amb = (err + delta >= 0.5 * up) || (err - delta <= -0.5 * dn);
  return s;
}

// Exact tile: triple-length sums of exact fp32 x fp32 products.
__device__ __forceinline__ void exact_tile(const double* sA, int n, int k0, Acc3 acc[16]) {
#pragma unroll
  for (int t = 0; t < 16; ++t) acc[t] = Acc3{0.0, 0.0, 0.0, true};
  const int ilo = k0 - (n - 1) > 0 ? k0 - (n - 1) : 0;
  const int ihi = k0 + 15 < n - 1 ? k0 + 15 : n - 1;
  for (int J = ilo >> 4; J <= (ihi >> 4); ++J) {
    double x[16], w[31];
#pragma unroll
    for (int s = 0; s < 16; ++s) x[s] = sA[16 * J + s];
    const int wb = k0 - 16 * J - 15;
#pragma unroll
    for (int m = 0; m < 31; ++m) w[m] = sA[wb + m];
#pragma unroll
    for (int s = 0; s < 16; ++s)
#pragma unroll
      for (int t = 0; t < 16; ++t) acc3_add(acc[t], x[s] * w[t - s + 15]);  // exact products
  }
}
#pragma clang fp contract(on)

template <int BD>
__global__ __launch_bounds__(BD) void k_eval_exact(const float* __restrict__ pop, int n,
                                                   double* F, double* Mo, double* So, int* amb) {
  constexpr int NW = BD / kWave;
  extern __shared__ __align__(16) unsigned char smem[];
  double* base = reinterpret_cast<double*>(smem);
  double* sA = base + EPAD;
  const int NT = eval_tiles(n);
  double* tmax = base + n + 2 * EPAD;  // rounded tile maxima
  double* slots = tmax + NT;
  double* sS = slots + 2 * NW;
  const int tid = threadIdx.x;
  const float* a = pop + (size_t)blockIdx.x * n;
  for (int i = tid; i < n + 2 * EPAD; i += BD) base[i] = 0.0;
  __syncthreads();
  for (int i = tid; i < n; i += BD) sA[i] = (double)a[i];
  __syncthreads();
  // bound on the error of lo2 (the only rounded part) relative to a sum of m
  // non-negative terms: m^3 u^3 (u = 2^-53), with slack.
  const double m = (double)n;
  const double rel = (2.0 * m * m * m + 8.0) * 0x1p-159;
  if (tid == 0) {
    Acc3 acc{0.0, 0.0, 0.0, true};
    for (int i = 0; i < n; ++i) acc3_add(acc, sA[i]);
    int ab = 0;
    const double S = round_certified(acc, rel, ab);
    sS[0] = S;
    sS[1] = (double)ab;
  }
  const int nb = 2 * n - 1;
  double loc = 0.0;
  for (int q = tid; q < (NT + 1) / 2; q += BD) {
    for (int side = 0; side < 2; ++side) {
      const int tq = side ? NT - 1 - q : q;
      if (side && tq == q) break;
      Acc3 acc[16];
      exact_tile(sA, n, 16 * tq, acc);
      double mx = 0.0;
#pragma unroll
      for (int t = 0; t < 16; ++t) {
        if (16 * tq + t >= nb) continue;
        int ab = 0;
        const double r = round_certified(acc[t], rel, ab);
        mx = r > mx ? r : mx;
      }
      tmax[tq] = mx;
      loc = mx > loc ? mx : loc;
    }
  }
  const double M = block_max<double, NW>(loc, slots);
  // certify the rounding of every b_k that rounds to M
  double cnt = 0.0;
  for (int q = tid; q < NT; q += BD) {
    if (tmax[q] != M) continue;
    Acc3 acc[16];
    exact_tile(sA, n, 16 * q, acc);
#pragma unroll
    for (int t = 0; t < 16; ++t) {
      if (16 * q + t >= nb) continue;
      int ab = 0;
      const double r = round_certified(acc[t], rel, ab);
      if (r == M && ab) cnt += 1.0;
    }
  }
  const double c = block_sum<double, NW>(cnt, slots + NW);
  if (tid == 0) {
    const double S = sS[0];
    F[blockIdx.x] = S > 0.0 ? 2.0 * (double)n * M / (S * S) : 1e30;
    if (Mo) Mo[blockIdx.x] = M;
    if (So) So[blockIdx.x] = S;
    if (amb) amb[blockIdx.x] = (int)(c + sS[1] + 0.5);
  }
}

template <typename Tin>
int eval_common(const Tin* pop, long long stride, const int* ns, int n, int P, double* F,
                int* nActive, double tol, hipStream_t stream) {
  if (!pop || !F || n < 1 || P < 0 || stride < n) return SIDON_ERR_ARG;
  if (P == 0) return SIDON_OK;
  int rc = SIDON_OK;
  DeviceState* ds = device_state(&rc);
  if (!ds) return rc;
  constexpr int BD = 256;
  const long long lds = eval_lds_elems(n, BD) * (long long)sizeof(double);
  if (lds > ds->lds_max) {
    set_last_error_msg("eval: n needs more LDS than a workgroup may use");
    return SIDON_ERR_LDS;
  }
  auto kfn = k_eval<Tin, BD>;
  SIDON_HIP_TRY(hipFuncSetAttribute(reinterpret_cast<const void*>(kfn),
                                    hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  hipLaunchKernelGGL(kfn, dim3(P), dim3(BD), (size_t)lds, stream, pop, stride, ns, n, F, nActive,
                     tol);
  SIDON_HIP_TRY(hipGetLastError());
  return SIDON_OK;
}

}  // namespace
}  // namespace sidon

using namespace sidon;

extern "C" long long sidon_eval_lds_bytes_impl(int n) {
  if (n < 1) return -1;
  return eval_lds_elems(n, 256) * (long long)sizeof(double);
}

extern "C" int sidon_eval_f64(const float* pop, int n, int P, double* F, int* nActive,
                              double activeTol, hipStream_t stream) {
  return eval_common<float>(pop, n, nullptr, n, P, F, nActive, activeTol, stream);
}

extern "C" int sidon_eval_f64_ragged(const float* pop, int stride, const int* ns, int P, int nmax,
                                     double* F, int* nActive, double activeTol,
                                     hipStream_t stream) {
  if (!ns) return SIDON_ERR_ARG;
  return eval_common<float>(pop, stride, ns, nmax, P, F, nActive, activeTol, stream);
}

extern "C" int sidon_eval_f64_d(const double* pop, int n, int P, double* F, int* nActive,
                                double activeTol, hipStream_t stream) {
  return eval_common<double>(pop, n, nullptr, n, P, F, nActive, activeTol, stream);
}

extern "C" int sidon_eval_exact(const float* pop, int n, int P, double* F, double* M, double* S,
                                int* ambiguous, hipStream_t stream) {
  if (!pop || !F || n < 1 || P < 0) return SIDON_ERR_ARG;
  if (P == 0) return SIDON_OK;
  int rc = SIDON_OK;
  DeviceState* ds = device_state(&rc);
  if (!ds) return rc;
  constexpr int BD = 256;
  const long long lds = eval_lds_elems(n, BD) * (long long)sizeof(double);
  if (lds > ds->lds_max) {
    set_last_error_msg("eval_exact: n needs more LDS than a workgroup may use");
    return SIDON_ERR_LDS;
  }
  auto kfn = k_eval_exact<BD>;
  SIDON_HIP_TRY(hipFuncSetAttribute(reinterpret_cast<const void*>(kfn),
                                    hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  hipLaunchKernelGGL(kfn, dim3(P), dim3(BD), (size_t)lds, stream, pop, n, F, M, S, ambiguous);
  SIDON_HIP_TRY(hipGetLastError());
  return SIDON_OK;
}
