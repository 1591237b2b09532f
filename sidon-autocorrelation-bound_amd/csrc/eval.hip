// eval.hip -- FP64 objective: F, near-active peak count, and a correctly
// rounded F for the cell-split check.
//
// Reference: src/kernels.cuh:291-380 (kEval<T,Acc>), launch src/main.cu:161-166.
// One workgroup per candidate.  a (fp64) sits in LDS with 32 zero cells on
// each side, so no bounds test sits in the inner loop, and with one unused
// slot after every 16 heights, so lanes whose tiles lie one apart read
// addresses 17 doubles apart (no bank conflicts).  A tile is 16 consecutive
// outputs k; a lane evaluates it over aligned 16-blocks of i, one "trip" of
// 16 x 16 fp64 FMAs per 47 LDS reads.  Tiles are ranked from the middle of
// the triangle outwards (descending trip count) and dealt to the lanes in
// boustrophedon rounds (lane l of L takes ranks l, 2L-1-l, 2L+l, ...), which
// gives every lane the same number of trips to about one percent; a lane
// walks its tiles in one flat loop, so a wave never waits at a tile boundary.
// b is never stored: per-tile maxima are kept, and the active-peak pass
// recomputes only the outputs of tiles whose maximum reaches the threshold,
// one output per lane, summed in the same order.
//
// For fp32 heights every product is exact in fp64 and each b_k accumulates in
// ascending i, and S is summed sequentially, exactly as hostF
// (src/main.cu:94-105) does, so F matches hostF bit for bit.
#include <cfloat>
#include <cstdio>
#include <cmath>

#include "common.hpp"
#include "runtime.hpp"

namespace sidon {
namespace {

constexpr int EPAD = 32;
constexpr int kFlagRounds = 16;  // tiles per thread the flag compaction handles

__host__ __device__ inline int eval_tiles(int n) { return (2 * n - 1 + 15) / 16; }
// LDS slot of height x, x in [-EPAD, n + EPAD): one unused slot after every 16.
__device__ __forceinline__ int eslot(int x) {
  const int y = x + EPAD;
  return y + (y >> 4);
}
__host__ __device__ inline int eval_slots(int n) {
  const int y = n + 2 * EPAD;
  return y + (y >> 4) + 1;
}
// heights, tile maxima, 2 NW reduction slots, S / flag / counter
__host__ __device__ inline long long eval_lds_elems(int n, int block) {
  return (long long)eval_slots(n) + eval_tiles(n) + 2 * (block / kWave) + 4;
}

template <typename Tin>
__device__ __forceinline__ void load_heights(const Tin* a, int n, double* E, int tid,
                                             int nthreads) {
  const int Y = n + 2 * EPAD;
  for (int y = tid; y < Y; y += nthreads) {
    const int x = y - EPAD;
    E[y + (y >> 4)] = (x >= 0 && x < n) ? (double)a[x] : 0.0;
  }
}

// 16-blocks J0..J1 of i that tile tq (outputs 16tq .. 16tq+15) reads.
__device__ __forceinline__ void tile_blocks(int n, int tq, int& J0, int& J1) {
  const int k0 = 16 * tq;
  J0 = (k0 - (n - 1) > 0 ? k0 - (n - 1) : 0) >> 4;
  J1 = (k0 + 15 < n - 1 ? k0 + 15 : n - 1) >> 4;
}

// Tile of rank r: ranks alternate either side of the middle tile, outwards.
__device__ __forceinline__ int ranked_tile(int r, int NT) {
  const int m0 = (NT - 1) >> 1;
  return (r & 1) ? m0 + 1 + (r >> 1) : m0 - (r >> 1);
}

// Rank of the k-th tile of lane l out of L (boustrophedon rounds).
__device__ __forceinline__ int lane_rank(int k, int l, int L) {
  return k * L + ((k & 1) ? L - 1 - l : l);
}

// Lane l of L walks its tiles (ranks l, 2L-1-l, 2L+l, ...) in one flat loop
// over its trips: trip(J, d) once per 16-block J of i (d = tile - J),
// done(tile) after the tile's last trip.  The loop runs on a trip counter and
// switches tiles at the top of its body, so it has a single latch and is not
// split into a loop per tile (which would make a wave wait at every tile end).
template <class Trip, class Done>
__device__ __forceinline__ void walk_tiles(int n, int NT, int l, int L, Trip&& trip,
                                           Done&& done) {
  int total = 0;
  for (int k = 0;; ++k) {
    const int r = lane_rank(k, l, L);
    if (r >= NT) break;
    int J0, J1;
    tile_blocks(n, ranked_tile(r, NT), J0, J1);
    total += J1 - J0 + 1;
  }
  if (total == 0) return;
  int k = 0;
  int tq = ranked_tile(l, NT), J, J1;
  tile_blocks(n, tq, J, J1);
  for (int it = 0; it < total; ++it) {
    if (J > J1) {
      done(tq);
      tq = ranked_tile(lane_rank(++k, l, L), NT);
      tile_blocks(n, tq, J, J1);
    }
    trip(J, tq - J);
    ++J;
  }
  done(tq);
}

// One trip's operands: x[s] = a_{16J+s}, w[m] = a_{16d-15+m}, so that
// x[s] * w[t-s+15] = a_i a_{k-i} for i = 16J+s, k = 16(J+d)+t.
__device__ __forceinline__ void trip_operands(const double* E, int J, int d, double x[16],
                                              double w[31]) {
  const double* px = E + 17 * J + 34;  // eslot(16J)
  const double* pw = E + 17 * d + 18;  // eslot(16d-15)
#pragma unroll
  for (int s = 0; s < 16; ++s) x[s] = px[s];
#pragma unroll
  for (int m = 0; m < 31; ++m) w[m] = pw[m + (m >= 15 ? 1 : 0)];
}

__device__ __forceinline__ void eval_trip(const double* E, int J, int d, double acc[16]) {
  double x[16], w[31];
  trip_operands(E, J, d, x, w);
#pragma unroll
  for (int s = 0; s < 16; ++s)
#pragma unroll
    for (int t = 0; t < 16; ++t) acc[t] = fma(x[s], w[t - s + 15], acc[t]);
}

// b_k alone, ascending i, one FMA per term: the value eval_trip reaches (the
// zero products a tile adds before and after the range leave a sum unchanged).
__device__ double output_b(const double* E, int n, int k) {
  const int ilo = k - (n - 1) > 0 ? k - (n - 1) : 0;
  const int ihi = k < n - 1 ? k : n - 1;
  double s = 0.0;
#pragma unroll 8
  for (int i = ilo; i <= ihi; ++i) s = fma(E[eslot(i)], E[eslot(k - i)], s);
  return s;
}

// Writes the tiles q with keep(tmax[q]) to list (which may alias tmax) and
// returns their number; every thread of the workgroup calls it.  *cnt is 0 on
// entry.  NT <= kFlagRounds * nthreads.
template <class Keep>
__device__ int compact_tiles(const double* tmax, int NT, int* list, int* cnt, int tid,
                             int nthreads, Keep&& keep) {
  unsigned flags = 0;
  for (int j = 0; j < kFlagRounds; ++j) {
    const int q = tid + j * nthreads;
    if (q < NT && keep(tmax[q])) flags |= 1u << j;
  }
  __syncthreads();
  for (int j = 0; j < kFlagRounds; ++j)
    if ((flags >> j) & 1u) list[atomicAdd(cnt, 1)] = tid + j * nthreads;
  __syncthreads();
  return *cnt;
}

template <typename Tin, int BD>
__global__ __launch_bounds__(BD) void k_eval(const Tin* __restrict__ pop, long long stride,
                                             const int* __restrict__ ns, int n_uniform,
                                             double* F, int* nActive, double tol) {
  constexpr int NW = BD / kWave;
  const int n = ns ? ns[blockIdx.x] : n_uniform;
  extern __shared__ __align__(16) unsigned char smem[];
  double* E = reinterpret_cast<double*>(smem);
  const int NT = eval_tiles(n);
  double* tmax = E + eval_slots(n);
  double* slots = tmax + NT;
  double* sS = slots + 2 * NW;
  int* cnt = reinterpret_cast<int*>(sS + 2);
  const int tid = threadIdx.x;
  load_heights(pop + (size_t)blockIdx.x * stride, n, E, tid, BD);
  if (tid == 0) *cnt = 0;
  __syncthreads();
  if (tid == 0) {  // sequential S, hostF order
    double S = 0.0;
    for (int i = 0; i < n; ++i) S += E[eslot(i)];
    sS[0] = S;
  }
  double loc = 0.0;
  double acc[16];
#pragma unroll
  for (int t = 0; t < 16; ++t) acc[t] = 0.0;
  // outputs past 2n-2 read only zero cells, so they are 0 and never raise mx
  walk_tiles(
      n, NT, tid, BD, [&](int J, int d) { eval_trip(E, J, d, acc); },
      [&](int tq) {
        double mx = 0.0;
#pragma unroll
        for (int t = 0; t < 16; ++t) {
          mx = acc[t] > mx ? acc[t] : mx;
          acc[t] = 0.0;
        }
        tmax[tq] = mx;
        loc = mx > loc ? mx : loc;
      });
  const double M = block_max<double, NW>(loc, slots);
  if (nActive) {
    const int nb = 2 * n - 1;
    const double thr = M * (1.0 - tol);
    int* list = reinterpret_cast<int*>(tmax);
    const int nf = compact_tiles(tmax, NT, list, cnt, tid, BD,
                                 [thr](double v) { return v >= thr; });
    double c = 0.0;
    for (int g = tid; g < 16 * nf; g += BD) {
      const int k = 16 * list[g >> 4] + (g & 15);
      if (k < nb && output_b(E, n, k) >= thr) c += 1.0;
    }
    const double cc = block_sum<double, NW>(c, slots + NW);
    if (tid == 0) nActive[blockIdx.x] = (int)(cc + 0.5);
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

// Exact trip: triple-length sums of exact fp32 x fp32 products.
__device__ __forceinline__ void exact_trip(const double* E, int J, int d, Acc3 acc[16]) {
  double x[16], w[31];
  trip_operands(E, J, d, x, w);
#pragma unroll
  for (int s = 0; s < 16; ++s)
#pragma unroll
    for (int t = 0; t < 16; ++t) acc3_add(acc[t], x[s] * w[t - s + 15]);  // exact products
}

// b_k alone as exact_trip accumulates it (zero terms leave an Acc3 unchanged).
__device__ Acc3 output_exact(const double* E, int n, int k) {
  const int ilo = k - (n - 1) > 0 ? k - (n - 1) : 0;
  const int ihi = k < n - 1 ? k : n - 1;
  Acc3 s{0.0, 0.0, 0.0, true};
  for (int i = ilo; i <= ihi; ++i) acc3_add(s, E[eslot(i)] * E[eslot(k - i)]);
  return s;
}
#pragma clang fp contract(on)

template <int BD>
__global__ __launch_bounds__(BD) void k_eval_exact(const float* __restrict__ pop, int n,
                                                   double* F, double* Mo, double* So, int* amb) {
  constexpr int NW = BD / kWave;
  extern __shared__ __align__(16) unsigned char smem[];
  double* E = reinterpret_cast<double*>(smem);
  const int NT = eval_tiles(n);
  double* tmax = E + eval_slots(n);  // rounded tile maxima
  double* slots = tmax + NT;
  double* sS = slots + 2 * NW;
  int* cnt = reinterpret_cast<int*>(sS + 2);
  const int tid = threadIdx.x;
  load_heights(pop + (size_t)blockIdx.x * n, n, E, tid, BD);
  if (tid == 0) *cnt = 0;
  __syncthreads();
  // bound on the error of lo2 (the only rounded part) relative to a sum of m
  // non-negative terms: m^3 u^3 (u = 2^-53), with slack.
  const double m = (double)n;
  const double rel = (2.0 * m * m * m + 8.0) * 0x1p-159;
  if (tid == 0) {
    Acc3 acc{0.0, 0.0, 0.0, true};
    for (int i = 0; i < n; ++i) acc3_add(acc, E[eslot(i)]);
    int ab = 0;
    const double S = round_certified(acc, rel, ab);
    sS[0] = S;
    sS[1] = (double)ab;
  }
  const int nb = 2 * n - 1;
  double loc = 0.0;
  Acc3 acc[16];
#pragma unroll
  for (int t = 0; t < 16; ++t) acc[t] = Acc3{0.0, 0.0, 0.0, true};
  walk_tiles(
      n, NT, tid, BD, [&](int J, int d) { exact_trip(E, J, d, acc); },
      [&](int tq) {
        double mx = 0.0;
#pragma unroll
        for (int t = 0; t < 16; ++t) {
          if (16 * tq + t < nb) {
            int ab = 0;
            const double r = round_certified(acc[t], rel, ab);
            mx = r > mx ? r : mx;
          }
          acc[t] = Acc3{0.0, 0.0, 0.0, true};
        }
        tmax[tq] = mx;
        loc = mx > loc ? mx : loc;
      });
  const double M = block_max<double, NW>(loc, slots);
  // certify the rounding of every b_k that rounds to M
  int* list = reinterpret_cast<int*>(tmax);
  const int nf = compact_tiles(tmax, NT, list, cnt, tid, BD, [M](double v) { return v == M; });
  double c = 0.0;
  for (int g = tid; g < 16 * nf; g += BD) {
    const int k = 16 * list[g >> 4] + (g & 15);
    if (k >= nb) continue;
    int ab = 0;
    const double r = round_certified(output_exact(E, n, k), rel, ab);
    if (r == M && ab) c += 1.0;
  }
  const double cc = block_sum<double, NW>(c, slots + NW);
  if (tid == 0) {
    const double S = sS[0];
    F[blockIdx.x] = S > 0.0 ? 2.0 * (double)n * M / (S * S) : 1e30;
    if (Mo) Mo[blockIdx.x] = M;
    if (So) So[blockIdx.x] = S;
    if (amb) amb[blockIdx.x] = (int)(cc + sS[1] + 0.5);
  }
}

// Workgroup size: the largest of 256 / 128 / 64 lanes that still deals at
// least two tiles to every lane (with one tile a lane, most lanes wait for
// the longest tile); at small n the small workgroups keep the CU occupied
// through their small LDS footprint.
inline int eval_block(int n) {
  const int NT = eval_tiles(n);
  return NT >= 512 ? 256 : NT >= 256 ? 128 : 64;
}

template <int BD, class Kernel, class... Args>
int launch_eval(const char* what, Kernel kfn, int n, int P, hipStream_t stream, Args... args) {
  int rc = SIDON_OK;
  DeviceState* ds = device_state(&rc);
  if (!ds) return rc;
  const long long lds = eval_lds_elems(n, BD) * (long long)sizeof(double);
  if (lds > ds->lds_max) {
    static thread_local char msg[96];
    snprintf(msg, sizeof msg, "%s: n needs more LDS than a workgroup may use", what);
    set_last_error_msg(msg);
    return SIDON_ERR_LDS;
  }
  if (eval_tiles(n) > kFlagRounds * BD) {  // implied by the LDS bound; kept explicit
    set_last_error_msg("eval: n is too large for one workgroup");
    return SIDON_ERR_ARG;
  }
  SIDON_HIP_TRY(hipFuncSetAttribute(reinterpret_cast<const void*>(kfn),
                                    hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  hipLaunchKernelGGL(kfn, dim3(P), dim3(BD), (size_t)lds, stream, args...);
  SIDON_HIP_TRY(hipGetLastError());
  return SIDON_OK;
}

template <typename Tin>
int eval_common(const Tin* pop, long long stride, const int* ns, int n, int P, double* F,
                int* nActive, double tol, hipStream_t stream) {
  if (!pop || !F || n < 1 || P < 0 || stride < n) return SIDON_ERR_ARG;
  if (P == 0) return SIDON_OK;
  switch (eval_block(n)) {
    case 256:
      return launch_eval<256>("eval", k_eval<Tin, 256>, n, P, stream, pop, stride, ns, n, F,
                              nActive, tol);
    case 128:
      return launch_eval<128>("eval", k_eval<Tin, 128>, n, P, stream, pop, stride, ns, n, F,
                              nActive, tol);
    default:
      return launch_eval<64>("eval", k_eval<Tin, 64>, n, P, stream, pop, stride, ns, n, F,
                             nActive, tol);
  }
}

}  // namespace
}  // namespace sidon

using namespace sidon;

extern "C" long long sidon_eval_lds_bytes_impl(int n) {
  if (n < 1) return -1;
  return eval_lds_elems(n, eval_block(n)) * (long long)sizeof(double);
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
  switch (eval_block(n)) {
    case 256:
      return launch_eval<256>("eval_exact", k_eval_exact<256>, n, P, stream, pop, n, F, M, S,
                              ambiguous);
    case 128:
      return launch_eval<128>("eval_exact", k_eval_exact<128>, n, P, stream, pop, n, F, M, S,
                              ambiguous);
    default:
      return launch_eval<64>("eval_exact", k_eval_exact<64>, n, P, stream, pop, n, F, M, S,
                             ambiguous);
  }
}
