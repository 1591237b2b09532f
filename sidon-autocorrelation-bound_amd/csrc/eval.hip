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
// The workgroup is as wide as still deals two tiles a lane (up to 1024
// lanes), for four waves a SIMD at the LDS footprint of one candidate.  b is
// never stored: per-tile maxima are kept, and the outputs that may reach the
// active-peak threshold are listed as their tiles finish (see k_eval).
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
// heights; NT + 16 doubles of tile maxima (and, in k_eval, candidate
// values); 2 NW reduction slots; S, running maximum and counters
__host__ __device__ inline long long eval_lds_elems(int n, int block) {
  return (long long)eval_slots(n) + eval_tiles(n) + 16 + 2 * (block / kWave) + 4;
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

// Ascending sequential sum over i of term(x_i, y_i) with x_i = a_i and
// y_i = a_{k-i} (or 1 when k < 0), i from i0 to i1 rounded up to a multiple
// of 8: heights past either end read zero cells, and a zero term leaves the
// sum unchanged.  The next 8 pairs load while the current 8 are summed, so
// the chain waits on the adds, not on LDS.
template <class Acc, class Add>
__device__ __forceinline__ void seq_sum(const double* E, int k, int i0, int i1, Acc& s,
                                        Add&& add) {
  double x[8], y[8], nx[8], ny[8];
  auto load = [&](int i, double* X, double* Y) {
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      X[j] = E[eslot(i + j)];
      Y[j] = k < 0 ? 1.0 : E[eslot(k - i - j)];
    }
  };
  load(i0, nx, ny);
  for (int i = i0; i <= i1; i += 8) {
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      x[j] = nx[j];
      y[j] = ny[j];
    }
    if (i + 8 <= i1) load(i + 8, nx, ny);
#pragma unroll
    for (int j = 0; j < 8; ++j) add(s, x[j], y[j]);
  }
}

// b_k alone, ascending i, one FMA per term: the value eval_trip reaches (the
// zero products a tile adds before and after the range leave a sum unchanged).
__device__ double output_b(const double* E, int n, int k) {
  const int ilo = k - (n - 1) > 0 ? k - (n - 1) : 0;
  const int ihi = k < n - 1 ? k : n - 1;
  double s = 0.0;
  seq_sum(E, k, ilo, ihi, s, [](double& acc, double x, double y) { acc = fma(x, y, acc); });
  return s;
}

// S = a_0 + a_1 + ... in sequence, as hostF sums it.
__device__ double height_sum(const double* E, int n) {
  double s = 0.0;
  seq_sum(E, -1, 0, n - 1, s, [](double& acc, double x, double) { acc += x; });
  return s;
}

// Writes the tiles q with keep(tmax[q]) to list (which may alias tmax) and
// returns their number; every thread of the workgroup calls it.  *cnt is 0 on
// entry.  NT <= kFlagRounds * nthreads.
template <class T, class Keep>
__device__ int compact_tiles(const T* tmax, int NT, int* list, int* cnt, int tid, int nthreads,
                             Keep&& keep) {
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

// The active count needs the b_k at or above thr = M (1 - tol), M known only
// once every tile is done.  A tile keeps, as it finishes, the outputs at or
// above max(its maximum, the running maximum) * (1 - tol) in a candidate
// list: both maxima are at most M, so the list holds every output that
// reaches thr, with its value as computed, and the count needs no second
// pass.  When the list overflows (or 0 <= tol <= 1 does not hold), the
// outputs of the tiles whose maximum reaches thr are summed again instead;
// tile maxima are kept as floats rounded up for that test.
template <typename Tin, int BD>
__global__ __launch_bounds__(BD) void k_eval(const Tin* __restrict__ pop, long long stride,
                                             const int* __restrict__ ns, int n_uniform,
                                             double* F, int* nActive, double tol) {
  constexpr int NW = BD / kWave;
  const int n = ns ? ns[blockIdx.x] : n_uniform;
  extern __shared__ __align__(16) unsigned char smem[];
  double* E = reinterpret_cast<double*>(smem);
  const int NT = eval_tiles(n);
  const int nb = 2 * n - 1;
  float* tmaxf = reinterpret_cast<float*>(E + eval_slots(n));  // NT floats
  double* cand = E + eval_slots(n) + (NT + 1) / 2;
  const int C = NT - (NT + 1) / 2 + 16;
  double* slots = cand + C;
  double* sS = slots + 2 * NW;  // S, running maximum, counters
  double* mrun = sS + 1;
  int* cnt = reinterpret_cast<int*>(sS + 2);  // [0] tile list, [1] candidates
  const int tid = threadIdx.x;
  load_heights(pop + (size_t)blockIdx.x * stride, n, E, tid, BD);
  if (tid == 0) {
    cnt[0] = 0;
    cnt[1] = 0;
    *mrun = 0.0;
  }
  __syncthreads();
  const bool listed = nActive && tol >= 0.0 && tol <= 1.0;
  const double keep = 1.0 - tol;
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
        for (int t = 0; t < 16; ++t) mx = acc[t] > mx ? acc[t] : mx;
        float mf = (float)mx;
        if ((double)mf < mx) mf = nextafterf(mf, HUGE_VALF);
        tmaxf[tq] = mf;
        loc = mx > loc ? mx : loc;
        if (listed) {
          const double old = __longlong_as_double((long long)atomicMax(
              reinterpret_cast<unsigned long long*>(mrun),
              static_cast<unsigned long long>(__double_as_longlong(mx))));
          const double lb = (old > mx ? old : mx) * keep;
#pragma unroll
          for (int t = 0; t < 16; ++t)
            if (16 * tq + t < nb && acc[t] >= lb) {
              const int pos = atomicAdd(&cnt[1], 1);
              if (pos < C) cand[pos] = acc[t];
            }
        }
#pragma unroll
        for (int t = 0; t < 16; ++t) acc[t] = 0.0;
      });
  // sequential S, hostF order, after the tiles: the older waves of a SIMD win
  // the issue arbitration and finish their tiles well before the younger
  // ones, so the sum hides in that wait
  if (tid == 0) sS[0] = height_sum(E, n);
  const double M = block_max<double, NW>(loc, slots);
  if (nActive) {
    const double thr = M * keep;
    const int nc = cnt[1];
    double c = 0.0;
    if (listed && nc <= C) {
      for (int j = tid; j < nc; j += BD) c += cand[j] >= thr ? 1.0 : 0.0;
    } else {
      int* list = reinterpret_cast<int*>(tmaxf);
      const int nf = compact_tiles(tmaxf, NT, list, cnt, tid, BD,
                                   [thr](float v) { return (double)v >= thr; });
      for (int g = tid; g < 16 * nf; g += BD) {
        const int k = 16 * list[g >> 4] + (g & 15);
        if (k < nb && output_b(E, n, k) >= thr) c += 1.0;
      }
    }
    const double cc = block_sum<double, NW>(c, slots + NW);
    if (tid == 0) nActive[blockIdx.x] = (int)(cc + 0.5);
  }
  if (tid == 0) {
    const double S = sS[0];
    F[blockIdx.x] = S > 0.0 ? 2.0 * (double)n * M / (S * S) : 1e30;
  }
}

// ------------------------------------------------------- packed, small n --
// With fewer than 128 tiles a 64-lane workgroup deals one tile a lane, and
// the lanes wait for the longest (tile work is a triangle: 1 .. n/16 trips).
// Here G candidates share a 64-lane workgroup, L = 64 / G lanes each, so every
// lane walks at least two tiles in the balanced order of k_eval.  Each
// candidate has its own LDS region with k_eval's layout (sized for n_region,
// the largest length); per output the arithmetic is k_eval's, and the
// reductions run over the candidate's L lanes (xor butterflies below L stay
// inside the group).
template <int L>
__device__ __forceinline__ double lanes_max(double v) {
#pragma unroll
  for (int o = L / 2; o > 0; o >>= 1) {
    const double w = __shfl_xor(v, o, kWave);
    v = w > v ? w : v;
  }
  return v;
}
template <int L>
__device__ __forceinline__ double lanes_sum(double v) {
#pragma unroll
  for (int o = L / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
  return v;
}

template <typename Tin, int G>
__global__ __launch_bounds__(kWave) void k_eval_packed(const Tin* __restrict__ pop,
                                                       long long stride,
                                                       const int* __restrict__ ns, int n_region,
                                                       int P, double* F, int* nActive,
                                                       double tol) {
  constexpr int L = kWave / G;
  extern __shared__ __align__(16) unsigned char smem[];
  const int tid = threadIdx.x, g = tid / L, l = tid % L;
  const int c = G * (int)blockIdx.x + g;
  const bool live = c < P;
  const int n = live ? (ns ? ns[c] : n_region) : 1;
  double* E = reinterpret_cast<double*>(smem) + (long long)g * eval_lds_elems(n_region, kWave);
  const int NT = eval_tiles(n);
  const int nb = 2 * n - 1;
  float* tmaxf = reinterpret_cast<float*>(E + eval_slots(n));  // NT floats
  double* cand = E + eval_slots(n) + (NT + 1) / 2;
  const int C = NT - (NT + 1) / 2 + 16;
  double* mrun = cand + C + 2;                  // (k_eval's two reduction slots are unused)
  int* cnt = reinterpret_cast<int*>(mrun + 1);  // [1] candidates
  if (live) load_heights(pop + (size_t)c * stride, n, E, l, L);
  if (l == 0) {
    cnt[1] = 0;
    *mrun = 0.0;
  }
  __syncthreads();
  const bool listed = nActive && tol >= 0.0 && tol <= 1.0;
  const double keep = 1.0 - tol;
  double loc = 0.0;
  double acc[16];
#pragma unroll
  for (int t = 0; t < 16; ++t) acc[t] = 0.0;
  if (live)
    walk_tiles(
        n, NT, l, L, [&](int J, int d) { eval_trip(E, J, d, acc); },
        [&](int tq) {
          double mx = 0.0;
#pragma unroll
          for (int t = 0; t < 16; ++t) mx = acc[t] > mx ? acc[t] : mx;
          float mf = (float)mx;
          if ((double)mf < mx) mf = nextafterf(mf, HUGE_VALF);
          tmaxf[tq] = mf;
          loc = mx > loc ? mx : loc;
          if (listed) {
            const double old = __longlong_as_double((long long)atomicMax(
                reinterpret_cast<unsigned long long*>(mrun),
                static_cast<unsigned long long>(__double_as_longlong(mx))));
            const double lb = (old > mx ? old : mx) * keep;
#pragma unroll
            for (int t = 0; t < 16; ++t)
              if (16 * tq + t < nb && acc[t] >= lb) {
                const int pos = atomicAdd(&cnt[1], 1);
                if (pos < C) cand[pos] = acc[t];
              }
          }
#pragma unroll
          for (int t = 0; t < 16; ++t) acc[t] = 0.0;
        });
  const double S = (live && l == 0) ? height_sum(E, n) : 0.0;  // sequential, hostF order
  const double M = lanes_max<L>(loc);
  __syncthreads();  // every tile's list entries and maxima are in LDS
  if (nActive) {
    const double thr = M * keep;
    const int nc = cnt[1];
    double k_on = 0.0;
    if (live) {
      if (listed && nc <= C) {
        for (int j = l; j < nc; j += L) k_on += cand[j] >= thr ? 1.0 : 0.0;
      } else {
        for (int q = l; q < NT; q += L)
          if ((double)tmaxf[q] >= thr)
            for (int t = 0; t < 16; ++t) {
              const int k = 16 * q + t;
              if (k < nb && output_b(E, n, k) >= thr) k_on += 1.0;
            }
      }
    }
    const double total = lanes_sum<L>(k_on);
    if (live && l == 0) nActive[c] = (int)(total + 0.5);
  }
  if (live && l == 0) F[c] = S > 0.0 ? 2.0 * (double)n * M / (S * S) : 1e30;
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
  seq_sum(E, k, ilo, ihi, s, [](Acc3& acc, double x, double y) { acc3_add(acc, x * y); });
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
  double* slots = tmax + NT + 16;
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

// Workgroup size: the largest of 1024 / 512 / 256 / 128 / 64 lanes (up to
// maxBD) that still deals at least two tiles to every lane (with one tile a
// lane, most lanes wait for the longest tile).  Large workgroups give more
// waves per CU at the same LDS footprint (the fp64 FMAs need about four waves
// a SIMD); at small n the small workgroups keep the CU occupied through their
// small LDS footprint.
inline int eval_block(int n, int maxBD) {
  const int NT = eval_tiles(n);
  int bd = 64;
  while (bd < maxBD && NT >= 4 * bd) bd *= 2;
  return bd;
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

// G candidates a 64-lane workgroup (k_eval_packed) for n with fewer than 128
// tiles: the smallest G with at least two tiles a lane (G <= 8).
template <int G, typename Tin>
int launch_packed(const Tin* pop, long long stride, const int* ns, int n, int P, double* F,
                  int* nActive, double tol, hipStream_t stream) {
  int rc = SIDON_OK;
  DeviceState* ds = device_state(&rc);
  if (!ds) return rc;
  const long long lds = G * eval_lds_elems(n, kWave) * (long long)sizeof(double);
  if (lds > ds->lds_max) return SIDON_ERR_LDS;  // not reached: n < 1032 here
  auto kfn = k_eval_packed<Tin, G>;
  SIDON_HIP_TRY(hipFuncSetAttribute(reinterpret_cast<const void*>(kfn),
                                    hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  hipLaunchKernelGGL(kfn, dim3((P + G - 1) / G), dim3(kWave), (size_t)lds, stream, pop, stride,
                     ns, n, P, F, nActive, tol);
  SIDON_HIP_TRY(hipGetLastError());
  return SIDON_OK;
}

template <typename Tin>
int eval_common(const Tin* pop, long long stride, const int* ns, int n, int P, double* F,
                int* nActive, double tol, hipStream_t stream) {
  if (!pop || !F || n < 1 || P < 0 || stride < n) return SIDON_ERR_ARG;
  if (P == 0) return SIDON_OK;
  const int NT = eval_tiles(n);
  if (NT < 32) return launch_packed<8>(pop, stride, ns, n, P, F, nActive, tol, stream);
  if (NT < 64) return launch_packed<4>(pop, stride, ns, n, P, F, nActive, tol, stream);
  if (NT < 128) return launch_packed<2>(pop, stride, ns, n, P, F, nActive, tol, stream);
  switch (eval_block(n, 1024)) {
    case 1024:
      return launch_eval<1024>("eval", k_eval<Tin, 1024>, n, P, stream, pop, stride, ns, n, F,
                               nActive, tol);
    case 512:
      return launch_eval<512>("eval", k_eval<Tin, 512>, n, P, stream, pop, stride, ns, n, F,
                              nActive, tol);
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
  return eval_lds_elems(n, eval_block(n, 1024)) * (long long)sizeof(double);
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
  // the exact accumulators take about 220 VGPRs: at most two waves a SIMD
  switch (eval_block(n, 512)) {
    case 512:
      return launch_eval<512>("eval_exact", k_eval_exact<512>, n, P, stream, pop, n, F, M, S,
                              ambiguous);
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
