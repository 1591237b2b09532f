// polish.hip -- fp64 local search of a few candidates at any length, spread
// over the whole device: the fp64 polish of the best vectors.
//
// Reference: src/main.cu:798-846 (fp64 polish: kLocalSearch<double> over the
// elite, passes of rising exponent) and src/kernels.cuh:177-289 (the step).
// The fused fp64 local search keeps a, b/c and G in one workgroup's LDS and so
// stops at n = 4560 (3n doubles do not fit 160 KB at n = 8400).  Here one
// step is five kernels over global memory (L2-resident: ~3n doubles per
// candidate), and both products run on the fp64 matrix cores
// (v_mfma_f64_16x16x4f64) as blocked Toeplitz / Hankel tiles cut into work
// units that every CU can take:
//
//   unit (candidate, tile j of 16 output segments = 256 outputs, chunk of CH
//   block steps) = one wave, CH steps of four MFMAs;
//
// a workgroup holds UW consecutive chunks of one tile, stages their operands
// from L2 into LDS windows (the overlapped segment layout of the fused
// kernel) and adds its units' partial tiles in chunk order into one partial
// per chunk group.  The autoconvolution is the symmetric form (block offsets
// d below ceil(K/2) only); its reduce kernel adds the group partials of an
// output in order, then the at most 15 remaining products and the square
// term.  Step: conv units -> conv reduce (b, M) -> weights (c, R partials) ->
// corr units -> step (G, R, Adam or mirror step, clamp >= 0, S partials).
// The renormalisation a = x / S (uniform 1/n if S is not positive) is applied
// by every reader of a from the previous step's S partials, which each adds in
// the same order, so all sums are deterministic and the arithmetic is the
// fused kernel's step in fp64.
#include <cmath>
#include <cstdio>

#include "common.hpp"
#include "runtime.hpp"

namespace sidon {
namespace {

constexpr int CH = 16;      // block steps per work unit
constexpr int UW = 4;       // work units (waves, consecutive chunks) per workgroup
constexpr int RB = 256;     // threads per reduce / weights / step workgroup

struct Shape {
  int n, N16, NTc, NGc, NTg, NGg, LB;
};

__host__ __device__ inline Shape shape_of(int n) {
  Shape s;
  s.n = n;
  s.N16 = nseg16(n);                            // 16-segments of a
  s.NTc = conv_tiles(n);                        // autoconvolution tiles (256 outputs each)
  s.NGc = (s.N16 / CH + 1 + UW - 1) / UW;       // groups of UW chunks of d in [0, N16]
  s.NTg = corr_tiles(n);                        // correlation tiles (256 outputs each)
  s.NGg = ((s.N16 + 15 + CH - 1) / CH + UW - 1) / UW;  // groups of a tile's Hankel chunks
  s.LB = 16 * (16 * s.NTg + CH * UW * s.NGg) + 64;     // b / c, zero beyond 2n - 1
  return s;
}

// workspace layout (doubles unless noted), per candidate:
//   x[n] (the unnormalised heights of the last step), bc[LB], part[NGc * 256 NTc],
//   gpart[NGg * 256 NTg], mom[n], vel[n], rpart[ceil(LB / RB)], mpart[ceil(LB / RB)],
//   spart[2][ceil(n / RB)] (S partials of even / odd steps), then K u64 max bits and
//   K ints: 0 while running, 1 + the step at which M was not positive.
struct Ws {
  double *x, *bc, *part, *gpart, *mom, *vel, *rpart, *mpart, *spart;
  unsigned long long* mbits;
  int* stopped;
  long long sb, sp, sg, sn, sr, ss;  // per-candidate strides
};

__host__ __device__ inline int rblocks(const Shape& s) { return (s.LB + RB - 1) / RB; }
__host__ __device__ inline int sblocks(const Shape& s) { return (s.n + RB - 1) / RB; }

__host__ __device__ inline size_t ws_bytes(int n, int K) {
  const Shape s = shape_of(n);
  const size_t per = (size_t)n + s.LB + (size_t)s.NGc * 256 * s.NTc +
                     (size_t)s.NGg * 256 * s.NTg + 2 * (size_t)n + 2 * rblocks(s) +
                     2 * sblocks(s);
  return per * K * sizeof(double) + K * (sizeof(unsigned long long) + sizeof(int)) + 256;
}

__host__ __device__ inline Ws ws_of(void* base, int n, int K) {
  const Shape s = shape_of(n);
  Ws w;
  w.sb = s.LB;
  w.sp = (long long)s.NGc * 256 * s.NTc;
  w.sg = (long long)s.NGg * 256 * s.NTg;
  w.sn = n;
  w.sr = rblocks(s);
  w.ss = sblocks(s);
  double* d = reinterpret_cast<double*>(base);
  w.x = d; d += w.sn * K;
  w.bc = d; d += w.sb * K;
  w.part = d; d += w.sp * K;
  w.gpart = d; d += w.sg * K;
  w.mom = d; d += w.sn * K;
  w.vel = d; d += w.sn * K;
  w.rpart = d; d += w.sr * K;
  w.mpart = d; d += w.sr * K;
  w.spart = d; d += 2 * w.ss * K;
  w.mbits = reinterpret_cast<unsigned long long*>(d);
  w.stopped = reinterpret_cast<int*>(w.mbits + K);
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

// Sum of v[0 .. m) in one fixed order, the same in every lane and every
// caller (lane l adds v[l], v[l + 64], ... in ascending order, then a
// butterfly); the calling wave must have all lanes active.
__device__ __forceinline__ double fixed_sum(const double* v, int m) {
  double r = 0.0;
  for (int b = (int)threadIdx.x & (kWave - 1); b < m; b += kWave) r += v[b];
  return wave_sum(r);
}

// The normalisation of the heights entering step `it`: a_i = reset ? 1/n :
// x_i * inv with S the fixed-order sum of step it - 1's partials (the fused
// kernel's renormalisation); step 0 takes the input as it is.
struct Norm {
  double inv, uni;
  bool reset;
  __device__ __forceinline__ double of(double x) const { return reset ? uni : x * inv; }
};
__device__ __forceinline__ Norm norm_at(const Ws& w, int cand, int it, int n) {
  Norm r{1.0, 1.0 / (double)n, false};
  if (it == 0) return r;
  const double S = fixed_sum(w.spart + w.ss * (2 * cand + ((it - 1) & 1)), (int)w.ss);
  r.reset = !(S > 0.0);
  r.inv = r.reset ? 1.0 : 1.0 / S;
  return r;
}

// sum of p[c * stride] for c = c0 .. c1 in ascending c, loaded sixteen at a
// time so the loads are in flight together (absent terms add an exact zero)
__device__ __forceinline__ double chunk_sum(const double* p, long long stride, int c0, int c1) {
  double r = 0.0;
  for (int c = c0; c <= c1; c += 16) {
    double v[16];
#pragma unroll
    for (int u = 0; u < 16; ++u) v[u] = c + u <= c1 ? p[(long long)(c + u) * stride] : 0.0;
#pragma unroll
    for (int u = 0; u < 16; ++u) r += v[u];
  }
  return r;
}

// Toeplitz A-operand windows: a in the overlapped segment layout, without its
// left padding: window element x (a_{16 g + x}) sits at wpos(x) = pad_addr(x) -
// 15 kSegStride for x in [-4, 16 LEN), so the buffer holds LEN + 2 segments.
constexpr int kWinShift = 15 * kSegStride;
__host__ __device__ constexpr int win_len(int len) { return (len + 2) * kSegStride; }
__device__ __forceinline__ int wpos(int x) { return pad_addr(x) - kWinShift; }

// Matrix B-operand windows (a lane reads two pairs of heights inside one
// segment, so no copies are needed): segment x/16 of the window at 20 (x/16)
// doubles.  The 16 reduction indices of a block step are dealt to lane group
// kk and MFMA q as s = 2 kk + 8 (q / 2) + q % 2 (see polish_conv_tile), so a
// lane's B operand is two 16-byte reads at doubles 2 kk and 2 kk + 8 of its
// segment.  gfx950 serves a 16-byte LDS read 16 lanes a pass over 64
// four-byte banks; by the device's bank-conflict counters on probe patterns,
// this order with a 20-double stride keeps each pass on distinct banks, where
// four contiguous doubles per lane at an 18- or 20-double stride cost a
// second pass.  LEN segments, then one zero segment (the masked reads of the
// symmetric tile land there).
constexpr int kBStride = 20;
__host__ __device__ constexpr int bwin_len(int len) { return (len + 1) * kBStride; }
__device__ __forceinline__ int bpos(int x) { return (x >> 4) * kBStride + (x & 15); }

// Two pairs of doubles: o = p[0], p[1], p[8], p[9] (two 16-byte LDS reads).
__device__ __forceinline__ void load_pairs(const double* p, double (&o)[4]) {
  const double2 v0 = *reinterpret_cast<const double2*>(p);
  const double2 v1 = *reinterpret_cast<const double2*>(p + 8);
  o[0] = v0.x; o[1] = v0.y; o[2] = v1.x; o[3] = v1.y;
}

template <int NT, int LEN>
__device__ __forceinline__ void fill_bwindow(double* win, const double* __restrict__ xs, int n,
                                             int g, const Norm& nm) {
  constexpr int E = 16 * LEN, R = (E + NT - 1) / NT;
  double v[R];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int x = (int)threadIdx.x + r * NT;
    const int i = 16 * g + x;
    v[r] = (x < E && i >= 0 && i < n) ? nm.of(xs[i]) : 0.0;
  }
  if (threadIdx.x < kBStride) win[LEN * kBStride + threadIdx.x] = 0.0;
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int x = (int)threadIdx.x + r * NT;
    if (x < E) win[bpos(x)] = v[r];
  }
}

// Fills an A-operand window from the unnormalised heights x (normalised on the
// way in, zero outside [0, n)); NT threads, every load in flight before the
// first store.  (fill_bwindow above does the same for B-operand windows.)
template <int NT, int LEN>
__device__ __forceinline__ void fill_window(double* win, const double* __restrict__ xs, int n,
                                            int g, const Norm& nm) {
  constexpr int E = 16 * LEN + 4, R = (E + NT - 1) / NT;
  double v[R];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int x = (int)threadIdx.x + r * NT - 4;
    const int i = 16 * g + x;
    v[r] = (x < 16 * LEN && i >= 0 && i < n) ? nm.of(xs[i]) : 0.0;
  }
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int x = (int)threadIdx.x + r * NT - 4;
    if (x < 16 * LEN) {
      const int p = wpos(x);
      win[p] = v[r];
      if (((x + kPadSeg * kSeg) & 15) >= 12) win[p + 4] = v[r];  // copy heading the next segment
    }
  }
}

// Adds the UW units' 16 x 16 partial tiles of a workgroup in unit order and
// stores the sum (256 doubles at `o`, one per thread, coalesced).  `red` is
// UW * kRedTile doubles of LDS that no wave reads any more; every thread calls
// it.  Lane (kk, t) of a unit holds rows kk + 4r of column t: outputs
// 16 t + kk + 4 r.  Columns sit 17 doubles apart in `red`, so the sixteen
// columns a store instruction writes fall on distinct banks (at 16 doubles
// apart they shared two).
constexpr int kRedCol = 17, kRedTile = 16 * kRedCol;
__device__ __forceinline__ void group_sum_store(double* red, const f64x4& acc, double* o) {
  const int lane = (int)threadIdx.x & (kWave - 1), wv = (int)threadIdx.x / kWave;
  double* p = red + kRedTile * wv + kRedCol * (lane & 15) + (lane >> 4);
  p[0] = acc[0]; p[4] = acc[1]; p[8] = acc[2]; p[12] = acc[3];
  __syncthreads();
  const int k = threadIdx.x;  // UW * kWave == 256 threads, one output each
  const int q = kRedCol * (k >> 4) + (k & 15);
  double r = red[q];
#pragma unroll
  for (int u = 1; u < UW; ++u) r += red[kRedTile * u + q];
  o[k] = r;
}
static_assert(UW * kWave == 256, "one thread per tile output in group_sum_store");

__global__ void k_polish_setup(const double* __restrict__ pop, int n, Ws w) {
  const int cand = blockIdx.y;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    w.x[w.sn * cand + i] = pop[(size_t)cand * n + i];
    w.mom[w.sn * cand + i] = 0.0;
    w.vel[w.sn * cand + i] = 0.0;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    w.mbits[cand] = 0ull;
    w.stopped[cand] = 0;
  }
}

// --------------------------------------------------------- conv units ----
// Matrix share of the symmetric autoconvolution (the fused kernel's
// autoconv_tile) for output segments K = K0 + t over block offsets d in
// [da, db), from windows: wa starts at segment ga, wb at segment gb, and wb's
// zero segment follows its LENB segments.  Column t takes block d only while
// d < ceil((K0 + t) / 2); past that its B reads come from the zero segment.
// Reduction index s of lane group kk and MFMA q: s = 2 kk + 8 (q / 2) + q % 2
// (pairs of adjacent heights; see the B-window layout above).
template <int D, int LENB>
__device__ __forceinline__ f64x4 polish_conv_tile(const double* wa, int ga, const double* wb,
                                                  int gb, int K0, int da, int db) {
  constexpr int S = kSegStride;
  const int lane = opaque((int)threadIdx.x) & (kWave - 1);
  const int t = lane & 15, kk = lane >> 4;
  const int pa0 = wpos(16 * (da - ga) + t - 2 * kk);        // a_{16d + t - s}, s = 2kk (+1)
  const int pa1 = wpos(16 * (da - ga) + t - 2 * kk - 8);    // s = 2kk + 8 (+1)
  const int pb = bpos(16 * (K0 + t - da - gb)) + 2 * kk;    // a_{16(K0 + t - d) + s}
  const int live = ((K0 + t + 1) >> 1) - da;
  constexpr int zero = LENB * kBStride;
  return run_steps<double, D>(db - da, [&](int i, double (&x)[4], double (&y)[4]) {
    x[0] = wa[pa0 + i * S];
    x[1] = wa[pa0 + i * S - 1];
    x[2] = wa[pa1 + i * S];
    x[3] = wa[pa1 + i * S - 1];
    load_pairs(wb + (i < live ? pb - i * kBStride : zero), y);
  });
}

// part[cand][g][256 j + 16 t + row] = matrix share of output 16 (16 j + t) + row
// over the block offsets of tile j in chunk group g (d in [CH UW g, CH UW (g + 1))).
// Workgroup (j, g, cand); wave u takes chunk UW g + u.
__global__ __launch_bounds__(UW * kWave) void k_polish_conv_units(Shape s, Ws w, int it) {
  constexpr int LENA = CH * UW + 1, LENB = 16 + CH * UW - 1;
  __shared__ __align__(16) double wa[win_len(LENA)];
  __shared__ __align__(16) double wb[bwin_len(LENB)];
  static_assert(win_len(LENA) >= kRedTile * UW, "the group sum reuses wa");
  const int cand = blockIdx.z, g = blockIdx.y, j = blockIdx.x;
  int lo, hi;
  autoconv_drange(j, s.N16, lo, hi);
  const int d0 = CH * UW * g;
  if (!(lo <= hi && lo < d0 + CH * UW && hi >= d0)) return;  // the tile has no block here
  const double* xs = w.x + w.sn * cand;
  const Norm nm = norm_at(w, cand, it, s.n);
  const int ga = d0 - 1;                       // segments of a_{16d + t - s}
  const int gb = 16 * j - d0 - CH * UW + 1;    // segments of a_{16(K - d) + s}
  fill_window<UW * kWave, LENA>(wa, xs, s.n, ga, nm);
  fill_bwindow<UW * kWave, LENB>(wb, xs, s.n, gb, nm);
  __syncthreads();
  const int c0 = d0 + CH * wave_id();
  const int da = lo > c0 ? lo : c0;
  const int db = hi + 1 < c0 + CH ? hi + 1 : c0 + CH;
  f64x4 acc = {0.0, 0.0, 0.0, 0.0};
  if (da < db && !w.stopped[cand])
    acc = polish_conv_tile<Pipe<double>::kDepth, LENB>(wa, ga, wb, gb, 16 * j, da, db);
  __syncthreads();  // every wave is done with the windows
  group_sum_store(wa, acc, w.part + w.sp * cand + (size_t)g * 256 * s.NTc + 256 * (size_t)j);
}

// ---------------------------------------------------------- conv reduce --
// b_k = 2 (sum of the group partials of k in group order + the products
// 16(dp-1) + row < i < k/2, at most 15, in ascending i) + [k even] a_{k/2}^2
// (dp = ceil(K/2), K = k/16, row = k%16: the fused kernel's autoconv_finish).
// Workgroup x holds outputs [256 x, 256 x + 256), whose finishing operands
// lie in a_{128 x - 24 .. 128 x + 151}; its maximum goes to mpart (the weights
// take the max of those).
__global__ __launch_bounds__(RB) void k_polish_conv_reduce(Shape s, Ws w, int it) {
  constexpr int NW = RB / kWave;
  constexpr int NA = RB / 2 + 48;
  __shared__ double slots[NW];
  __shared__ double sa[NA];
  const int cand = blockIdx.y;
  if (w.stopped[cand]) return;
  const double* xs = w.x + w.sn * cand;
  const Norm nm = norm_at(w, cand, it, s.n);
  const int base = 128 * (int)blockIdx.x - 24;
  for (int e = threadIdx.x; e < NA; e += RB) {
    const int i = base + e;
    sa[e] = (i >= 0 && i < s.n) ? nm.of(xs[i]) : 0.0;
  }
  __syncthreads();
  const int k = blockIdx.x * RB + threadIdx.x;
  double b = 0.0;
  if (k < 2 * s.n - 1) {
    const int j = k >> 8;
    int lo, hi;
    autoconv_drange(j, s.N16, lo, hi);
    double r = lo <= hi ? chunk_sum(w.part + w.sp * cand + k, 256LL * s.NTc, lo / (CH * UW),
                                    hi / (CH * UW))
                        : 0.0;
    const int K = k >> 4, row = k & 15, dp = (K + 1) >> 1;
    const int i0 = 16 * (dp - 1) + row + 1;
    const int cnt = ((k + 1) >> 1) - i0;
#pragma unroll
    for (int u = 0; u < 15; ++u)
      if (u < cnt) r += sa[i0 + u - base] * sa[k - i0 - u - base];
    b = r + r;
    if (!(k & 1)) {
      const double x = sa[(k >> 1) - base];
      b += x * x;
    }
  }
  if (k < s.LB) w.bc[w.sb * cand + k] = b;
  const double m = block_max<double, NW>(b, slots);
  if (threadIdx.x == 0) w.mpart[w.sr * cand + blockIdx.x] = m;
}

// -------------------------------------------------------------- weights --
// c_k = exp((p-1) log1p((b_k - M) / M)) in place of b_k; R partial sums of
// c (1 + d) per workgroup (added in workgroup order by the step kernel).
__global__ __launch_bounds__(RB) void k_polish_weights(Shape s, Ws w, Sched S, int it) {
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
  if (!(M > 0.0)) {  // the fused kernel's break: later kernels of the call skip this candidate
    if (blockIdx.x == 0 && threadIdx.x == 0) w.stopped[cand] = it + 1;
    return;
  }
  double p, lr;
  sched_at(S, it, p, lr);
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
// Hankel tile (the fused kernel's corr_tile) of output segments I0 + t over
// steps e in [ea, eb): c from the plain window wc (element 0 = c_{cb}), a from
// the B-operand window wa starting at segment ga.
template <int D>
__device__ __forceinline__ f64x4 polish_corr_tile(const double* wc, int cb, const double* wa,
                                                  int ga, int I0, int ea, int eb) {
  const int lane = opaque((int)threadIdx.x) & (kWave - 1);
  const int t = lane & 15, kk = lane >> 4;
  const int pc = 16 * ea + t + 2 * kk - cb;                 // c_{16e + t + s}
  const int pb = bpos(16 * (ea - I0 - t - ga)) + 2 * kk;    // a_{16(e - I0 - t) + s}
  return run_steps<double, D>(eb - ea, [&](int i, double (&x)[4], double (&y)[4]) {
    x[0] = wc[pc + 16 * i];
    x[1] = wc[pc + 16 * i + 1];
    x[2] = wc[pc + 16 * i + 8];
    x[3] = wc[pc + 16 * i + 9];
    load_pairs(wa + pb + i * kBStride, y);
  });
}

// gpart[cand][g][256 j + 16 t + row] = the share of G_{16(16 j + t) + row} of
// the Hankel steps e in [16 j + CH UW g, 16 j + CH UW (g + 1)) (within the
// tile's N16 + 15).  Workgroup (j, g, cand); wave u takes chunk UW g + u.  A
// unit whose weights are all zero (most of them at high p) adds zeros.
__global__ __launch_bounds__(UW * kWave) void k_polish_corr_units(Shape s, Ws w, int it) {
  constexpr int LENC = 16 * CH * UW + 31, LENA = CH * UW + 15;
  constexpr int NT = UW * kWave, RC = (LENC + NT - 1) / NT;
  __shared__ __align__(16) double wc[LENC + 1];
  __shared__ __align__(16) double wa[bwin_len(LENA)];
  static_assert(bwin_len(LENA) >= kRedTile * UW, "the group sum reuses wa");
  const int cand = blockIdx.z, g = blockIdx.y, j = blockIdx.x;
  const double* xs = w.x + w.sn * cand;
  const double* c = w.bc + w.sb * cand;
  const Norm nm = norm_at(w, cand, it, s.n);
  const int I0 = 16 * j, e0 = I0 + CH * UW * g, eend = I0 + s.N16 + 15;
  const int cb = 16 * e0;
  const int ga = CH * UW * g - 15;
  double cv[RC];
#pragma unroll
  for (int r = 0; r < RC; ++r) {
    const int e = (int)threadIdx.x + r * NT;
    cv[r] = (e < LENC && cb + e < s.LB) ? c[cb + e] : 0.0;
  }
  fill_bwindow<NT, LENA>(wa, xs, s.n, ga, nm);
#pragma unroll
  for (int r = 0; r < RC; ++r) {
    const int e = (int)threadIdx.x + r * NT;
    if (e < LENC) wc[e] = cv[r];
  }
  __syncthreads();
  const int ea = e0 + CH * wave_id();
  const int eb = ea + CH < eend ? ea + CH : eend;
  f64x4 acc = {0.0, 0.0, 0.0, 0.0};
  if (ea < eb && !w.stopped[cand]) {
    // the weights this unit reads: wc[16 (ea - e0) .. + 16 (eb - ea) + 15)
    const int c0 = 16 * (ea - e0), cl = 16 * (eb - ea) + 15;
    bool nz = false;
    for (int e = (int)threadIdx.x & (kWave - 1); e < cl; e += kWave) nz |= wc[c0 + e] != 0.0;
    if (__ballot(nz)) acc = polish_corr_tile<Pipe<double>::kDepth>(wc, cb, wa, ga, I0, ea, eb);
  }
  __syncthreads();  // every wave is done with the windows
  group_sum_store(wa, acc, w.gpart + w.sg * cand + (size_t)g * 256 * s.NTg + 256 * (size_t)j);
}

// ----------------------------------------------------------------- step --
// Many workgroups per candidate: R (rpart, fixed order), G_i (group partials in
// order), g = G/R - M, the Adam step (lr/n) mhat / (sqrt(vhat) + 1e-30) and
// clamp >= 0 (mode 0) or the mirror step a exp(clamp(-lr g / M, +-4)) (mode
// 1) into x, and partial sums of the new heights per workgroup (the next
// step's S).  b1t, b2t are beta^(it+1), multiplied up once a step by the host
// as the fused kernel does.
__global__ __launch_bounds__(RB) void k_polish_step(Shape s, Ws w, Sched S, int mode, int it,
                                                    double b1t, double b2t) {
  constexpr int NW = RB / kWave;
  __shared__ double slots[NW];
  const int cand = blockIdx.y;
  if (w.stopped[cand]) return;
  const int n = s.n;
  const Norm nm = norm_at(w, cand, it, n);
  const double M = __longlong_as_double((long long)w.mbits[cand]);
  const double R = fixed_sum(w.rpart + w.sr * cand, (int)w.sr);
  double p, lr;
  sched_at(S, it, p, lr);
  const double bc1 = 1.0 / (1.0 - b1t), bc2 = 1.0 / (1.0 - b2t);
  const int i = blockIdx.x * RB + threadIdx.x;
  double xx = 0.0;
  if (i < n) {
    const double G = chunk_sum(w.gpart + w.sg * cand + i, 256LL * s.NTg, 0, s.NGg - 1);
    const double gi = G / R - M;
    double* xs = w.x + w.sn * cand;
    const double ai = nm.of(xs[i]);
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
    xs[i] = xx;  // this thread alone reads and writes x_i
  }
  const double part = block_sum<double, NW>(xx, slots);
  if (threadIdx.x == 0) w.spart[w.ss * (2 * cand + (it & 1)) + blockIdx.x] = part;
}

// The heights after the last step a candidate took, normalised, into pop.
__global__ void k_polish_finish(Shape s, Ws w, int iters, double* __restrict__ pop) {
  const int cand = blockIdx.y;
  const int st = w.stopped[cand];
  const Norm nm = norm_at(w, cand, st ? st - 1 : iters, s.n);
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < s.n; i += gridDim.x * blockDim.x)
    pop[(size_t)cand * s.n + i] = nm.of(w.x[w.sn * cand + i]);
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

namespace {

int check_polish_args(const double* pop, int n, int K, int iters, double pStart, double pEnd,
                      double lrStart, double lrEnd, int mode) {
  if (!pop || n < 1 || K < 0 || iters < 0) return SIDON_ERR_ARG;
  if (!(pStart >= 1.0) || !(pEnd >= 1.0) || !(lrStart > 0.0) || !(lrEnd > 0.0) ||
      (mode != 0 && mode != 1))
    return SIDON_ERR_ARG;
  return SIDON_OK;
}

// The polish's kernel sequence on a workspace of at least ws_bytes(n, K)
// bytes (arguments already checked, K and iters positive).
int polish_run(double* pop, double* mom, double* vel, int n, int K, int iters, double pStart,
               double pEnd, double lrStart, double lrEnd, double beta1, double beta2, int mode,
               void* workspace, hipStream_t stream) {
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
  const int ux = (n + 255) / 256 < 64 ? (n + 255) / 256 : 64;
  hipLaunchKernelGGL(k_polish_setup, dim3(ux, K), dim3(256), 0, stream, pop, n, w);
  SIDON_HIP_TRY(hipGetLastError());
  const dim3 gconv(s.NTc, s.NGc, K);
  const dim3 gcorr(s.NTg, s.NGg, K);
  const dim3 gred(rblocks(s), K);
  const dim3 gupd(sblocks(s), K);
  double b1t = 1.0, b2t = 1.0;
  for (int it = 0; it < iters; ++it) {
    b1t *= beta1;
    b2t *= beta2;
    hipLaunchKernelGGL(k_polish_conv_units, gconv, dim3(UW * kWave), 0, stream, s, w, it);
    hipLaunchKernelGGL(k_polish_conv_reduce, gred, dim3(RB), 0, stream, s, w, it);
    hipLaunchKernelGGL(k_polish_weights, gred, dim3(RB), 0, stream, s, w, S, it);
    hipLaunchKernelGGL(k_polish_corr_units, gcorr, dim3(UW * kWave), 0, stream, s, w, it);
    hipLaunchKernelGGL(k_polish_step, gupd, dim3(RB), 0, stream, s, w, S, mode, it, b1t, b2t);
    SIDON_HIP_TRY(hipGetLastError());
  }
  hipLaunchKernelGGL(k_polish_finish, dim3(ux, K), dim3(256), 0, stream, s, w, iters, pop);
  SIDON_HIP_TRY(hipGetLastError());
  if (mom || vel) {
    hipLaunchKernelGGL(k_polish_moments, dim3(64), dim3(256), 0, stream, w, n, K, mom, vel);
    SIDON_HIP_TRY(hipGetLastError());
  }
  return SIDON_OK;
}

}  // namespace

extern "C" int sidon_polish_f64(double* pop, double* mom, double* vel, int n, int K, int iters,
                                double pStart, double pEnd, double lrStart, double lrEnd,
                                double beta1, double beta2, int mode, void* workspace,
                                size_t workspace_bytes, hipStream_t stream) {
  if (int rc = check_polish_args(pop, n, K, iters, pStart, pEnd, lrStart, lrEnd, mode)) return rc;
  if (!workspace) return SIDON_ERR_ARG;
  if (K == 0 || iters == 0) return SIDON_OK;
  if (workspace_bytes < ws_bytes(n, K)) {
    set_last_error_msg("polish: workspace smaller than sidon_polish_f64_workspace(n, K)");
    return SIDON_ERR_ARG;
  }
  return polish_run(pop, mom, vel, n, K, iters, pStart, pEnd, lrStart, lrEnd, beta1, beta2, mode,
                    workspace, stream);
}

// ------------------------------------------------------------ no-alloc --
// sidon_polish_f64_noalloc takes its workspace from `scratch`, a file-scope
// device array: every GPU the library runs on holds its own copy until the
// process ends, and every call on that GPU shares it (calls that overlap on
// one GPU race).  Host code never passes the array's name to a kernel (on
// the host it is a stand-in, not a device address): sidon_init resolves each
// device's address once with hipGetSymbolAddress, and the entry point uses
// that address.  The kernels rewrite every part of the workspace they read
// (the setup kernel the heights, moments and flags; each step its b / c,
// partials and sums), so what a previous call left there does no harm.
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace sidon {
int noalloc_scratch_address(void** out) {
  void* p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess) {
    set_last_error("hipGetSymbolAddress(scratch)", e);
    return SIDON_ERR_HIP;
  }
  *out = p;
  return SIDON_OK;
}
}  // namespace sidon

extern "C" int sidon_polish_f64_noalloc(double* pop, double* mom, double* vel, int n, int K,
                                        int iters, double pStart, double pEnd, double lrStart,
                                        double lrEnd, double beta1, double beta2, int mode,
                                        hipStream_t stream) {
  if (int rc = check_polish_args(pop, n, K, iters, pStart, pEnd, lrStart, lrEnd, mode)) return rc;
  if (K == 0 || iters == 0) return SIDON_OK;
  if (ws_bytes(n, K) > kNoAllocScratchBytes) {
    char msg[200];
    std::snprintf(msg, sizeof msg,
                  "polish (no-alloc): n = %d, K = %d need %zu bytes of scratch, above the "
                  "%zu-byte static array (use sidon_polish_f64 with a workspace)",
                  n, K, ws_bytes(n, K), kNoAllocScratchBytes);
    set_last_error_msg(msg);
    return SIDON_ERR_ARG;
  }
  int rc = SIDON_OK;
  DeviceState* ds = device_state(&rc);
  if (!ds) return rc;
  return polish_run(pop, mom, vel, n, K, iters, pStart, pEnd, lrStart, lrEnd, beta1, beta2, mode,
                    ds->scratch, stream);
}
