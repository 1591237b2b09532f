// common.hpp -- device building blocks shared by every Sidon kernel (gfx950).
//
// Objective: F(a) = 2n * max_k b_k / (sum a)^2, b = a*a (autoconvolution),
// a >= 0.  The hot loops are the autoconvolution b_k = sum_i a_i a_{k-i} and
// its adjoint, the correlation G_i = sum_j c_{i+j} a_j
// (reference: src/kernels.cuh:92-158).
//
// MI355X design (not a translation of the reference's warp32 register tiles):
//  * wave64 butterfly reductions; every lane ends with the identical value, so
//    block reductions are deterministic;
//  * both O(n^2) products run on the matrix cores as blocked Toeplitz /
//    Hankel GEMMs (v_mfma_f32_16x16x4f32, v_mfma_f64_16x16x4f64):
//      conv:  b[16K + t] = sum_d sum_s a[16d + t - s] * a[16(K-d) + s]
//             -> for a fixed block offset d the 16x16 Toeplitz block
//                H_d[t][s] = a[16d + t - s] multiplies 16 consecutive input
//                segments at once (one MFMA tile = 16 output segments);
//      corr:  G[16I + t] = sum_e sum_s c[16e + t + s] * a[16(e-I) + s]
//             -> Hankel block C_e[t][s] = c[16e + t + s] times 16 segments;
//  * the vector `a` lives in LDS in an overlapped segment layout: 20
//    elements per 16-element segment (the last four heights of the previous
//    segment, then its own sixteen), 16 zero segments on each side.  The
//    MFMA B-operand reads (one 16-byte read per lane, stride 16 heights
//    across lanes) hit disjoint banks, the four Toeplitz A-operand heights
//    of a lane are always four adjacent elements, and every out-of-range
//    index reads an exact zero;
//  * the work of each product is cut into W equal contiguous ranges of
//    (tile, block-offset) units, one per wave ("stream-K").  A tile is never
//    split among more than two waves, and two partial sums added onto an
//    exactly-zero buffer commute, so results are bit-reproducible no matter
//    which wave finishes first.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace sidon {

constexpr int kWave = 64;
constexpr int kSeg = 16;          // MFMA tile edge / segment length
constexpr int kSegStride = 20;    // padded elements per segment in LDS (16-byte aligned)
constexpr int kPadSeg = 16;       // zero segments on each side of `a`

typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef double f64x4 __attribute__((ext_vector_type(4)));

// 16x16x4 accumulator and its lane layout: lane (kk = lane/16, t = lane%16)
// holds output rows kRowMul*kk + kRowStride*r (r = 0..3) of column t.
template <typename T> struct Acc4;
template <> struct Acc4<float> {
  using type = f32x4;
  static constexpr int kRowMul = 4, kRowStride = 1;
};
template <> struct Acc4<double> {
  using type = f64x4;
  static constexpr int kRowMul = 1, kRowStride = 4;
};

__device__ __forceinline__ f32x4 mfma16x16x4(float a, float b, f32x4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ f64x4 mfma16x16x4(double a, double b, f64x4 c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

// Wave index as a wave-uniform (scalar) value, so loops over a wave's share
// of tiles compile to scalar control flow.
__device__ __forceinline__ int wave_id() {
  return __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
}

// ------------------------------------------------------------- reductions --
template <typename T>
__device__ __forceinline__ T wave_max(T v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    const T w = __shfl_xor(v, o, kWave);
    v = w > v ? w : v;
  }
  return v;
}
template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
  return v;
}

// Block reductions.  `slots` holds NW values; callers alternate between two
// slot arrays so that no trailing barrier is needed (a slot array is only
// rewritten after every thread has passed the barrier of the reduction that
// used the other array).  Result is identical in every thread.
template <typename T, int NW>
__device__ __forceinline__ T block_sum(T v, T* slots) {
  v = wave_sum(v);
  if ((threadIdx.x & (kWave - 1)) == 0) slots[threadIdx.x / kWave] = v;
  __syncthreads();
  T r = slots[0];
#pragma unroll
  for (int i = 1; i < NW; ++i) r += slots[i];
  return r;
}
template <typename T, int NW>
__device__ __forceinline__ T block_max(T v, T* slots) {
  v = wave_max(v);
  if ((threadIdx.x & (kWave - 1)) == 0) slots[threadIdx.x / kWave] = v;
  __syncthreads();
  T r = slots[0];
#pragma unroll
  for (int i = 1; i < NW; ++i) r = slots[i] > r ? slots[i] : r;
  return r;
}

// ------------------------------------------------------------ LDS layout --
// Overlapped segment layout of a length-n vector: segment g (heights
// 16(g-16) .. 16(g-16)+15) occupies elements [20g, 20g+20): first copies of
// the previous segment's last four heights, then its own sixteen.  Element
// x (x may be negative or >= n) lives at pad_addr(x), so x-3 .. x are the
// four elements pad_addr(x)-3 .. pad_addr(x); valid for x in
// [-256, 16*(N16+16)).  Writes go through pad_put, which also refreshes the
// copy held by the next segment.
__host__ __device__ __forceinline__ int pad_addr(int x) {
  const int y = x + kPadSeg * kSeg;
  return (y >> 4) * kSegStride + 4 + (y & 15);
}
template <typename T>
__device__ __forceinline__ void pad_put(T* s, int x, T v) {
  const int p = pad_addr(x);
  s[p] = v;
  if (((x + kPadSeg * kSeg) & 15) >= 12) s[p + 4] = v;  // copy at the head of the next segment
}
__host__ __device__ __forceinline__ int nseg16(int n) { return (n + kSeg - 1) / kSeg; }
// floats (elements) of the padded vector buffer
__host__ __device__ __forceinline__ int pad_len(int n) {
  return (nseg16(n) + 2 * kPadSeg) * kSegStride;
}
// elements of the weight / autoconvolution buffer (normal layout, zero tail)
__host__ __device__ __forceinline__ int cbuf_len(int n) { return kSeg * (2 * nseg16(n) + 16); }
// correlation output tiles and buffer
__host__ __device__ __forceinline__ int corr_tiles(int n) { return (nseg16(n) + 15) / 16; }
__host__ __device__ __forceinline__ int gbuf_len(int n) { return 256 * corr_tiles(n); }
// autoconvolution output tiles (16 output segments each)
__host__ __device__ __forceinline__ int conv_tiles(int n) {
  const int NK = (2 * n - 1 + 15) / 16;
  return (NK + 15) / 16;
}

// Block offsets d of the Toeplitz blocks that tile j (output segments
// K in [16j, 16j+16)) needs: d in [lo, hi].
__device__ __forceinline__ void conv_drange(int j, int N16, int& lo, int& hi) {
  const int K0 = 16 * j;
  lo = K0 - N16 + 1 > 0 ? K0 - N16 + 1 : 0;
  hi = K0 + 15 < N16 ? K0 + 15 : N16;
}
__device__ __forceinline__ int conv_cost(int j, int N16) {
  int lo, hi;
  conv_drange(j, N16, lo, hi);
  return hi - lo + 1;
}
// corr tile j (outputs I in [16j, 16j+16)) needs Hankel blocks e in
// [16j, 16j + 15 + N16 - 1].
__device__ __forceinline__ int corr_cost(int N16) { return N16 + 15; }

// A wave's share of a product: units [u0, u1) of the concatenated tile costs.
struct Share {
  int u0, u1;     // unit range
  int j, base;    // first tile overlapping u0 and its starting unit
  int active;     // 0: this wave has no share
};

// Splits total units U over at most `nwaves` waves so that no tile (cost <=
// maxcost) is shared by more than two waves: each share is >= maxcost.
__device__ __forceinline__ Share make_share(int U, int maxcost, int wave, int nwaves) {
  Share s;
  int W = maxcost > 0 ? U / maxcost : 1;
  W = W < 1 ? 1 : (W > nwaves ? nwaves : W);
  s.active = wave < W;
  s.u0 = (int)((long long)U * wave / W);
  s.u1 = (int)((long long)U * (wave + 1) / W);
  s.j = 0;
  s.base = 0;
  return s;
}

// ------------------------------------------------- MFMA tile accumulators --
// The 16 reduction indices s of a block are dealt to the four MFMAs as
// s = 4*kk + q (kk = lane/16 is the operand's k row, q = 0..3 the MFMA), so
// the four B-operand values a lane feeds are four consecutive heights of one
// segment: a single 16-byte LDS read (segments are 20 elements apart, so the
// 16 columns of a row hit disjoint banks).  Operand reads for step d+1 are
// issued before the MFMAs of step d, and even / odd steps accumulate into
// separate registers so consecutive MFMAs do not wait on each other.

__device__ __forceinline__ void load4(const float* p, float (&o)[4]) {
  const float4 v = *reinterpret_cast<const float4*>(p);
  o[0] = v.x; o[1] = v.y; o[2] = v.z; o[3] = v.w;
}
__device__ __forceinline__ void load4(const double* p, double (&o)[4]) {
  const double2 v0 = *reinterpret_cast<const double2*>(p);
  const double2 v1 = *reinterpret_cast<const double2*>(p + 2);
  o[0] = v0.x; o[1] = v0.y; o[2] = v1.x; o[3] = v1.y;
}

// Two independent accumulations issued MFMA by MFMA in turn: a wave running
// alone on its SIMD (the tail of a phase) still has two MFMAs in flight.
template <typename A4, typename T>
__device__ __forceinline__ void mfma4x2(const T (&a0)[4], const T (&b0)[4], A4& acc0,
                                        const T (&a1)[4], const T (&b1)[4], A4& acc1) {
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    acc0 = mfma16x16x4(a0[q], b0[q], acc0);
    acc1 = mfma16x16x4(a1[q], b1[q], acc1);
  }
}
template <typename A4, typename T>
__device__ __forceinline__ A4 mfma4(const T (&a)[4], const T (&b)[4], A4 acc) {
#pragma unroll
  for (int q = 0; q < 4; ++q) acc = mfma16x16x4(a[q], b[q], acc);
  return acc;
}

// Toeplitz tile: the 16x16 output block D[row][t] = (sA (*) sB)[16(K0+t) + row]
// accumulated over block offsets d in [da, db) (rows per lane: Acc4<T>).
//   A operand (lane t, kk; MFMA q): H_d[t][s] = sA[16d + t - s]
//   B operand:                      sB[16(K0 + t - d) + s]      (column = t)
// Steps run in pairs (d -> acc0, d+1 -> acc1) with two operand sets used in
// turn: the next pair's operands are read while the current pair multiplies.
// Nothing is read outside [da, db).  LDS addresses are kept as integer
// element indices so the reads stay ds_read.
template <typename T>
__device__ __forceinline__ typename Acc4<T>::type conv_tile(const T* __restrict__ sA,
                                                           const T* __restrict__ sB, int K0,
                                                           int da, int db) {
  using A4 = typename Acc4<T>::type;
  constexpr int S = kSegStride;
  const int lane = threadIdx.x & (kWave - 1);
  const int t = lane & 15, kk = lane >> 4;
  A4 acc0 = {0, 0, 0, 0}, acc1 = {0, 0, 0, 0};
  int pa = pad_addr(16 * da + t - 4 * kk);  // A operand q at pa - q
  int pb = pad_addr(16 * (K0 + t - da)) + 4 * kk;
  // operands of the pair of steps starting `off` steps after the current one
  auto load_pair = [&](int off, T (&x0)[4], T (&y0)[4], T (&x1)[4], T (&y1)[4]) {
#pragma unroll
    for (int q = 0; q < 4; ++q) { x0[q] = sA[pa + off * S - q]; x1[q] = sA[pa + (off + 1) * S - q]; }
    load4(sB + pb - off * S, y0);
    load4(sB + pb - (off + 1) * S, y1);
  };
  const int cnt = db - da, pairs = cnt >> 1;
  T a0[4], b0[4], a1[4], b1[4], a2[4], b2[4], a3[4], b3[4];
  if (pairs > 0) {
    load_pair(0, a0, b0, a1, b1);
    int k = 0;
    for (; k + 2 < pairs; k += 2) {
      load_pair(2, a2, b2, a3, b3);
      __builtin_amdgcn_sched_barrier(0);  // keep the prefetch ahead of the MFMAs
      mfma4x2(a0, b0, acc0, a1, b1, acc1);
      __builtin_amdgcn_sched_barrier(0);
      load_pair(4, a0, b0, a1, b1);
      __builtin_amdgcn_sched_barrier(0);
      mfma4x2(a2, b2, acc0, a3, b3, acc1);
      __builtin_amdgcn_sched_barrier(0);
      pa += 4 * S;
      pb -= 4 * S;
    }
    if (k + 1 < pairs) {
      load_pair(2, a2, b2, a3, b3);
      __builtin_amdgcn_sched_barrier(0);
      mfma4x2(a0, b0, acc0, a1, b1, acc1);
      mfma4x2(a2, b2, acc0, a3, b3, acc1);
      pa += 4 * S;
      pb -= 4 * S;
    } else {
      mfma4x2(a0, b0, acc0, a1, b1, acc1);
      pa += 2 * S;
      pb -= 2 * S;
    }
  }
  if (cnt & 1) {
#pragma unroll
    for (int q = 0; q < 4; ++q) a0[q] = sA[pa - q];
    load4(sB + pb, b0);
    acc0 = mfma4(a0, b0, acc0);
  }
  return acc0 + acc1;
}

// Hankel tile: D[row][t] = sum over e in [ea, eb) of
// sum_s c[16e + row + s] * a[16(e - I0 - t) + s]  = G[16(I0 + t) + row].
//   A operand: C_e[t][s] = c[16e + t + s]       (four consecutive elements)
//   B operand: sA[16(e - I0 - t) + s]           (column = t)
template <typename T>
__device__ __forceinline__ typename Acc4<T>::type corr_tile(const T* __restrict__ c,
                                                           const T* __restrict__ sA, int I0,
                                                           int ea, int eb) {
  using A4 = typename Acc4<T>::type;
  constexpr int S = kSegStride;
  const int lane = threadIdx.x & (kWave - 1);
  const int t = lane & 15, kk = lane >> 4;
  A4 acc0 = {0, 0, 0, 0}, acc1 = {0, 0, 0, 0};
  int pc = 16 * ea + t + 4 * kk;
  int pb = pad_addr(16 * (ea - I0 - t)) + 4 * kk;
  auto load_pair = [&](int off, T (&x0)[4], T (&y0)[4], T (&x1)[4], T (&y1)[4]) {
#pragma unroll
    for (int q = 0; q < 4; ++q) { x0[q] = c[pc + 16 * off + q]; x1[q] = c[pc + 16 * (off + 1) + q]; }
    load4(sA + pb + off * S, y0);
    load4(sA + pb + (off + 1) * S, y1);
  };
  const int cnt = eb - ea, pairs = cnt >> 1;
  T a0[4], b0[4], a1[4], b1[4], a2[4], b2[4], a3[4], b3[4];
  if (pairs > 0) {
    load_pair(0, a0, b0, a1, b1);
    int k = 0;
    for (; k + 2 < pairs; k += 2) {
      load_pair(2, a2, b2, a3, b3);
      __builtin_amdgcn_sched_barrier(0);  // keep the prefetch ahead of the MFMAs
      mfma4x2(a0, b0, acc0, a1, b1, acc1);
      __builtin_amdgcn_sched_barrier(0);
      load_pair(4, a0, b0, a1, b1);
      __builtin_amdgcn_sched_barrier(0);
      mfma4x2(a2, b2, acc0, a3, b3, acc1);
      __builtin_amdgcn_sched_barrier(0);
      pc += 64;
      pb += 4 * S;
    }
    if (k + 1 < pairs) {
      load_pair(2, a2, b2, a3, b3);
      __builtin_amdgcn_sched_barrier(0);
      mfma4x2(a0, b0, acc0, a1, b1, acc1);
      mfma4x2(a2, b2, acc0, a3, b3, acc1);
      pc += 64;
      pb += 4 * S;
    } else {
      mfma4x2(a0, b0, acc0, a1, b1, acc1);
      pc += 32;
      pb += 2 * S;
    }
  }
  if (cnt & 1) {
#pragma unroll
    for (int q = 0; q < 4; ++q) a0[q] = c[pc + q];
    load4(sA + pb, b0);
    acc0 = mfma4(a0, b0, acc0);
  }
  return acc0 + acc1;
}

// Stores (whole tile) or atomically adds (half of a shared tile) a lane's 4
// results of output column t: out[16 * col + row(kk, r)].
template <typename T, typename A4>
__device__ __forceinline__ void put4(T* out, int col, const A4& acc, bool whole) {
  const int kk = (threadIdx.x & (kWave - 1)) >> 4;
  constexpr int RS = Acc4<T>::kRowStride;
  T* o = out + 16 * col + Acc4<T>::kRowMul * kk;
  if (whole) {
    o[0] = acc[0]; o[RS] = acc[1]; o[2 * RS] = acc[2]; o[3 * RS] = acc[3];
  } else {
    atomicAdd(o, acc[0]); atomicAdd(o + RS, acc[1]);
    atomicAdd(o + 2 * RS, acc[2]); atomicAdd(o + 3 * RS, acc[3]);
  }
}

// -------------------------------------------- MFMA autoconvolution (split) --
// out[k] += (sA (*) sB)[k] for all k < 16 * 16 * conv_tiles(n), where sA and
// sB are padded vectors of length n.  `out` must be exactly zero before the
// call (two commuting partial sums per element at most).  Every wave calls
// this with its wave index; returns without a barrier.
template <typename T>
__device__ void conv_mfma(const T* __restrict__ sA, const T* __restrict__ sB, T* out, int n,
                          int wave, int nwaves) {
  const int N16 = nseg16(n);
  const int NT = conv_tiles(n);
  int U = 0, maxc = 0;
  for (int j = 0; j < NT; ++j) {
    const int c = conv_cost(j, N16);
    U += c;
    maxc = c > maxc ? c : maxc;
  }
  Share sh = make_share(U, maxc, wave, nwaves);
  if (!sh.active) return;
  const int lane = threadIdx.x & (kWave - 1);
  const int t = lane & 15;
  int j = 0, base = 0;
  while (base + conv_cost(j, N16) <= sh.u0) { base += conv_cost(j, N16); ++j; }
  while (base < sh.u1 && j < NT) {
    int dlo, dhi;
    conv_drange(j, N16, dlo, dhi);
    const int cost = dhi - dlo + 1;
    const int ua = sh.u0 > base ? sh.u0 : base;
    const int ub = sh.u1 < base + cost ? sh.u1 : base + cost;
    const int K0 = 16 * j;
    const auto acc = conv_tile<T>(sA, sB, K0, dlo + (ua - base), dlo + (ub - base));
    put4(out, K0 + t, acc, (ua == base) && (ub == base + cost));
    base += cost;
    ++j;
  }
}

// ------------------------------------- symmetric MFMA autoconvolution -----
// b_k = sum_i a_i a_{k-i} counts every product a_i a_j (i != j) twice, so
//   b_k = 2 * sum_{i < k-i} a_i a_{k-i} + [k even] a_{k/2}^2.
// In the blocked form, block d of output segment K holds the products with
// i in [16d + row - 15, 16d + row], j = k - i in [16(K-d), 16(K-d) + 15]; all
// of them have i < j while d < dp(K) = ceil(K/2), none once d > dp(K).  The
// matrix cores therefore run only d < dp(K) (half the blocks): tile j needs
// d <= 8j + 7, and inside the last eight offsets (d >= 8j, the "zone")
// column t = K - 16j drops out once d >= dp(K).  What is left of each output,
// the products with 16(dp-1) + row < i < k/2 (at most 15) and the square
// term, is added by autoconv_finish.
__device__ __forceinline__ void autoconv_drange(int j, int N16, int& lo, int& hi) {
  conv_drange(j, N16, lo, hi);
  const int z = 8 * j + 7;
  hi = hi < z ? hi : z;
}
__device__ __forceinline__ int autoconv_cost(int j, int N16) {
  int lo, hi;
  autoconv_drange(j, N16, lo, hi);
  return hi >= lo ? hi - lo + 1 : 0;
}

// Zone steps d in [da, db) of tile K0 = 16j (8j <= da): column t keeps its B
// operand only while d < dp(K0 + t) = 8j + ceil(t/2).
template <typename T, typename A4>
__device__ __forceinline__ A4 autoconv_zone(const T* __restrict__ sA, int K0, int da, int db,
                                            A4 acc) {
  const int lane = threadIdx.x & (kWave - 1);
  const int t = lane & 15, kk = lane >> 4;
  const int lim = (K0 >> 1) + ((t + 1) >> 1);  // dp of this lane's column
  for (int d = da; d < db; ++d) {
    T x[4], y[4];
    const int pa = pad_addr(16 * d + t - 4 * kk);
#pragma unroll
    for (int q = 0; q < 4; ++q) x[q] = sA[pa - q];
    load4(sA + pad_addr(16 * (K0 + t - d)) + 4 * kk, y);
    if (d >= lim) {
#pragma unroll
      for (int q = 0; q < 4; ++q) y[q] = T(0);
    }
    acc = mfma4(x, y, acc);
  }
  return acc;
}

// out[k] += sum over the matrix-core share of the symmetric autoconvolution
// (see above) for all k < 16 * 16 * conv_tiles(n); sA padded, `out` exactly
// zero before the call.  Same split and determinism as conv_mfma.
template <typename T>
__device__ void autoconv_mfma(const T* __restrict__ sA, T* out, int n, int wave, int nwaves) {
  const int N16 = nseg16(n);
  const int NT = conv_tiles(n);
  int U = 0, maxc = 0;
  for (int j = 0; j < NT; ++j) {
    const int c = autoconv_cost(j, N16);
    U += c;
    maxc = c > maxc ? c : maxc;
  }
  if (U == 0) return;
  Share sh = make_share(U, maxc, wave, nwaves);
  if (!sh.active) return;
  const int lane = threadIdx.x & (kWave - 1);
  const int t = lane & 15;
  int j = 0, base = 0;
  while (base + autoconv_cost(j, N16) <= sh.u0) { base += autoconv_cost(j, N16); ++j; }
  while (base < sh.u1 && j < NT) {
    int dlo, dhi;
    autoconv_drange(j, N16, dlo, dhi);
    const int cost = dhi >= dlo ? dhi - dlo + 1 : 0;
    if (cost > 0) {
      const int ua = sh.u0 > base ? sh.u0 : base;
      const int ub = sh.u1 < base + cost ? sh.u1 : base + cost;
      const int K0 = 16 * j, z = 8 * j;
      const int da = dlo + (ua - base), db = dlo + (ub - base);
      const int dm = db < z ? db : z;
      typename Acc4<T>::type acc = {0, 0, 0, 0};
      if (da < dm) acc = conv_tile<T>(sA, sA, K0, da, dm);
      const int dz = da > z ? da : z;
      if (dz < db) acc = autoconv_zone<T>(sA, K0, dz, db, acc);
      put4(out, K0 + t, acc, (ua == base) && (ub == base + cost));
    }
    base += cost;
    ++j;
  }
}

// b_k from the matrix-core share h = out[k] of autoconv_mfma: adds the
// products 16(dp-1) + row < i < k/2 (dp = ceil(K/2), K = k/16, row = k%16;
// at most 15 of them) in ascending i, doubles, and adds a_{k/2}^2 for even k.
// The loop runs its full 15 trips with the surplus terms masked, so the reads
// of all trips are issued together; every read stays inside the padded
// vector (i >= -15, k - i <= k/2 + 15).
template <typename T>
__device__ __forceinline__ T autoconv_finish(const T* __restrict__ sA, T h, int k) {
  const int K = k >> 4, row = k & 15;
  const int i0 = 16 * (((K + 1) >> 1) - 1) + row + 1;
  const int cnt = ((k + 1) >> 1) - i0;
  T r = h;
#pragma unroll
  for (int u = 0; u < 15; ++u) {
    const T x = sA[pad_addr(i0 + u)];
    const T y = sA[pad_addr(k - i0 - u)];
    r += u < cnt ? x * y : T(0);
  }
  T b = r + r;
  if (!(k & 1)) {
    const T x = sA[pad_addr(k >> 1)];
    b += x * x;
  }
  return b;
}

// ---------------------------------------------- MFMA correlation (split) ---
// out[i] += sum_j c[i + j] a[j] for i < 256 * corr_tiles(n); c in the normal
// layout, zero for indices >= 2n-1 up to cbuf_len(n); a padded.  `out` must
// be exactly zero before the call.
template <typename T>
__device__ void corr_mfma(const T* __restrict__ c, const T* __restrict__ sA, T* out, int n,
                          int wave, int nwaves) {
  const int N16 = nseg16(n);
  const int NT = corr_tiles(n);
  const int cost = corr_cost(N16);
  const int U = NT * cost;
  Share sh = make_share(U, cost, wave, nwaves);
  if (!sh.active) return;
  const int lane = threadIdx.x & (kWave - 1);
  const int t = lane & 15;
  int j = sh.u0 / cost;
  int base = j * cost;
  while (base < sh.u1 && j < NT) {
    const int ua = sh.u0 > base ? sh.u0 : base;
    const int ub = sh.u1 < base + cost ? sh.u1 : base + cost;
    const int I0 = 16 * j;
    const auto acc = corr_tile<T>(c, sA, I0, I0 + (ua - base), I0 + (ub - base));
    put4(out, I0 + t, acc, (ua == base) && (ub == base + cost));
    base += cost;
    ++j;
  }
}

// ------------------------------------------------------------ L^p weights --
// c = (b/M)^(p-1) in the log1p-stable form exp((p-1) log1p((b-M)/M))
// (reference: src/kernels.cuh:35-44, :221-230); u^p = c (1 + d).
__device__ __forceinline__ float lp_weight(float d, float pm1) {
  return d > -1.0f ? expf(pm1 * log1pf(d)) : 0.0f;
}
__device__ __forceinline__ double lp_weight(double d, double pm1) {
  return d > -1.0 ? exp(pm1 * log1p(d)) : 0.0;
}
__device__ __forceinline__ float dexp(float x) { return expf(x); }
__device__ __forceinline__ double dexp(double x) { return exp(x); }
__device__ __forceinline__ float dsqrt(float x) { return sqrtf(x); }
__device__ __forceinline__ double dsqrt(double x) { return sqrt(x); }

}  // namespace sidon
