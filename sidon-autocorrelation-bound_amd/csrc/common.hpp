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
//  * the vector `a` lives in LDS in an overlapped segment layout: SS (20 or
//    24) elements per 16-element segment (the last four heights of the
//    previous segment, then its own sixteen, then SS - 20 unused slots), 16
//    zero segments on each side.  The four Toeplitz A-operand heights of a
//    lane are always four adjacent elements and every out-of-range index
//    reads an exact zero.  The MFMA B operand is one 16-byte read per lane
//    from 16 consecutive segments; gfx950 serves such a read 16 lanes a pass
//    over 64 four-byte banks, and measured on the device a 20-element stride
//    costs a second pass (2-way conflict) where a 24-element stride costs
//    none, so kernels take SS = 24 wherever the LDS budget allows;
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
constexpr int kSegStrideWide = 24;  // the same with 4 unused slots: no LDS bank conflicts
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

// Returns x unchanged but opaque to the optimiser, so values derived from it
// inside a loop body are computed there rather than hoisted out of the loop
// (where they would stay live, and spill, across the matrix-core loops).
__device__ __forceinline__ int opaque(int x) {
  asm volatile("" : "+v"(x));
  return x;
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
template <int SS = kSegStride>
__host__ __device__ __forceinline__ int pad_addr(int x) {
  const int y = x + kPadSeg * kSeg;
  return (y >> 4) * SS + 4 + (y & 15);
}
template <int SS = kSegStride, typename T>
__device__ __forceinline__ void pad_put(T* s, int x, T v) {
  const int p = pad_addr<SS>(x);
  s[p] = v;
  // copy at the head of the next segment (SS = 16: the layout is contiguous,
  // the "copy" is the element itself)
  if constexpr (SS > kSeg)
    if (((x + kPadSeg * kSeg) & 15) >= 12) s[p + (SS - kSeg)] = v;
}
__host__ __device__ __forceinline__ int nseg16(int n) { return (n + kSeg - 1) / kSeg; }
// floats (elements) of the padded vector buffer.  SS = 16 is the plain
// contiguous layout (pad_addr(x) = x + 260): the four leading slots of the
// overlapped layouts then shift the last segment four elements past
// (N16 + 32) * 16, so it takes four more.
template <int SS = kSegStride>
__host__ __device__ __forceinline__ int pad_len(int n) {
  return (nseg16(n) + 2 * kPadSeg) * SS + (SS == kSeg ? 4 : 0);
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
// 16 columns of a row hit disjoint banks).  Even / odd steps accumulate into
// separate registers.

__device__ __forceinline__ void load4(const float* p, float (&o)[4]) {
  const float4 v = *reinterpret_cast<const float4*>(p);
  o[0] = v.x; o[1] = v.y; o[2] = v.z; o[3] = v.w;
}
__device__ __forceinline__ void load4(const double* p, double (&o)[4]) {
  const double2 v0 = *reinterpret_cast<const double2*>(p);
  const double2 v1 = *reinterpret_cast<const double2*>(p + 2);
  o[0] = v0.x; o[1] = v0.y; o[2] = v1.x; o[3] = v1.y;
}

template <typename A4, typename T>
__device__ __forceinline__ A4 mfma4(const T (&a)[4], const T (&b)[4], A4 acc) {
#pragma unroll
  for (int q = 0; q < 4; ++q) acc = mfma16x16x4(a[q], b[q], acc);
  return acc;
}

// Operand steps kept in flight by the tile loops (registers: 8 values/step).
// Lane indices inside the tile code come from opaque(threadIdx.x), so the
// lane-dependent address terms are rebuilt per phase instead of being
// hoisted out of the step loop; that keeps a 16-wave workgroup (128
// registers per lane) within its registers at this depth.
template <typename T> struct Pipe { static constexpr int kDepth = 4; };
template <> struct Pipe<double> { static constexpr int kDepth = 2; };

// Runs `cnt` block steps of one output tile: step i multiplies the operands
// that ld(i, a, b) reads (four A and four B values per lane, one MFMA each);
// even steps accumulate into one register set, odd steps into another.  The
// operands of D steps are in flight at once: slot s is refilled for step
// i + D right after the MFMAs of step i have issued (the scheduling barriers
// keep the reads there), so each read has D - 1 steps of MFMA time to land.
// Nothing is read for steps outside [0, cnt).
template <typename T, int D, typename Ld, typename A4>
__device__ __forceinline__ void run_steps_into(int cnt, const Ld& ld, A4& acc0, A4& acc1) {
  static_assert(D % 2 == 0, "even depth keeps the accumulator parity static");
  T x[D][4], y[D][4];
  if (cnt < 2 * D) {  // short range (shares' edges, small n)
    for (int i = 0; i < cnt; ++i) {
      ld(i, x[0], y[0]);
      if (i & 1) acc1 = mfma4(x[0], y[0], acc1);
      else acc0 = mfma4(x[0], y[0], acc0);
    }
    return;
  }
#pragma unroll
  for (int s = 0; s < D; ++s) {
    ld(s, x[s], y[s]);
    __builtin_amdgcn_sched_barrier(0);
  }
  int i = 0;
  for (; i + 2 * D <= cnt; i += D) {
#pragma unroll
    for (int s = 0; s < D; ++s) {
      if (s & 1) acc1 = mfma4(x[s], y[s], acc1);
      else acc0 = mfma4(x[s], y[s], acc0);
      __builtin_amdgcn_sched_barrier(0);
      ld(i + s + D, x[s], y[s]);
      __builtin_amdgcn_sched_barrier(0);
    }
  }
  // D <= cnt - i < 2D steps remain; refill only what lies in range
#pragma unroll
  for (int s = 0; s < D; ++s) {
    if (s & 1) acc1 = mfma4(x[s], y[s], acc1);
    else acc0 = mfma4(x[s], y[s], acc0);
    if (i + s + D < cnt) {
      __builtin_amdgcn_sched_barrier(0);
      ld(i + s + D, x[s], y[s]);
      __builtin_amdgcn_sched_barrier(0);
    }
  }
  i += D;
#pragma unroll
  for (int s = 0; s < D; ++s) {
    if (i + s < cnt) {
      if (s & 1) acc1 = mfma4(x[s], y[s], acc1);
      else acc0 = mfma4(x[s], y[s], acc0);
    }
  }
}
template <typename T, int D, typename Ld>
__device__ __forceinline__ typename Acc4<T>::type run_steps(int cnt, const Ld& ld) {
  typename Acc4<T>::type acc0 = {0, 0, 0, 0}, acc1 = {0, 0, 0, 0};
  run_steps_into<T, D>(cnt, ld, acc0, acc1);
  return acc0 + acc1;
}

// Toeplitz tile: the 16x16 output block D[row][t] = (sA (*) sB)[16(K0+t) + row]
// accumulated over block offsets d in [da, db) (rows per lane: Acc4<T>).
//   A operand (lane t, kk; MFMA q): H_d[t][s] = sA[16d + t - s]
//                                   (four adjacent elements, pad_addr(x) - q)
//   B operand:                      sB[16(K0 + t - d) + s]      (column = t)
// LDS addresses are kept as integer element indices so the reads stay
// ds_read.
template <typename T, int D = Pipe<T>::kDepth, int SS = kSegStride>
__device__ __forceinline__ typename Acc4<T>::type conv_tile(const T* __restrict__ sA,
                                                           const T* __restrict__ sB, int K0,
                                                           int da, int db) {
  constexpr int S = SS;
  const int lane = opaque((int)threadIdx.x) & (kWave - 1);
  const int t = lane & 15, kk = lane >> 4;
  const int pa = pad_addr<SS>(16 * da + t - 4 * kk);
  const int pb = pad_addr<SS>(16 * (K0 + t - da)) + 4 * kk;
  return run_steps<T, D>(db - da, [&](int i, T (&x)[4], T (&y)[4]) {
#pragma unroll
    for (int q = 0; q < 4; ++q) x[q] = sA[pa + i * S - q];
    load4(sB + pb - i * S, y);
  });
}

// Hankel tile: D[row][t] = sum over e in [ea, eb) of
// sum_s c[16e + row + s] * a[16(e - I0 - t) + s]  = G[16(I0 + t) + row].
//   A operand: C_e[t][s] = c[16e + t + s]       (four consecutive elements)
//   B operand: sA[16(e - I0 - t) + s]           (column = t)
template <typename T, int D = Pipe<T>::kDepth, int SS = kSegStride>
__device__ __forceinline__ typename Acc4<T>::type corr_tile(const T* __restrict__ c,
                                                           const T* __restrict__ sA, int I0,
                                                           int ea, int eb) {
  constexpr int S = SS;
  const int lane = opaque((int)threadIdx.x) & (kWave - 1);
  const int t = lane & 15, kk = lane >> 4;
  const int pc = 16 * ea + t + 4 * kk;
  const int pb = pad_addr<SS>(16 * (ea - I0 - t)) + 4 * kk;
  return run_steps<T, D>(eb - ea, [&](int i, T (&x)[4], T (&y)[4]) {
#pragma unroll
    for (int q = 0; q < 4; ++q) x[q] = c[pc + 16 * i + q];
    load4(sA + pb + i * S, y);
  });
}

// ------------------------------------------------ sparse weight vectors --
// At high p most weights c_k = (b_k/M)^(p-1) are exact zeros.  The weight
// pass records one bit per 16-element segment g of c (bit set: some c_k,
// 16g <= k < 16g+16, is non-zero); Hankel step e reads c[16e .. 16e+30], so
// it is needed only when segment e or e+1 is flagged.  Steps that are not
// needed would add exact zeros, so skipping them leaves the sums unchanged
// except for their grouping into runs.
__host__ __device__ __forceinline__ int segmask_words(int n) {
  return (cbuf_len(n) / kSeg + 31) / 32 + 1;
}
// Marks the segments of 64 consecutive weights c[k0 .. k0+63] (k0 a multiple
// of 64, one per lane of the calling wave; lanes past the end inactive).
__device__ __forceinline__ void mark_segments(unsigned* segmask, int k0, bool nonzero) {
  const unsigned long long bal = __ballot(nonzero);
  unsigned bits = 0;
#pragma unroll
  for (int q = 0; q < 4; ++q) bits |= ((bal >> (16 * q)) & 0xFFFFull) ? (1u << q) : 0u;
  const int g = k0 >> 4;  // multiple of 4: the four bits share a word
  if (bits && (threadIdx.x & (kWave - 1)) == __builtin_ctzll(__ballot(true)))
    atomicOr(segmask + (g >> 5), bits << (g & 31));
}
// Needed-step bits of steps 32w .. 32w+31 (wave-uniform).
__device__ __forceinline__ unsigned needed_word(const unsigned* segmask, int w) {
  const unsigned s0 = __builtin_amdgcn_readfirstlane(segmask[w]);
  const unsigned s1 = __builtin_amdgcn_readfirstlane(segmask[w + 1]);
  return s0 | (s0 >> 1) | (s1 << 31);
}
// First step in [e, eb) whose needed bit equals `want` (eb if none).
__device__ __forceinline__ int next_step(const unsigned* segmask, int e, int eb, bool want) {
  while (e < eb) {
    const int w = e >> 5;
    unsigned bits = needed_word(segmask, w);
    if (!want) bits = ~bits;
    bits >>= (e & 31);
    if (bits) {
      const int f = e + __builtin_ctz(bits);
      return f < eb ? f : eb;
    }
    e = (w + 1) << 5;
  }
  return eb;
}

// Hankel tile over the needed steps of [ea, eb) only (see corr_tile).
template <typename T, int D = Pipe<T>::kDepth, int SS = kSegStride>
__device__ __forceinline__ typename Acc4<T>::type corr_tile_sparse(const T* __restrict__ c,
                                                                  const T* __restrict__ sA,
                                                                  int I0, int ea, int eb,
                                                                  const unsigned* segmask,
                                                                  int& steps) {
  constexpr int S = SS;
  const int lane = opaque((int)threadIdx.x) & (kWave - 1);
  const int t = lane & 15, kk = lane >> 4;
  typename Acc4<T>::type acc0 = {0, 0, 0, 0}, acc1 = {0, 0, 0, 0};
  int e = next_step(segmask, ea, eb, true);
  while (e < eb) {
    const int e1 = next_step(segmask, e + 1, eb, false);
    const int pc = 16 * e + t + 4 * kk;
    const int pb = pad_addr<SS>(16 * (e - I0 - t)) + 4 * kk;
    run_steps_into<T, D>(e1 - e, [&](int i, T (&x)[4], T (&y)[4]) {
#pragma unroll
      for (int q = 0; q < 4; ++q) x[q] = c[pc + 16 * i + q];
      load4(sA + pb + i * S, y);
    }, acc0, acc1);
    steps += e1 - e;
    e = next_step(segmask, e1, eb, true);
  }
  return acc0 + acc1;
}

// Stores (whole tile) or atomically adds (half of a shared tile) a lane's 4
// results of output column t: out[16 * col + row(kk, r)].
template <typename T, typename A4>
__device__ __forceinline__ void put4(T* out, int col, const A4& acc, bool whole) {
  const int kk = (opaque((int)threadIdx.x) & (kWave - 1)) >> 4;
  constexpr int RS = Acc4<T>::kRowStride;
  T* o = out + 16 * col + Acc4<T>::kRowMul * kk;
  if (whole) {
    o[0] = acc[0]; o[RS] = acc[1]; o[2 * RS] = acc[2]; o[3 * RS] = acc[3];
  } else {
    atomicAdd(o, acc[0]); atomicAdd(o + RS, acc[1]);
    atomicAdd(o + 2 * RS, acc[2]); atomicAdd(o + 3 * RS, acc[3]);
  }
}

// Device-scope (agent) accesses: coherent between workgroups on different
// XCDs, whose L2 caches are separate (the two halves of a split candidate).
template <typename T>
__device__ __forceinline__ void coherent_store(T* p, T v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <typename T>
__device__ __forceinline__ T coherent_load(const T* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// Workgroup-scope accesses to a buffer in global memory that one workgroup
// owns (a candidate's correlation output row): coherent among the waves of
// the workgroup, which share one CU and one L2, so the lines stay in L2
// instead of being written through to memory as device-scope ones are.
template <typename T>
__device__ __forceinline__ void wg_store(T* p, T v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
template <typename T>
__device__ __forceinline__ T wg_load(const T* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
template <typename T>
__device__ __forceinline__ void wg_add(T* p, T v) {
  [[clang::atomic(no_remote_memory, no_fine_grained_memory, ignore_denormal_mode)]] {
    __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
  }
}
// put4 into a candidate's row in global memory (outputs i < n only; rows past
// n belong to the next candidate): whole tiles by stores, halves of a shared
// tile by atomic adds onto an exact zero, both at workgroup scope.  Readers in
// the same workgroup wait for these (s_waitcnt vmcnt(0) before their barrier,
// whose workgroup-scope fences order them) and read with wg_load.
template <typename T, typename A4>
__device__ __forceinline__ void put4_global(T* out, int col, const A4& acc, bool whole, int n) {
  const int kk = (opaque((int)threadIdx.x) & (kWave - 1)) >> 4;
  constexpr int RS = Acc4<T>::kRowStride;
  const int i0 = 16 * col + Acc4<T>::kRowMul * kk;
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const int i = i0 + RS * r;
    if (i < n) {
      if (whole) wg_store(out + i, acc[r]);
      else wg_add(out + i, acc[r]);
    }
  }
}
// The two-row form: every item stores its sums (outputs i < n), the first
// item of a tile into `out`, the second into the second row, and the reader
// adds the two rows.  Float atomic adds in global memory are performed past
// the L2 (each one a 32-byte write to the memory side), stores stay in it;
// fl(x + y) is the value 0 + x + y of two atomic adds in either order.
template <typename T, typename A4>
__device__ __forceinline__ void put4_rows(T* out, int col, const A4& acc, int n) {
  const int kk = (opaque((int)threadIdx.x) & (kWave - 1)) >> 4;
  constexpr int RS = Acc4<T>::kRowStride;
  const int i0 = 16 * col + Acc4<T>::kRowMul * kk;
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const int i = i0 + RS * r;
    if (i < n) wg_store(out + i, acc[r]);
  }
}
// This wave's global stores and atomics have completed (before a barrier
// that publishes them to the other waves of the workgroup).
__device__ __forceinline__ void wait_global_writes() {
  asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
}

// -------------------------------------------- MFMA autoconvolution (split) --
// out[k] += (sA (*) sB)[k] for all k < 16 * 16 * conv_tiles(n), where sA and
// sB are padded vectors of length n.  `out` must be exactly zero before the
// call (two commuting partial sums per element at most).  Every wave calls
// this with its wave index; returns without a barrier.
template <typename T, int SS = kSegStride>
__device__ __forceinline__ void conv_mfma(const T* __restrict__ sA, const T* __restrict__ sB, T* out, int n,
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
  const int lane = opaque((int)threadIdx.x) & (kWave - 1);
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
    const auto acc =
        conv_tile<T, Pipe<T>::kDepth, SS>(sA, sB, K0, dlo + (ua - base), dlo + (ub - base));
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

// Toeplitz steps d in [da, db) of autoconvolution tile K0 = 16j.  In the
// zone d >= 8j, column t keeps its B operand only while d < dp(K0 + t) =
// 8j + ceil(t/2); a lane past that reads four zeros from the left padding
// instead.  The steps before the zone (all columns live) run as one
// pipelined sequence with constant-stride addresses, the zone steps (at most
// eight) as a second one with the per-lane select; the second continues the
// first's even / odd accumulator assignment, so the sums are the ones of a
// single sequence over [da, db).
template <typename T, int D = Pipe<T>::kDepth, int SS = kSegStride>
__device__ __forceinline__ typename Acc4<T>::type autoconv_tile(const T* __restrict__ sA, int K0,
                                                               int da, int db) {
  constexpr int S = SS;
  const int lane = opaque((int)threadIdx.x) & (kWave - 1);
  const int t = lane & 15, kk = lane >> 4;
  const int pa = pad_addr<SS>(16 * da + t - 4 * kk);
  const int pb = pad_addr<SS>(16 * (K0 + t - da)) + 4 * kk;
  const int dz = K0 >> 1;                               // first zone offset, 8j
  const int za = da > dz ? da : (dz < db ? dz : db);    // zone steps are [za, db)
  const int na = za - da;                               // steps before the zone
  typename Acc4<T>::type acc0 = {0, 0, 0, 0}, acc1 = {0, 0, 0, 0};
  if (na > 0)
    run_steps_into<T, D>(na, [&](int i, T (&x)[4], T (&y)[4]) {
#pragma unroll
      for (int q = 0; q < 4; ++q) x[q] = sA[pa + i * S - q];
      load4(sA + pb - i * S, y);
    }, acc0, acc1);
  if (db > za) {
    const int qa = pa + na * S, qb = pb - na * S;
    const int live = dz + ((t + 1) >> 1) - za;          // zone steps i < live use this column
    const int zero = pad_addr<SS>(-kPadSeg * kSeg);     // first padding segment: always zero
    auto ld = [&](int i, T (&x)[4], T (&y)[4]) {
#pragma unroll
      for (int q = 0; q < 4; ++q) x[q] = sA[qa + i * S - q];
      load4(sA + (i < live ? qb - i * S : zero), y);
    };
    if (na & 1) run_steps_into<T, D>(db - za, ld, acc1, acc0);
    else run_steps_into<T, D>(db - za, ld, acc0, acc1);
  }
  return acc0 + acc1;
}

// ------------------------------------------ dynamic, deterministic split --
// The waves of a workgroup take work items in a fixed order from an LDS
// counter (zero when the phase starts), so waves that the matrix-core
// arbiter serves faster simply take more items and all finish together.
// One wave computes each item in a fixed order; a tile is at most two items
// whose partial sums are added onto an exactly-zero output (0 + x + y ==
// 0 + y + x), so the result has the same bits whichever wave takes which
// item.  A wave asks for its next ticket when it has finished an item: a
// ticket taken earlier (when the item starts) is held by a wave that may be
// the slowest of its SIMD, and at the end of a phase the other waves then
// wait for that item (measured: 27 k cycles of spread at the end of the
// correlation at n = 8400, against 5 k without).
__device__ __forceinline__ int take_ticket(int* ctr) {
  int it = 0;
  if ((threadIdx.x & (kWave - 1)) == 0) it = atomicAdd(ctr, 1);
  return it;  // lane 0 holds the ticket
}
__device__ __forceinline__ int ticket_value(int it) {
  return __builtin_amdgcn_readfirstlane(it);  // every lane is active: lane 0's ticket
}

// Work items of a phase: (tile j, block steps [u0, u1) of the tile), packed
// j << 22 | u0 << 11 | u1.  With `split`, tile j (cost c_j steps, rank r_j
// in decreasing cost, ties by index) is cut into a head [0, c_j - s_j) and
// a tail [c_j - s_j, c_j), s_j = c_j (1 + 24 r_j / (NT - 1)) / 50: 2 % of
// the largest tile up to half of the smallest, none below 8 steps; without
// it every tile is one item.  The list holds the heads (or whole tiles) in
// decreasing size, then the tails in decreasing size (2 NT slots; empty
// tails sort last and are skipped), so the items handed out last are the
// smallest.  Equal-cost tiles (the correlation) need the split to end
// together; the autoconvolution's tiles already range from 8 steps to N16/2
// and run fastest whole.
constexpr int kItemBits = 11;  // steps per tile < 2048, tiles < 1024
__device__ __forceinline__ unsigned pack_item(int j, int u0, int u1) {
  return ((unsigned)j << (2 * kItemBits)) | ((unsigned)u0 << kItemBits) | (unsigned)u1;
}
__device__ __forceinline__ void unpack_item(unsigned v, int& j, int& u0, int& u1) {
  constexpr unsigned m = (1u << kItemBits) - 1;
  j = (int)(v >> (2 * kItemBits));
  u0 = (int)((v >> kItemBits) & m);
  u1 = (int)(v & m);
}

// Builds the 2 NT items of a phase whose tile j costs cost(j) steps.  Every
// thread of the workgroup calls this (it holds barriers); `scratch` is 3 NT
// ints of LDS that nothing else uses meanwhile; ends with a barrier, so
// `items` is published and `scratch` free on return.
template <typename Cost>
__device__ void build_items(int NT, const Cost& cost, bool split, unsigned* items, int* scratch,
                            int tid, int nthreads) {
  int* c = scratch;
  int* h = scratch + NT;
  int* tl = scratch + 2 * NT;
  for (int j = tid; j < NT; j += nthreads) c[j] = cost(j);
  __syncthreads();
  for (int j = tid; j < NT; j += nthreads) {
    const int cj = c[j];
    int r = 0;
    for (int i = 0; i < NT; ++i) r += (c[i] > cj) || (c[i] == cj && i < j);
    int sj = 0;
    if (split && cj >= 8) {
      sj = NT > 1 ? (int)((long long)cj * (NT - 1 + 24 * r) / (50LL * (NT - 1))) : cj / 50;
      sj = sj < 1 ? 1 : (sj > cj - 1 ? cj - 1 : sj);
    }
    h[j] = cj - sj;
    tl[j] = sj;
  }
  __syncthreads();
  for (int j = tid; j < NT; j += nthreads) {
    int rh = 0, rt = 0;
    for (int i = 0; i < NT; ++i) {
      rh += (h[i] > h[j]) || (h[i] == h[j] && i < j);
      rt += (tl[i] > tl[j]) || (tl[i] == tl[j] && i < j);
    }
    items[rh] = pack_item(j, 0, h[j]);
    items[NT + rt] = pack_item(j, h[j], c[j]);
  }
  __syncthreads();
}

// Work items of the symmetric autoconvolution (2 conv_tiles(n) entries).
__device__ __forceinline__ void autoconv_items(int n, unsigned* items, int* scratch, int tid,
                                               int nthreads) {
  const int N16 = nseg16(n);
  build_items(conv_tiles(n), [N16](int j) { return autoconv_cost(j, N16); }, false, items, scratch,
              tid, nthreads);
}

// b_k from the matrix-core share h = out[k] of autoconv_mfma: adds the
// products 16(dp-1) + row < i < k/2 (dp = ceil(K/2), K = k/16, row = k%16;
// at most 15 of them) in ascending i, doubles, and adds a_{k/2}^2 for even k.
// The loop runs its full 15 trips with the surplus terms masked, so the reads
// of all trips are issued together; every read stays inside the padded
// vector (i >= -15, k - i <= k/2 + 15).
template <typename T, int SS = kSegStride>
__device__ __forceinline__ T autoconv_finish(const T* __restrict__ sA, T h, int k) {
  const int K = k >> 4, row = k & 15;
  const int dp = (K + 1) >> 1;
  const int i0 = 16 * (dp - 1) + row + 1;
  const int cnt = ((k + 1) >> 1) - i0;
  // a_{i0+u}: from pad_addr(i0), one segment boundary at most (skip the SS - 16
  // copy / padding slots there); a_{k-i0-u} = a_{16(K-dp)+15-u}: one segment,
  // descending
  const int xs = pad_addr<SS>(i0), r1 = (i0 + kPadSeg * kSeg) & 15;
  const int ys = pad_addr<SS>(16 * (K - dp) + 15);
  T r = h;
#pragma unroll
  for (int u = 0; u < 15; ++u) {
    const T x = sA[xs + u + ((r1 + u) >> 4) * (SS - kSeg)];
    const T y = sA[ys - u];
    r += u < cnt ? x * y : T(0);
  }
  T b = r + r;
  if (!(k & 1)) {
    const T x = sA[pad_addr<SS>(k >> 1)];
    b += x * x;
  }
  return b;
}

// autoconv_finish for the four outputs of a lane in the fp32 accumulator
// layout (rows 4 kk + r of column K): output r adds a_{i0 + r + u} a_{16(K-dp)
// + 15 - u} for u < cnt(r), with i0 = 16(dp-1) + 4kk + 1 the first index of
// output 0, so the eighteen heights a_{i0 .. i0+17} and the fifteen a_{16(K-dp)
// + 15 - u} are read once for all four outputs (the same elements, the same
// products in the same order as autoconv_finish).
template <typename T, int SS = kSegStride>
__device__ __forceinline__ void autoconv_finish_lane(const T* __restrict__ sA,
                                                     typename Acc4<T>::type& acc, int K, int kk) {
  static_assert(Acc4<T>::kRowMul == 4 && Acc4<T>::kRowStride == 1, "fp32 accumulator layout");
  const int dp = (K + 1) >> 1;
  const int i0 = 16 * (dp - 1) + 4 * kk + 1;
  const int xs = pad_addr<SS>(i0), r1 = (i0 + kPadSeg * kSeg) & 15;
  const int ys = pad_addr<SS>(16 * (K - dp) + 15);
  T xv[18], yv[15];
#pragma unroll
  for (int v = 0; v < 18; ++v) xv[v] = sA[xs + v + ((r1 + v) >> 4) * (SS - kSeg)];
#pragma unroll
  for (int u = 0; u < 15; ++u) yv[u] = sA[ys - u];
  // output r takes u < cnt(r) = ((k + 1) >> 1) - (i0 + r) terms: cnt(0) = cnt(1) = C,
  // cnt(2) = cnt(3) = C - 1 (k = 16K + 4kk + r, 16K + 4kk even); the surplus terms
  // multiply a zeroed y (heights are >= 0, so they add exactly +0 as before)
  const int C = ((16 * K + 4 * kk) >> 1) - i0;
  T ya[15], yb[15];
#pragma unroll
  for (int u = 0; u < 15; ++u) {
    ya[u] = u < C ? yv[u] : T(0);
    yb[u] = u < C - 1 ? yv[u] : T(0);
  }
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const int k = 16 * K + 4 * kk + r;
    T s = acc[r];
#pragma unroll
    for (int u = 0; u < 15; ++u) s += xv[r + u] * (r < 2 ? ya[u] : yb[u]);
    T b = s + s;
    if (!(k & 1)) {
      const T x = sA[pad_addr<SS>(k >> 1)];
      b += x * x;
    }
    acc[r] = b;
  }
}

// Atomic max of a non-negative value into an LDS slot holding one (IEEE bits
// of non-negative floats order like unsigned integers).
__device__ __forceinline__ void atomic_max_nonneg(float* p, float v) {
  atomicMax(reinterpret_cast<unsigned*>(p), __float_as_uint(v));
}
__device__ __forceinline__ void atomic_max_nonneg(double* p, double v) {
  atomicMax(reinterpret_cast<unsigned long long*>(p),
            static_cast<unsigned long long>(__double_as_longlong(v)));
}

// out[k] = the matrix-core share of the symmetric autoconvolution (see
// above) for all k < 16 * 16 * conv_tiles(n), item by item from the list of
// autoconv_items.  The autoconvolution's items are whole tiles, each stored
// once (plain stores), so `out` needs no clearing between calls.  With
// `bmax` (an LDS slot that holds 0 when the phase starts) each tile's wave
// also finishes its outputs (autoconv_finish: out[k] = b_k) and raises
// *bmax to the tile's largest b_k, so the finishing work overlaps the other
// waves' matrix work.  Every wave of the workgroup calls this; returns
// without a barrier.
// With `half` >= 0 the workgroup is one of a pair sharing the candidate: it
// computes only the items it owns (conv_item_owner) and passes the others.
__device__ __forceinline__ int conv_item_owner(int rank) {
  // items are ranked by decreasing cost: dealt 0 1 1 0 0 1 1 0 ... the two
  // halves' totals differ by at most the largest tile
  return ((rank & 3) == 1 || (rank & 3) == 2) ? 1 : 0;
}
__device__ __forceinline__ int corr_tile_owner(int j) { return j & 1; }

template <typename T, int D = Pipe<T>::kDepth, int SS = kSegStride>
__device__ __forceinline__ void autoconv_mfma(const T* __restrict__ sA, T* out, int n, int* ctr,
                              const unsigned* items, unsigned* steps = nullptr,
                              T* bmax = nullptr, int half = -1) {
  const int N16 = nseg16(n);
  const int NT = conv_tiles(n);
  const int t = opaque((int)threadIdx.x) & 15;
  int it = ticket_value(take_ticket(ctr));
  int done = 0;  // block steps this wave ran
  while (it < 2 * NT) {
    int j, u0, u1;
    unpack_item(__builtin_amdgcn_readfirstlane(items[it]), j, u0, u1);  // wave-uniform
    const bool mine = half < 0 || conv_item_owner(it) == half;
    if (mine) done += u1 - u0;
    if (mine && (it < NT || u0 < u1)) {  // heads always store, so every tile is written
      int dlo, dhi;
      autoconv_drange(j, N16, dlo, dhi);
      const int cost = dhi >= dlo ? dhi + 1 - dlo : 0;
      const int K0 = 16 * j;
      auto acc = autoconv_tile<T, D, SS>(sA, K0, dlo + u0, dlo + u1);
      if (bmax) {  // whole tiles: this wave owns the outputs, so it finishes them
        const int kk = (opaque((int)threadIdx.x) & (kWave - 1)) >> 4;
        T m = T(0);
        if constexpr (Acc4<T>::kRowStride == 1) {
          autoconv_finish_lane<T, SS>(sA, acc, K0 + t, kk);
#pragma unroll
          for (int r = 0; r < 4; ++r) m = acc[r] > m ? acc[r] : m;
        } else {
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const int k = 16 * (K0 + t) + Acc4<T>::kRowMul * kk + Acc4<T>::kRowStride * r;
            acc[r] = autoconv_finish<T, SS>(sA, acc[r], k);
            m = acc[r] > m ? acc[r] : m;
          }
        }
        m = wave_max(m);
        if ((threadIdx.x & (kWave - 1)) == 0) atomic_max_nonneg(bmax, m);
      }
      put4(out, K0 + t, acc, u0 == 0 && u1 == cost);
    }
    it = ticket_value(take_ticket(ctr));
  }
  if (steps && done && (threadIdx.x & (kWave - 1)) == 0) atomicAdd(steps, (unsigned)done);
}


// ------------------------------------------------------ MFMA correlation --
// Work items of the correlation (2 corr_tiles(n) entries): the tiles are all
// equally long, so heads and tails ramp with the tile index.
__device__ __forceinline__ void corr_items(int n, unsigned* items, int* scratch, int tid,
                                           int nthreads) {
  const int cost = corr_cost(nseg16(n));
  build_items(corr_tiles(n), [cost](int) { return cost; }, true, items, scratch, tid, nthreads);
}

// out[i] += sum_j c[i + j] a[j] for i < 256 * corr_tiles(n), item by item
// from the list of corr_items; c in the normal layout, zero for indices
// >= 2n-1 up to cbuf_len(n); a padded.  `out` must be exactly zero before
// the call.  With `segmask` (the weight segments' non-zero bits) steps that
// read only zero weights are skipped.
// With GOUT, `out` is the candidate's row in global memory (put4_global);
// with `out2` as well, the two-row form (put4_rows): neither row needs to be
// zero, and G_i is out[i] + out2[i] once out2 is zero where no tile's second
// item writes.
template <typename T, int D = Pipe<T>::kDepth, int SS = kSegStride, bool GOUT = false>
__device__ __forceinline__ void corr_mfma(const T* __restrict__ c, const T* __restrict__ sA, T* out, int n,
                          int* ctr, const unsigned* items, const unsigned* segmask,
                          unsigned* steps = nullptr, int half = -1, T* out2 = nullptr) {
  const int N16 = nseg16(n);
  const int NT = corr_tiles(n);
  const int cost = corr_cost(N16);
  const int t = opaque((int)threadIdx.x) & 15;
  int it = ticket_value(take_ticket(ctr));
  int done = 0;  // block steps this wave ran
  while (it < 2 * NT) {
    int j, u0, u1;
    unpack_item(__builtin_amdgcn_readfirstlane(items[it]), j, u0, u1);  // wave-uniform
    if (u0 < u1 && (half < 0 || corr_tile_owner(j) == half)) {
      const int I0 = 16 * j;
      typename Acc4<T>::type acc;
      if (segmask) {
        acc = corr_tile_sparse<T, D, SS>(c, sA, I0, I0 + u0, I0 + u1, segmask, done);
      } else {
        acc = corr_tile<T, D, SS>(c, sA, I0, I0 + u0, I0 + u1);
        done += u1 - u0;
      }
      if constexpr (GOUT) {
        if (out2) put4_rows(u0 == 0 ? out : out2, I0 + t, acc, n);
        else put4_global(out, I0 + t, acc, u0 == 0 && u1 == cost, n);
      } else {
        put4(out, I0 + t, acc, u0 == 0 && u1 == cost);
      }
    }
    it = ticket_value(take_ticket(ctr));
  }
  if (steps && done && (threadIdx.x & (kWave - 1)) == 0) atomicAdd(steps, (unsigned)done);
}

// ------------------------------------------------------------ L^p weights --
// c = (b/M)^(p-1) in the log1p-stable form exp((p-1) log1p((b-M)/M))
// (reference: src/kernels.cuh:35-44, :221-230); u^p = c (1 + d).
__device__ __forceinline__ float lp_weight(float d, float pm1) {
  return d > -1.0f ? expf(pm1 * log1pf(d)) : 0.0f;
}
// Search-kernel form: c = 2^((p-1) log2(1+d)) on the hardware base-2
// logarithm and exponential (v_log_f32, v_exp_f32; the reference's fp32
// search uses a fast exponential at this point, src/kernels.cuh:43).
// log2(1+d) stays accurate for small |d|, so the weights of near-maximal b_k
// keep their digits: u = fl(1+d) is rounded, but u - 1 is exact (Sterbenz),
// so e = d - (u - 1) is the rounding error and log2(1+d) = log2(u) +
// log2(1 + e/u) ~ log2(u) + e log2(e) (for -1 < d <= 0, |e| <= 2^-25 and the
// dropped 1/u and second-order terms stay below one ulp).  Measured on
// gfx950 against fp64 log1p over d in (-1, -2^-40]: relative error of
// log2(1+d) <= 3 * 2^-24 (the libm log1pf path: <= 2 * 2^-24, at five
// times the instructions).  Weights below 2^-126 are flushed to zero (as a
// fast exponential with denormals flushed does), which makes most weights
// exact zeros at high p.
__device__ __forceinline__ float lp_weight_fast(float d, float pm1) {
  if (!(d > -1.0f)) return 0.0f;
  const float u = 1.0f + d;
  const float l2 = __builtin_amdgcn_logf(u) + (d - (u - 1.0f)) * 1.44269504088896341f;
  const float y = pm1 * l2;
  return y >= -126.0f ? __builtin_amdgcn_exp2f(y) : 0.0f;  // flushed like a fast exp: exact zeros
}
__device__ __forceinline__ double lp_weight_fast(double d, double pm1) {
  return d > -1.0 ? exp(pm1 * log1p(d)) : 0.0;
}
__device__ __forceinline__ double lp_weight(double d, double pm1) {
  return d > -1.0 ? exp(pm1 * log1p(d)) : 0.0;
}
__device__ __forceinline__ float dexp(float x) { return expf(x); }
__device__ __forceinline__ double dexp(double x) { return exp(x); }
__device__ __forceinline__ float dsqrt(float x) { return sqrtf(x); }
__device__ __forceinline__ double dsqrt(double x) { return sqrt(x); }
// a / b for the search update: fp32 through the hardware reciprocal (about one
// ulp, for normal b), fp64 the correctly rounded quotient
__device__ __forceinline__ float dquot(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }
__device__ __forceinline__ double dquot(double a, double b) { return a / b; }

}  // namespace sidon
