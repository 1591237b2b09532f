// local_search.hip -- fused L^p-annealed projected Adam / mirror descent and
// the gradient of the smoothed objective, on gfx950 matrix cores.
//
// Reference: src/kernels.cuh:160-289 (kLocalSearch), launch src/main.cu:152-159.
// One workgroup owns one candidate for all `iters` steps; a (padded), the
// autoconvolution / weight vector and the correlation output live in LDS
// (about 140 KB at n = 8400), Adam moments and the new heights live in
// registers (each thread owns heights i = tid + BD*q), so a step moves no HBM
// bytes.  Per step:
//   conv (MFMA) -> barrier -> M = max b -> weights c, R -> corr (MFMA) ->
//   barrier -> Adam/mirror update -> S -> a = x/S, clear buffers -> barrier.
#include <cmath>
#include <cstdio>
#include <mutex>

#include "common.hpp"
#include "runtime.hpp"

namespace sidon {

// Matrix block steps (4 MFMAs of 16x16x4 each) run by the local-search and
// gradient kernels since the last sidon_ls_block_steps(reset) call: the work
// the kernels actually executed, for roofline accounting (correlation steps
// that would read only zero weights are skipped and not counted).
__device__ unsigned long long g_block_steps = 0ull;
__device__ unsigned long long g_block_steps_taken = 0ull;

// Reads the counter (and with `reset` zeroes it) in one device atomic, so no
// step a running launch adds in between is lost.
__global__ void k_take_block_steps(int reset) {
  g_block_steps_taken = reset ? atomicExch(&g_block_steps, 0ull) : atomicAdd(&g_block_steps, 0ull);
}

// ------------------------------------------------------- phase clocks --
// Profiling builds only (SIDON_PHASE_CLOCKS, the library's `clocks` make
// target; the product library compiles none of this): shader-clock cycles of
// each phase of a step, from barrier to barrier as wave 0 sees them, summed
// over steps and workgroups, and the wave-cycles the waves spend at the end
// of each product phase between finishing their last work item and leaving
// the barrier.  Slots: [0] autoconvolution + finish + max, [1] weights + R,
// [2] correlation, [3] update + S + renormalisation, [4] / [5] idle
// wave-cycles at the end of the autoconvolution / correlation, [6] steps,
// [7] candidates.
#ifdef SIDON_PHASE_CLOCKS
constexpr int kPhaseSlots = 8;
__device__ unsigned long long g_phase_clk[kPhaseSlots];
__device__ unsigned long long g_phase_clk_taken[kPhaseSlots];
__global__ void k_take_phase_clocks(int reset) {
  for (int i = 0; i < kPhaseSlots; ++i)
    g_phase_clk_taken[i] = reset ? atomicExch(&g_phase_clk[i], 0ull)
                                 : atomicAdd(&g_phase_clk[i], 0ull);
}
struct PhaseClock {
  unsigned long long t = 0, acc[4] = {0, 0, 0, 0};
  unsigned long long* idle;  // LDS: [0] autoconvolution end, [1] correlation end
  unsigned long long done = 0;
  __device__ void start(unsigned long long* lds_idle) {
    idle = lds_idle;
    if (threadIdx.x == 0) { idle[0] = 0; idle[1] = 0; }
    t = clock64();
  }
  __device__ void mark(int ph) {
    const unsigned long long now = clock64();
    acc[ph] += now - t;
    t = now;
  }
  __device__ void item_end() { done = clock64(); }  // this wave's last item is done
  __device__ void after_barrier(int which) {       // wave-cycles spent waiting
    if ((threadIdx.x & (kWave - 1)) == 0) atomicAdd(&idle[which], clock64() - done);
  }
  __device__ void finish(int steps) {
    __syncthreads();
    if (threadIdx.x == 0) {
      for (int i = 0; i < 4; ++i) atomicAdd(&g_phase_clk[i], acc[i]);
      atomicAdd(&g_phase_clk[4], idle[0]);
      atomicAdd(&g_phase_clk[5], idle[1]);
      atomicAdd(&g_phase_clk[6], (unsigned long long)steps);
      atomicAdd(&g_phase_clk[7], 1ull);
    }
  }
};
#else
struct PhaseClock {
  __device__ void start(unsigned long long*) {}
  __device__ void mark(int) {}
  __device__ void item_end() {}
  __device__ void after_barrier(int) {}
  __device__ void finish(int) {}
};
#endif

template <typename T>
struct LsArgs {
  T* pop;
  long long stride;     // elements between candidates
  const int* ns;        // per-candidate length (nullptr: uniform n)
  int n;                // uniform length
  int P;
  T* mom;
  T* vel;
  const int* skip;
  const int* mode;
  int modeDefault;
  const int* order;     // processing order (nullptr: identity)
  int* queue;           // work-queue counter (nullptr: one candidate per block)
  int iters;
  T pStart, pEnd, lrStart, lrEnd, beta1, beta2;
  // persistent launches: candidates that may be split over two workgroups
  // (0: none; 1: the longest, see split_count) and the pair workspace
  int split;
  int* pair_flags;      // 16 ints per candidate: [0], [1] each half's progress
  int* pair_err;        // set when a half stopped waiting for its partner
  int* split_out;       // the number of candidates split (written by workgroup 0)
  T* xchg;              // per candidate xstride elements: the b_k of each half's tiles
  long long xstride;
  // lengths whose correlation output does not fit the LDS beside a and the
  // weights: each candidate's G in global memory, gstride elements a row
  T* gscr;
  long long gstride;
  // a second set of rows (same stride): the two-row form, G_i = row + row2
  // (nullptr: one row, the split tiles' halves added by atomic adds)
  T* gscr2;
};

// ------------------------------------------------------- paired halves --
// A candidate of a persistent launch may be shared by two workgroups (halves
// 0 and 1), each holding all of a in LDS: each computes the autoconvolution
// tiles it owns (conv_item_owner), the halves exchange those b_k through
// global memory, both form all weights and R (the same arithmetic in the
// same order, so the same bits), each computes the correlation tiles it owns
// (corr_tile_owner) and updates their heights, and the halves exchange the
// new heights through the candidate's row.  S is summed over all heights in
// the one-workgroup order, so a paired candidate ends bit-identical to the
// same candidate run by one workgroup.  The two workgroups may sit on
// different XCDs, whose L2 caches are not coherent with each other, so every
// exchanged value is written and read with device-coherent (agent-scope
// relaxed atomic, sc1) accesses: an exchange is those stores, a wait for
// their completion in every wave, a barrier, a flag raised to a step-numbered
// value (coherent store), and on the other side a poll of the flag (coherent
// load), a barrier and coherent loads of the data.  No cache is flushed or
// invalidated, so the other workgroups of the XCD are not disturbed.  A wait
// gives up after kPairWaitTicks of the 100 MHz wall clock (a partner that
// never runs), counts that in *pair_err and stops waiting, so every wave
// reaches the end.
constexpr unsigned long long kPairWaitTicks = 2000000000ull;  // 20 s
constexpr int kPairFinal = 1 << 30;

__device__ __forceinline__ void pair_release(int* flag, int value) {
  asm volatile("s_waitcnt vmcnt(0)" ::: "memory");  // this wave's coherent stores are done
  __syncthreads();
  if (threadIdx.x == 0) coherent_store(flag, value);
}
// Returns when the partner's flag has reached `value` (its coherent stores
// before that are then visible to coherent loads); `broken` (LDS) turns on
// after a timeout and makes later waits return at once.
__device__ __forceinline__ void pair_acquire(const int* flag, int value, int* err, int* broken) {
  if (threadIdx.x == 0 && !*broken) {
    const unsigned long long t0 = wall_clock64();
    while (coherent_load(flag) < value) {
      if (wall_clock64() - t0 > kPairWaitTicks) {
        *broken = 1;
        atomicAdd(err, 1);
        break;
      }
      __builtin_amdgcn_s_sleep(1);
    }
  }
  __syncthreads();
}

// Bytes of dynamic LDS for one candidate of length n (GG: the correlation
// output G in global memory, not in LDS).
template <typename T, int SS = kSegStride, bool GG = false>
__host__ __device__ inline long long ls_lds_bytes(int n, int block) {
  // + reduction slots (2 per wave), the running max of b, three counters
  // (work items of each product, matrix steps run), the work items of both
  // products and the weight segments' bits (32-bit each; T is at least as
  // wide)
  // (conv_tiles(n) more: the item rank of each autoconvolution tile)
  const long long elems = (long long)pad_len<SS>(n) + cbuf_len(n) + (GG ? 0 : gbuf_len(n)) +
                          2 * (block / kWave) + 1 + 3 + 3 * conv_tiles(n) + 2 * corr_tiles(n) +
                          segmask_words(n);
  return elems * (long long)sizeof(T);
}

// GG: the correlation output G lives in the candidate's row of A.gscr
// (global memory: required for lengths whose G does not fit the LDS beside a
// and the weights, and faster for the 1024-thread launches, which take it
// whenever the caller provides the rows).  The correlation writes it there
// (put4_global), the update reads and re-zeroes it at workgroup scope and
// parks the unnormalised new height in a's own LDS slot (only its owner
// thread reads that slot during the update) until S is known.  The two halves
// of a split candidate share the row: each writes, reads and zeroes only the
// outputs of the correlation tiles it owns.
template <typename T, int BD, int QMAX, int SS, int D, bool GG = false>
__device__ __forceinline__ void ls_candidate(const LsArgs<T>& A, int cand, int half, unsigned char* smem,
                             int* broken) {
  constexpr int NW = BD / kWave;
  // D: operand steps in flight in the tile loops (a tuning choice)
  const int n = A.ns ? A.ns[cand] : A.n;
  if (A.skip && A.skip[cand]) return;
  if (n <= 0) return;
  const int mode = A.mode ? A.mode[cand] : A.modeDefault;
  T* __restrict__ a = A.pop + (size_t)cand * (size_t)A.stride;

  T* sA = reinterpret_cast<T*>(smem);
  T* sC = sA + pad_len<SS>(n);
  T* sG = sC + cbuf_len(n);
  T* slots = sG + (GG ? 0 : gbuf_len(n));  // 2 * NW
  T* gG = GG ? A.gscr + (size_t)cand * (size_t)A.gstride : nullptr;
  T* gG2 = GG && A.gscr2 ? A.gscr2 + (size_t)cand * (size_t)A.gstride : nullptr;
  T* bmax = slots + 2 * NW;     // max_k b_k of the step (raised by the autoconvolution)
  int* ctr = reinterpret_cast<int*>(bmax + 1);  // [0] conv items, [1] corr items
  unsigned* nsteps = reinterpret_cast<unsigned*>(ctr + 2);  // matrix block steps run
  unsigned* citems = nsteps + 1;
  unsigned* gitems = citems + 2 * conv_tiles(n);
  unsigned* segmask = gitems + 2 * corr_tiles(n);
  int* crank = reinterpret_cast<int*>(segmask + segmask_words(n));  // conv tile -> item rank
  const int tid = threadIdx.x;
  const int LA = pad_len<SS>(n), LC = cbuf_len(n), LG = gbuf_len(n);
  const int SW = segmask_words(n);
  const int NTc = conv_tiles(n);
  // pair state (half >= 0): own progress flag, the partner's, exchanged b_k
  int* myflag = half >= 0 ? A.pair_flags + 16 * (size_t)cand + half : nullptr;
  const int* peerflag = half >= 0 ? A.pair_flags + 16 * (size_t)cand + (1 - half) : nullptr;
  T* xb = half >= 0 ? A.xchg + (size_t)cand * (size_t)A.xstride : nullptr;

  if (tid == 0) { ctr[0] = 0; ctr[1] = 0; *nsteps = 0u; *bmax = T(0); }
  autoconv_items(n, citems, reinterpret_cast<int*>(sC), tid, BD);  // sC is scratch until zeroed
  corr_items(n, gitems, reinterpret_cast<int*>(sC), tid, BD);
  for (int r = tid; r < NTc; r += BD) crank[citems[r] >> (2 * kItemBits)] = r;
  for (int i = tid; i < LA; i += BD) sA[i] = T(0);
  for (int i = tid; i < LC; i += BD) sC[i] = T(0);
  if constexpr (GG) {
    for (int i = tid; i < n; i += BD)
      if (half < 0 || corr_tile_owner(i >> 8) == half) {
        wg_store(gG + i, T(0));
        if (gG2) wg_store(gG2 + i, T(0));  // stays zero where no second item writes
      }
    wait_global_writes();  // zero before any wave's correlation adds onto it
  } else {
    for (int i = tid; i < LG; i += BD) sG[i] = T(0);
  }
  for (int w = tid; w < SW; w += BD) segmask[w] = 0u;
  __syncthreads();
  for (int i = tid; i < n; i += BD) pad_put<SS>(sA, i, a[i]);
  if (half >= 0)  // both halves have read the row before either writes heights into it
    pair_release(myflag, 1), pair_acquire(peerflag, 1, A.pair_err, broken);

  T m[QMAX], v[QMAX];
#pragma unroll
  for (int q = 0; q < QMAX; ++q) { m[q] = T(0); v[q] = T(0); }
  PhaseClock clk;
#ifdef SIDON_PHASE_CLOCKS
  __shared__ unsigned long long s_idle[2];
  clk.start(s_idle);
#endif
  int steps_run = 0;
  __syncthreads();

  const int nb = 2 * n - 1;
  const T lnPr = T(log((double)A.pEnd / (double)A.pStart));
  const T lnLr = T(log((double)A.lrEnd / (double)A.lrStart));
  const T beta1 = A.beta1, beta2 = A.beta2;
  T b1t = T(1), b2t = T(1);
  const T nT = T(n);

  for (int it = 0; it < A.iters; ++it) {
    const T frac = A.iters > 1 ? T(it) / T(A.iters - 1) : T(0);
    const T pexp = A.pStart * dexp(lnPr * frac);
    const T lr = A.lrStart * dexp(lnLr * frac);

    // sC[k] = b_k, *bmax = M (this half's tiles only when paired)
    autoconv_mfma<T, D, SS>(sA, sC, n, ctr, citems, nsteps, bmax, half);
    clk.item_end();
    __syncthreads();
    clk.after_barrier(0);
    if (half >= 0) {  // exchange b_k of the owned tiles and the tile maxima
      for (int k = tid; k < 256 * NTc; k += BD)
        if (conv_item_owner(crank[k >> 8]) == half) coherent_store(xb + k, sC[k]);
      if (tid == 0) coherent_store(xb + 256 * NTc + half, *bmax);
      pair_release(myflag, 2 + 2 * it);
      pair_acquire(peerflag, 2 + 2 * it, A.pair_err, broken);
      for (int k = tid; k < 256 * NTc; k += BD)
        if (conv_item_owner(crank[k >> 8]) != half) sC[k] = coherent_load(xb + k);
      const T other = coherent_load(xb + 256 * NTc + 1 - half);
      __syncthreads();  // every thread has read *bmax
      if (tid == 0 && other > *bmax) *bmax = other;
      __syncthreads();
    }

    if (tid == 0) ctr[0] = 0;  // every wave is past the autoconvolution
    const int st = opaque(tid);  // per-step copy: nothing derived from it is hoisted
    const T M = *bmax;
    if (!(M > T(0))) break;
    clk.mark(0);
    ++steps_run;

    const T pm1 = pexp - T(1);
    // fp32 weights flush to zero once (p-1) log2(b/M) < -126, i.e. b < M exp(-126 ln2 / (p-1));
    // below 0.999 of that bound (a margin of >= 1.4e-3 in the exponent for p >= 2, far above
    // its rounding error) the weight is certainly zero and its transcendentals are skipped
    T cut = T(0);
    if constexpr (sizeof(T) == 4)
      if (pm1 >= T(1)) cut = M * T(0.999) * dexp(T(-87.33654475055310898657) / pm1);
    // fp32: d from one reciprocal of M a step (within an ulp of the quotient);
    // fp64 keeps the quotient of the reference's fp64 polish
    const T invM = T(1) / M;
    T rloc = T(0);
    for (int k = st; k < nb; k += BD) {
      const T b = sC[k];
      T c = T(0);
      if (b >= cut) {
        const T d = sizeof(T) == 4 ? (b - M) * invM : (b - M) / M;
        c = lp_weight_fast(d, pm1);
        rloc += c * (T(1) + d);
      }
      sC[k] = c;
      mark_segments(segmask, k - (st & (kWave - 1)), c != T(0));
    }
    const T R = block_sum<T, NW>(rloc, slots + NW);  // barrier also publishes segmask
    clk.mark(1);

    b1t *= beta1;
    b2t *= beta2;
    const T bc1 = T(1) / (T(1) - b1t), bc2 = T(1) / (T(1) - b2t);

    if constexpr (GG) {
      corr_mfma<T, D, SS, true>(sC, sA, gG, n, ctr + 1, gitems, segmask, nsteps, half, gG2);
      wait_global_writes();  // G complete in memory before the barrier
    } else {
      corr_mfma<T, D, SS>(sC, sA, sG, n, ctr + 1, gitems, segmask, nsteps, half);
    }
    const int ut = opaque(tid);  // element addresses below are built per step, not hoisted
    clk.item_end();
    __syncthreads();
    clk.after_barrier(1);
    clk.mark(2);

    if (tid == 0) { ctr[1] = 0; *bmax = T(0); }  // every wave is past the correlation (and M)
    // the correlation is done with the segment bits; clearing them here, three
    // barriers before the next step's weight pass sets them, keeps the clear
    // from racing with another wave's atomicOr
    for (int w = tid; w < SW; w += BD) segmask[w] = 0u;
    T sloc = T(0);
#pragma unroll
    for (int q = 0; q < QMAX; ++q) {
      const int i = ut + BD * q;
      if (i < n && (half < 0 || corr_tile_owner(i >> 8) == half)) {
        const T ai = sA[pad_addr<SS>(i)];
        T Gi;
        if constexpr (GG) {
          if (gG2) {  // the two-row form: each step's items overwrite the rows
            Gi = wg_load(gG + i) + wg_load(gG2 + i);
          } else {
            Gi = wg_load(gG + i);
            wg_store(gG + i, T(0));  // zero again for the next step's correlation
          }
        } else {
          Gi = sG[i];
        }
        const T gi = dquot(Gi, R) - M;
        T xx;
        if (mode == 0) {
          const T mm = beta1 * m[q] + (T(1) - beta1) * gi;
          const T vv = beta2 * v[q] + (T(1) - beta2) * gi * gi;
          m[q] = mm;
          v[q] = vv;
          const T step = dquot((lr / nT) * (mm * bc1), dsqrt(vv * bc2) + T(1e-30));
          xx = ai - step;
          if (xx < T(0)) xx = T(0);
        } else {
          T e = -lr * (gi / M);
          if (e > T(4)) e = T(4);
          if (e < -T(4)) e = -T(4);
          xx = ai * dexp(e);
        }
        // the owner thread parks the new height in place of G_i (GG: in its
        // own slot of a, which only this thread reads until S is known)
        if constexpr (GG) pad_put<SS>(sA, i, xx);
        else sG[i] = xx;
        sloc += xx;
        if (half >= 0) coherent_store(a + i, xx);
      }
    }
    if (half >= 0) {  // exchange the new heights; S over all of them, in the same order
      pair_release(myflag, 3 + 2 * it);
      pair_acquire(peerflag, 3 + 2 * it, A.pair_err, broken);
      sloc = T(0);
#pragma unroll
      for (int q = 0; q < QMAX; ++q) {
        const int i = ut + BD * q;
        if (i < n) {
          if constexpr (GG) {
            if (corr_tile_owner(i >> 8) != half) pad_put<SS>(sA, i, coherent_load(a + i));
            sloc += sA[pad_addr<SS>(i)];
          } else {
            if (corr_tile_owner(i >> 8) != half) sG[i] = coherent_load(a + i);
            sloc += sG[i];
          }
        }
      }
    }
    T S = block_sum<T, NW>(sloc, slots);
    const bool reset = !(S > T(0));
    if (reset) S = T(1);
    const T inv = T(1) / S;
#pragma unroll
    for (int q = 0; q < QMAX; ++q) {
      const int i = ut + BD * q;
      if (i < n) {
        const T xi = GG ? sA[pad_addr<SS>(i)] : sG[i];
        pad_put<SS>(sA, i, reset ? T(1) / nT : xi * inv);
      }
    }
    // sC needs no clearing: the autoconvolution stores every tile whole
    if constexpr (GG) wait_global_writes();  // G's zeros land before the next correlation
    else for (int i = tid; i < LG; i += BD) sG[i] = T(0);
    __syncthreads();
    clk.mark(3);
  }
  clk.finish(steps_run);

  if (half >= 0)  // the partner has read this half's last heights from the row
    pair_release(myflag, kPairFinal), pair_acquire(peerflag, kPairFinal, A.pair_err, broken);
  for (int i = tid; i < n; i += BD)
    if (half < 0 || corr_tile_owner(i >> 8) == half) a[i] = sA[pad_addr<SS>(i)];
  __syncthreads();  // every wave has added its matrix steps
  if (tid == 0) atomicAdd(&g_block_steps, (unsigned long long)*nsteps);
  if (A.mom) {
    T* mo = A.mom + (size_t)cand * (size_t)A.stride;
#pragma unroll
    for (int q = 0; q < QMAX; ++q) {
      const int i = tid + BD * q;
      if (i < n && (half < 0 || corr_tile_owner(i >> 8) == half)) mo[i] = m[q];
    }
  }
  if (A.vel) {
    T* vo = A.vel + (size_t)cand * (size_t)A.stride;
#pragma unroll
    for (int q = 0; q < QMAX; ++q) {
      const int i = tid + BD * q;
      if (i < n && (half < 0 || corr_tile_owner(i >> 8) == half)) vo[i] = v[q];
    }
  }
  __syncthreads();  // LDS is reused by the next candidate of a persistent block
}

// How many of the longest candidates a persistent launch splits over two
// workgroups.  A candidate's time on one workgroup is modelled as
// t = n (n + kLinCost) (fitted to measured per-length times of the
// 1024-thread kernel: the O(n) phases and the few work items of short
// candidates weigh more than n^2 says), a half of a split candidate as
// h = c t / 2, c the measured cost of a pair relative to one workgroup (the
// exchanges, and half as many work items per workgroup).  Only candidates of
// the longest length n0 (the first L entries of `order`, longest first) are
// split, and only if n0 >= kSplitMinN.  Workgroups take tickets in order: the
// 2s halves first, then the L - s whole long candidates, then the rest.  On G
// workgroups the halves leave r = 2s mod G workgroups one half later than the
// others, and the whole long candidates go round-robin to the earliest free,
// so the long work ends at
//   M = max(a h + (k + [q > 0]) t0, [r > 0] ((a + 1) h + (k + [q > G - r]) t0)),
// a = floor(2s / G), k = floor((L - s) / G), q = (L - s) mod G; the rest of
// the work W_r fills the idle time before M and only the excess, spread over
// G, adds to it.  s minimises that makespan (smallest s on ties).  More long
// candidates than workgroups (L > G) then split just enough to end the
// second round of whole candidates with the first halves (s = L - G);
// uniform lengths with one candidate per workgroup or more split none; a few
// long candidates among many short ones split them all.  Every workgroup
// computes the same count (the sums over the workgroup's threads, the
// minimisation by thread 0, O(L)).
constexpr int kSplitMinN = 4096;
constexpr double kLinCost = 11000.0;
constexpr double kPairCost = 1.5;
__host__ __device__ inline double split_makespan(int s, int L, int G, double t0, double h,
                                                 double rest) {
  const long long A = 2LL * s, B = (long long)L - s;
  const long long a = A / G, r = A % G, k = B / G, q = B % G;
  double M = (double)a * h + (double)(k + (q > 0 ? 1 : 0)) * t0;
  if (r > 0) {
    const double m2 = (double)(a + 1) * h + (double)(k + (q > G - r ? 1 : 0)) * t0;
    if (m2 > M) M = m2;
  }
  const double idle = (double)G * M - ((double)A * h + (double)B * t0);
  if (rest > idle) M += (rest - idle) / (double)G;
  return M;
}
template <typename T, int BD>
__device__ int split_count(const LsArgs<T>& A, int* s_out, unsigned char* smem) {
  if (!A.split || !A.order || !A.ns) return 0;
  constexpr int NW = BD / kWave;
  // L and the rest's cost over the workgroup's threads (a fixed-order block
  // sum: every workgroup forms the same values); smem is free until the
  // first candidate is loaded
  double* red = reinterpret_cast<double*>(smem);
  const int n0 = A.ns[A.order[0]];
  double L = 0.0, rest = 0.0;
  for (int p = (int)threadIdx.x; p < A.P; p += BD) {
    const int n = A.ns[A.order[p]];
    if (n == n0) L += 1.0;
    else rest += (double)n * ((double)n + kLinCost);
  }
  L = block_sum<double, NW>(L, red);
  rest = block_sum<double, NW>(rest, red + NW);
  if (threadIdx.x == 0) {
    int bs = 0;
    if (n0 >= kSplitMinN) {
      const double t0 = (double)n0 * ((double)n0 + kLinCost);
      const int Li = (int)L;
      const int G = (int)gridDim.x;
      const double h = kPairCost * t0 / 2.0;
      double best = split_makespan(0, Li, G, t0, h, rest);
      for (int sc = 1; sc <= Li; ++sc) {
        const double m = split_makespan(sc, Li, G, t0, h, rest);
        if (m < best) { best = m; bs = sc; }
      }
    }
    *s_out = bs;
  }
  __syncthreads();
  const int r = *s_out;
  __syncthreads();  // every thread has read the count and the reduction slots
  return r;
}

// One candidate per workgroup (grid = candidates), or workgroups pulling
// tickets from a device queue (persistent): ticket t < 2 s is half t % 2 of
// candidate order[t / 2] (s = split_count), later tickets whole candidates.
template <typename T, int BD, int QMAX, int SS, int D, bool GG = false>
__device__ __forceinline__ void ls_body(const LsArgs<T>& A, unsigned char* smem, int& s_item,
                                        int* broken) {
  if (threadIdx.x == 0) *broken = 0;
  if (!A.queue) {
    const int cand = A.order ? A.order[blockIdx.x] : (int)blockIdx.x;
    ls_candidate<T, BD, QMAX, SS, D, GG>(A, cand, -1, smem, broken);
    return;
  }
  const int split = split_count<T, BD>(A, &s_item, smem);
  if (A.split_out && blockIdx.x == 0 && threadIdx.x == 0) *A.split_out = split;
  for (;;) {
    if (threadIdx.x == 0) s_item = atomicAdd(A.queue, 1);
    __syncthreads();
    const int item = s_item;
    __syncthreads();
    if (item >= A.P + split) break;
    int cand, half = -1;
    if (item < 2 * split) {
      cand = A.order[item >> 1];
      half = item & 1;
    } else {
      cand = A.order ? A.order[item - split] : item - split;
    }
    ls_candidate<T, BD, QMAX, SS, D, GG>(A, cand, half, smem, broken);
  }
}

template <typename T, int BD, int QMAX, int SS, int D>
__global__ __launch_bounds__(BD) void k_local_search(LsArgs<T> A) {
  extern __shared__ __align__(16) unsigned char smem[];
  __shared__ int s_item, s_broken;
  ls_body<T, BD, QMAX, SS, D>(A, smem, s_item, &s_broken);
}

// The correlation output G in global memory (A.gscr, the caller's rows),
// with the plain 16-float segment layout (whose B-operand reads are served in
// four passes; measured faster all the same).
template <typename T, int BD, int QMAX, int SS, int D>
__global__ __launch_bounds__(BD) void k_local_search_gg(LsArgs<T> A) {
  extern __shared__ __align__(16) unsigned char smem[];
  __shared__ int s_item, s_broken;
  ls_body<T, BD, QMAX, SS, D, true>(A, smem, s_item, &s_broken);
}

// The same kernel held to 128 registers a lane (four waves a SIMD): two
// 512-thread or four 256-thread candidates then share a CU where their LDS
// allows it, and one candidate's elementwise phases and barrier waits overlap
// the others' matrix work.
template <typename T, int BD, int QMAX, int SS, int D>
__global__ __launch_bounds__(BD) __attribute__((amdgpu_waves_per_eu(4, 4))) void
k_local_search_shared(LsArgs<T> A) {
  extern __shared__ __align__(16) unsigned char smem[];
  __shared__ int s_item, s_broken;
  ls_body<T, BD, QMAX, SS, D>(A, smem, s_item, &s_broken);
}
// The same with the correlation output in global memory (A.gscr).
template <typename T, int BD, int QMAX, int SS, int D>
__global__ __launch_bounds__(BD) __attribute__((amdgpu_waves_per_eu(4, 4))) void
k_local_search_shared_gg(LsArgs<T> A) {
  extern __shared__ __align__(16) unsigned char smem[];
  __shared__ int s_item, s_broken;
  ls_body<T, BD, QMAX, SS, D, true>(A, smem, s_item, &s_broken);
}

// ---------------------------------------------------------------- lp_grad --
template <typename T>
struct GradArgs {
  const T* a;
  int n, B;
  T p;
  T* J;
  T* grad;
};

// GG: lengths whose correlation output does not fit the LDS; the
// correlation accumulates straight into the candidate's gradient row.
template <typename T, int BD, int SS, bool GG = false>
__global__ __launch_bounds__(BD) void k_lp_grad(GradArgs<T> A) {
  constexpr int NW = BD / kWave;
  constexpr int D = (BD >= 1024 || sizeof(T) == 8) ? 2 : 4;
  extern __shared__ __align__(16) unsigned char smem[];
  const int n = A.n, cand = blockIdx.x;
  const T* a = A.a + (size_t)cand * n;
  T* g = A.grad ? A.grad + (size_t)cand * n : nullptr;
  T* sA = reinterpret_cast<T*>(smem);
  T* sC = sA + pad_len<SS>(n);
  T* sG = sC + cbuf_len(n);
  T* slots = sG + (GG ? 0 : gbuf_len(n));
  T* bmax = slots + 2 * NW;
  int* ctr = reinterpret_cast<int*>(bmax + 1);
  unsigned* nsteps = reinterpret_cast<unsigned*>(ctr + 2);  // matrix block steps run
  unsigned* citems = reinterpret_cast<unsigned*>(ctr + 3);
  unsigned* gitems = citems + 2 * conv_tiles(n);
  unsigned* segmask = gitems + 2 * corr_tiles(n);
  const int tid = threadIdx.x;
  const int LA = pad_len<SS>(n), LC = cbuf_len(n), LG = gbuf_len(n);
  if (tid == 0) { ctr[0] = 0; ctr[1] = 0; *nsteps = 0u; *bmax = T(0); }
  for (int w = tid; w < segmask_words(n); w += BD) segmask[w] = 0u;
  autoconv_items(n, citems, reinterpret_cast<int*>(sC), tid, BD);
  corr_items(n, gitems, reinterpret_cast<int*>(sC), tid, BD);
  for (int i = tid; i < LA; i += BD) sA[i] = T(0);
  for (int i = tid; i < LC; i += BD) sC[i] = T(0);
  if constexpr (GG) {
    if (g) {
      for (int i = tid; i < n; i += BD) wg_store(g + i, T(0));
      wait_global_writes();
    }
  } else {
    for (int i = tid; i < LG; i += BD) sG[i] = T(0);
  }
  __syncthreads();
  T sl = T(0);
  for (int i = tid; i < n; i += BD) { const T x = a[i]; pad_put<SS>(sA, i, x); sl += x; }
  const T S = block_sum<T, NW>(sl, slots);  // barrier also publishes sA
  const int nb = 2 * n - 1;
  autoconv_mfma<T, D, SS>(sA, sC, n, ctr, citems, nsteps, bmax);  // sC[k] = b_k, *bmax = M
  __syncthreads();
  const T M = *bmax;
  const T pm1 = A.p - T(1);
  T rloc = T(0);
  for (int k = tid; k < nb; k += BD) {
    const T d = (sC[k] - M) / M;
    const T c = lp_weight(d, pm1);
    rloc += c * (T(1) + d);
    sC[k] = c;
    mark_segments(segmask, k - (tid & (kWave - 1)), c != T(0));
  }
  const T R = block_sum<T, NW>(rloc, slots + NW);
  const T Rp = dexp(log(R) / A.p);
  if (A.J && tid == 0) A.J[cand] = T(2) * T(n) * M * Rp / (S * S);
  if (!g) {  // the autoconvolution's steps are counted (the barrier of R is past them)
    if (tid == 0) atomicAdd(&g_block_steps, (unsigned long long)*nsteps);
    return;
  }
  if constexpr (GG) {
    corr_mfma<T, D, SS, true>(sC, sA, g, n, ctr + 1, gitems, segmask, nsteps);
    wait_global_writes();
  } else {
    corr_mfma<T, D, SS>(sC, sA, sG, n, ctr + 1, gitems, segmask, nsteps);
  }
  __syncthreads();
  if (tid == 0) atomicAdd(&g_block_steps, (unsigned long long)*nsteps);
  const T scale = T(4) * T(n) * Rp / (S * S);
  const T MS = M / S;
  for (int i = tid; i < n; i += BD) g[i] = scale * ((GG ? wg_load(g + i) : sG[i]) / R - MS);
}

// ------------------------------------------------------------- host side --
namespace {

// Launch configuration of the fp32 local search at one length: waves per
// candidate (workgroup size), operand steps in flight in the tile loops, and
// candidates a CU (2 or 4: the 512- or 256-thread workgroup held to 128
// registers a lane, so two or four share a CU where their LDS allows; 1: the
// kernel without that cap).
struct LsTile { int block, depth, per_cu; };

// Per-length overrides set by the autotuner (n -> configuration).
struct TileRule { int n; LsTile tile; };
TileRule g_rules[64];
int g_nrules = 0;
std::mutex g_rules_mu;  // the table is process-wide; launches read it from any thread

// The built-in rule, a function of n alone (measured, the anneal p 4 -> 1e4):
// four 256-thread candidates a CU with two operand steps in flight below
// n = 1024 (faster than every other configuration whether a launch holds two
// or four candidates a CU: 1.1-1.7x at n = 128, 1.03-1.26x at 600); two
// 512-thread candidates a CU with four steps up to 4096 (at 1200-1600 four
// 256-thread candidates a CU win only when the launch has four a CU to give
// them, which the autotuner sees and the rule does not); 1024 threads above,
// where one candidate fills a CU's LDS.
constexpr int kQuadMaxN = 1023;
LsTile default_tile(int n) {
  if (n > 4096) return LsTile{1024, 4, 1};
  if (n > kQuadMaxN) return LsTile{512, 4, 2};
  return LsTile{256, 2, 4};
}

LsTile pick_tile(int nmax) {
  std::lock_guard<std::mutex> lk(g_rules_mu);
  for (int r = 0; r < g_nrules; ++r)
    if (g_rules[r].n == nmax) return g_rules[r].tile;
  return default_tile(nmax);
}

bool tile_valid(const LsTile& c) {
  if (c.depth != 2 && c.depth != 4) return false;
  switch (c.block) {
    case 256: return c.per_cu == 1 || c.per_cu == 4;
    case 512: return c.per_cu == 1 || c.per_cu == 2;
    case 1024: return c.per_cu == 1;
    default: return false;
  }
}

template <typename T, int BD, int QMAX, bool SHARED, int SS, int D, bool GG = false>
int launch_ls_layout(const LsArgs<T>& A, int nmax, hipStream_t stream) {
  if (nmax > BD * QMAX) return SIDON_ERR_UNSUPPORTED;
  int rc = SIDON_OK;
  DeviceState* ds = device_state(&rc);
  if (!ds) return rc;
  const long long lds = ls_lds_bytes<T, SS, GG>(nmax, BD);
  if (lds > ds->lds_max) {
    set_last_error_msg("local search: n needs more LDS than a workgroup may use");
    return SIDON_ERR_LDS;
  }
  auto kfn = [] {
    if constexpr (GG && SHARED) return k_local_search_shared_gg<T, BD, QMAX, SS, D>;
    else if constexpr (GG) return k_local_search_gg<T, BD, QMAX, SS, D>;
    else if constexpr (SHARED) return k_local_search_shared<T, BD, QMAX, SS, D>;
    else return k_local_search<T, BD, QMAX, SS, D>;
  }();
  SIDON_HIP_TRY(hipFuncSetAttribute(reinterpret_cast<const void*>(kfn),
                                    hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  int grid = A.P;
  if (A.queue) {
    // persistent: as many workgroups as are resident at once (two or four a
    // CU for the register-capped kernel where LDS allows), at most P
    int per_cu = 0;
    SIDON_HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(
        &per_cu, reinterpret_cast<const void*>(kfn), BD, (size_t)lds));
    if (per_cu < 1) per_cu = 1;
    const long long resident = (long long)per_cu * ds->cu_count;
    const long long tickets = A.split && A.order ? 2LL * A.P : A.P;  // at most P + P halves
    grid = resident < tickets ? (int)resident : (int)tickets;
  }
  hipLaunchKernelGGL(kfn, dim3(grid), dim3(BD), (size_t)lds, stream, A);
  SIDON_HIP_TRY(hipGetLastError());
  return SIDON_OK;
}

// fp32 with the caller's rows for the correlation output (A.gscr): G in
// global memory and the plain 16-float layout (faster at every configuration
// measured: 6-10 % for two 512-thread candidates a CU at n = 1200-4096, 4-5 %
// for 1024 threads at 4200-8400); otherwise the wide (bank-conflict-free)
// segment layout where it fits the LDS, else the narrow one.
template <typename T, int BD, int QMAX, bool SHARED, int D>
int launch_ls(const LsArgs<T>& A, int nmax, hipStream_t stream) {
  int rc = SIDON_OK;
  DeviceState* ds = device_state(&rc);
  if (!ds) return rc;
  if constexpr (sizeof(T) == 4) {
    if constexpr (BD < 1024)  // 1024 threads: launch_ls_gg, 13 heights a thread
      if (A.gscr) return launch_ls_layout<T, BD, QMAX, SHARED, kSeg, D, true>(A, nmax, stream);
    if (ls_lds_bytes<T, kSegStrideWide>(nmax, BD) <= ds->lds_max)
      return launch_ls_layout<T, BD, QMAX, SHARED, kSegStrideWide, D>(A, nmax, stream);
  }
  return launch_ls_layout<T, BD, QMAX, SHARED, kSegStride, D>(A, nmax, stream);
}

// The 1024-thread fp32 launch with G in global memory (A.gscr) and the plain
// 16-float segment layout: required above the on-chip form's lengths, and
// taken at every length the 1024-thread configuration runs when the caller
// provides the rows (measured at n = 8400: 714 k candidate-steps/s against
// 676 k with G in LDS and 24-float segments; the correlation's epilogues and
// the update no longer go through the LDS the matrix operands are read from).
constexpr int kGgQmax = 13;  // n <= 13312; the plain layout's LDS ends at 13216
template <typename T>
bool ls_needs_gg(int nmax, long long lds_max) {
  return ls_lds_bytes<T, kSegStride>(nmax, 1024) > lds_max;
}
int launch_ls_gg(const LsArgs<float>& A, int nmax, int depth, hipStream_t stream) {
  if (!A.gscr) {
    set_last_error_msg("local search: n above the on-chip form needs mom or vel "
                       "(the launch's scratch for the correlation output)");
    return SIDON_ERR_LDS;
  }
  if (depth == 2)
    return launch_ls_layout<float, 1024, kGgQmax, false, kSeg, 2, true>(A, nmax, stream);
  return launch_ls_layout<float, 1024, kGgQmax, false, kSeg, 4, true>(A, nmax, stream);
}

// fp32 configurations by operand depth
template <int BD, int QMAX, bool SHARED>
int launch_ls_depth(const LsArgs<float>& A, int nmax, int depth, hipStream_t stream) {
  if (depth == 2) return launch_ls<float, BD, QMAX, SHARED, 2>(A, nmax, stream);
  if (depth == 4) return launch_ls<float, BD, QMAX, SHARED, 4>(A, nmax, stream);
  return SIDON_ERR_UNSUPPORTED;
}

// Workspace of a persistent launch (bytes): the queue counter and the pair
// timeout count, then 16 progress ints per candidate, then per candidate the
// b_k exchange of a split pair (256 per autoconvolution tile, and the two
// halves' maxima).
// Then two sets of correlation output rows (stride floats per candidate each:
// the two-row form of the correlation output).
constexpr size_t kWsHeader = 256;
inline long long xchg_stride(int stride) { return 256LL * conv_tiles(stride) + 64; }
inline bool ragged_ws_has_g(int) { return true; }
inline size_t ragged_ws_g_offset(int P, int stride) {
  return kWsHeader + (size_t)P * 64 + (size_t)P * (size_t)xchg_stride(stride) * sizeof(float);
}
inline size_t ragged_ws_bytes(int P, int stride) {
  return ragged_ws_g_offset(P, stride) +
         (ragged_ws_has_g(stride) ? 2 * (size_t)P * (size_t)stride * sizeof(float) : 0);
}

// Clears the queue counter, the timeout count and the candidates' progress
// flags (a kernel, not a memset node, so the launch can be captured).
__global__ void k_clear_ints(int* p, long long count) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < count;
       i += (long long)gridDim.x * blockDim.x)
    p[i] = 0;
}

// `ws` non-null: persistent launch pulling tickets from the workspace queue.
template <typename T>
int dispatch_ls(LsArgs<T>& A, int nmax, const LsTile& tile, void* ws, hipStream_t stream) {
  A.queue = nullptr;
  if (ws) {
    int* wi = static_cast<int*>(ws);
    const long long ints = (long long)(kWsHeader / sizeof(int)) + 16LL * A.P;
    const int grid = (int)((ints + 255) / 256 < 1024 ? (ints + 255) / 256 : 1024);
    hipLaunchKernelGGL(k_clear_ints, dim3(grid), dim3(256), 0, stream, wi, ints);
    SIDON_HIP_TRY(hipGetLastError());
    A.queue = wi;
    A.pair_err = wi + 1;
    A.split_out = wi + 2;
    A.pair_flags = wi + kWsHeader / sizeof(int);
    A.xchg = reinterpret_cast<T*>(static_cast<char*>(ws) + kWsHeader + (size_t)A.P * 64);
    A.xstride = xchg_stride((int)A.stride);
    if (sizeof(T) == 4 && ragged_ws_has_g((int)A.stride)) {
      A.gscr = reinterpret_cast<T*>(static_cast<char*>(ws) +
                                    ragged_ws_g_offset(A.P, (int)A.stride));
      A.gstride = A.stride;
      A.gscr2 = A.gscr + (size_t)A.P * (size_t)A.stride;
    }
  }
  if constexpr (sizeof(T) == 4) {
    int rc = SIDON_OK;
    DeviceState* ds = device_state(&rc);
    if (!ds) return rc;
    // above the on-chip form: 1024 threads, G in global memory (rows needed);
    // 1024-thread launches with rows: the same form
    if (ls_needs_gg<float>(nmax, ds->lds_max))
      return launch_ls_gg(A, nmax, tile.block == 1024 ? tile.depth : 4, stream);
    if (tile.block == 1024 && A.gscr) return launch_ls_gg(A, nmax, tile.depth, stream);
  }
  if constexpr (sizeof(T) == 8) {  // fp64: fixed configurations, two steps in flight
    int rc = SIDON_OK;
    DeviceState* ds = device_state(&rc);
    if (!ds) return rc;
    // lengths whose fp64 G does not fit the LDS beside a and the weights: the
    // correlation output in the caller's rows, 512 threads, plain layout
    if (ls_lds_bytes<T, kSegStride>(nmax, tile.block) > ds->lds_max) {
      if (!A.gscr) {
        set_last_error_msg("local search (fp64): n above the on-chip form needs mom or vel "
                           "(the launch's scratch for the correlation output)");
        return SIDON_ERR_LDS;
      }
      return launch_ls_layout<T, 512, kGgQmax, false, kSeg, 2, true>(A, nmax, stream);
    }
    switch (tile.block) {
      case 256: return launch_ls<T, 256, 16, false, 2>(A, nmax, stream);
      case 512: return launch_ls<T, 512, 12, false, 2>(A, nmax, stream);
      default: return SIDON_ERR_UNSUPPORTED;
    }
  } else {
    switch (tile.block) {
      case 256:
        if (tile.per_cu == 4) {
          if (nmax > 256 * 9) return SIDON_ERR_UNSUPPORTED;
          return launch_ls_depth<256, 9, true>(A, nmax, tile.depth, stream);
        }
        return launch_ls_depth<256, 16, false>(A, nmax, tile.depth, stream);
      case 512:
        if (tile.per_cu == 2) {
          if (nmax > 512 * 8) return SIDON_ERR_UNSUPPORTED;
          return launch_ls_depth<512, 8, true>(A, nmax, tile.depth, stream);
        }
        return launch_ls_depth<512, 20, false>(A, nmax, tile.depth, stream);
      case 1024: return launch_ls_depth<1024, 10, false>(A, nmax, tile.depth, stream);
      default: return SIDON_ERR_UNSUPPORTED;
    }
  }
}

// The moments' buffers may serve as the correlation output's scratch, so
// neither may overlap the population (all three are [P x n]).
template <typename T>
bool rows_overlap(const T* pop, const T* buf, int n, int P) {
  if (!buf || P <= 0) return false;
  const size_t len = (size_t)n * (size_t)P;
  return buf < pop + len && pop < buf + len;
}

template <typename T>
int check_schedule(int iters, T pS, T pE, T lS, T lE) {
  if (iters < 0) return SIDON_ERR_ARG;
  if (iters == 0) return SIDON_OK;
  if (!(pS > T(0)) || !(pE > T(0)) || !(lS > T(0)) || !(lE > T(0))) return SIDON_ERR_ARG;
  return SIDON_OK;
}

template <typename T, int BD, int SS, bool GG = false>
int launch_grad_layout(const GradArgs<T>& A, hipStream_t stream) {
  int rc = SIDON_OK;
  DeviceState* ds = device_state(&rc);
  if (!ds) return rc;
  const long long lds = ls_lds_bytes<T, SS, GG>(A.n, BD);
  if (lds > ds->lds_max) {
    set_last_error_msg("lp_grad: n needs more LDS than a workgroup may use");
    return SIDON_ERR_LDS;
  }
  auto kfn = k_lp_grad<T, BD, SS, GG>;
  SIDON_HIP_TRY(hipFuncSetAttribute(reinterpret_cast<const void*>(kfn),
                                    hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  hipLaunchKernelGGL(kfn, dim3(A.B), dim3(BD), (size_t)lds, stream, A);
  SIDON_HIP_TRY(hipGetLastError());
  return SIDON_OK;
}

template <typename T, int BD>
int launch_grad(const GradArgs<T>& A, hipStream_t stream) {
  int rc = SIDON_OK;
  DeviceState* ds = device_state(&rc);
  if (!ds) return rc;
  if constexpr (sizeof(T) == 4)
    if (ls_lds_bytes<T, kSegStrideWide>(A.n, BD) <= ds->lds_max)
      return launch_grad_layout<T, BD, kSegStrideWide>(A, stream);
  if (ls_lds_bytes<T, kSegStride>(A.n, BD) <= ds->lds_max)
    return launch_grad_layout<T, BD, kSegStride>(A, stream);
  // the correlation output does not fit beside a and the weights: it goes to
  // the gradient rows (none when only J is asked for), with the widest
  // segment stride that fits (lengths that long run 1024 threads)
  if constexpr (BD == 1024) {
    if constexpr (sizeof(T) == 4)
      if (ls_lds_bytes<T, kSegStrideWide, true>(A.n, BD) <= ds->lds_max)
        return launch_grad_layout<T, BD, kSegStrideWide, true>(A, stream);
    if (ls_lds_bytes<T, kSegStride, true>(A.n, BD) <= ds->lds_max)
      return launch_grad_layout<T, BD, kSegStride, true>(A, stream);
    return launch_grad_layout<T, BD, kSeg, true>(A, stream);
  }
  return launch_grad_layout<T, BD, kSegStride>(A, stream);
}

}  // namespace
}  // namespace sidon

using namespace sidon;

extern "C" int sidon_set_ls_tile(int n, int block, int depth, int per_cu) {
  const bool clear = block == 0;
  const LsTile tile{block, depth, per_cu};
  if (!clear && !tile_valid(tile)) return SIDON_ERR_ARG;
  std::lock_guard<std::mutex> lk(g_rules_mu);
  for (int r = 0; r < g_nrules; ++r)
    if (g_rules[r].n == n) {
      if (clear) { g_rules[r] = g_rules[--g_nrules]; return SIDON_OK; }
      g_rules[r].tile = tile;
      return SIDON_OK;
    }
  if (clear) return SIDON_OK;
  if (g_nrules >= 64) return SIDON_ERR_ARG;
  g_rules[g_nrules++] = TileRule{n, tile};
  return SIDON_OK;
}

extern "C" int sidon_get_ls_tile(int n, int* block, int* depth, int* per_cu) {
  const LsTile t = pick_tile(n);
  if (block) *block = t.block;
  if (depth) *depth = t.depth;
  if (per_cu) *per_cu = t.per_cu;
  return SIDON_OK;
}

extern "C" int sidon_set_ls_block(int n, int block) {
  if (block == 0) return sidon_set_ls_tile(n, 0, 0, 0);
  if (block != 256 && block != 512 && block != 1024) return SIDON_ERR_ARG;
  // the built-in configuration when it has this workgroup size, else four
  // steps in flight and the CU-sharing form where the rule's lengths allow it
  const LsTile d = default_tile(n);
  if (d.block == block) return sidon_set_ls_tile(n, d.block, d.depth, d.per_cu);
  const int per_cu = block == 512 && n <= 4096 ? 2 : block == 256 && n <= kQuadMaxN ? 4 : 1;
  return sidon_set_ls_tile(n, block, 4, per_cu);
}

extern "C" int sidon_get_ls_block(int n) { return pick_tile(n).block; }

extern "C" int sidon_local_search(float* pop, float* mom, float* vel, const int* skip,
                                  const int* mode, int n, int P, int iters, float pStart,
                                  float pEnd, float lrStart, float lrEnd, float beta1,
                                  float beta2, int modeDefault, hipStream_t stream) {
  if (!pop || n < 1 || P < 0) return SIDON_ERR_ARG;
  if (rows_overlap(pop, mom, n, P) || rows_overlap(pop, vel, n, P)) return SIDON_ERR_ARG;
  if (int rc = check_schedule<float>(iters, pStart, pEnd, lrStart, lrEnd)) return rc;
  if (P == 0 || iters == 0) return SIDON_OK;
  LsArgs<float> A{pop, n, nullptr, n, P, mom, vel, skip, mode, modeDefault, nullptr, nullptr,
                  iters, pStart, pEnd, lrStart, lrEnd, beta1, beta2};
  // above the on-chip form the moments' buffers are the correlation output's
  // scratch until the final moments are written (both, when both are given)
  A.gscr = mom ? mom : vel;
  A.gstride = n;
  A.gscr2 = mom && vel && !rows_overlap(mom, vel, n, P) ? vel : nullptr;
  return dispatch_ls<float>(A, n, pick_tile(n), nullptr, stream);
}

extern "C" int sidon_local_search_f64(double* pop, double* mom, double* vel, const int* skip,
                                      const int* mode, int n, int P, int iters, double pStart,
                                      double pEnd, double lrStart, double lrEnd, double beta1,
                                      double beta2, int modeDefault, hipStream_t stream) {
  if (!pop || n < 1 || P < 0) return SIDON_ERR_ARG;
  if (rows_overlap(pop, mom, n, P) || rows_overlap(pop, vel, n, P)) return SIDON_ERR_ARG;
  if (int rc = check_schedule<double>(iters, pStart, pEnd, lrStart, lrEnd)) return rc;
  if (P == 0 || iters == 0) return SIDON_OK;
  LsArgs<double> A{pop, n, nullptr, n, P, mom, vel, skip, mode, modeDefault, nullptr, nullptr,
                   iters, pStart, pEnd, lrStart, lrEnd, beta1, beta2};
  // above the on-chip form the moments' buffers are the correlation output's
  // scratch until the final moments are written (both, when both are given)
  A.gscr = mom ? mom : vel;
  A.gstride = n;
  A.gscr2 = mom && vel && !rows_overlap(mom, vel, n, P) ? vel : nullptr;
  const LsTile tile{n >= 1024 ? 512 : 256, 2, 1};
  return dispatch_ls<double>(A, n, tile, nullptr, stream);
}

extern "C" size_t sidon_local_search_ragged_workspace(int P, int stride) {
  if (P < 0 || stride < 1) return 0;
  return ragged_ws_bytes(P, stride);
}

extern "C" int sidon_local_search_ragged(float* pop, int stride, const int* ns, int P, int nmax,
                                         const int* order, const int* skip, const int* mode,
                                         int iters, float pStart, float pEnd, float lrStart,
                                         float lrEnd, float beta1, float beta2, int modeDefault,
                                         int split, void* workspace, size_t workspace_bytes,
                                         hipStream_t stream) {
  if (!pop || !ns || !workspace || nmax < 1 || stride < nmax || P < 0) return SIDON_ERR_ARG;
  if (split != 0 && split != 1) return SIDON_ERR_ARG;
  if (int rc = check_schedule<float>(iters, pStart, pEnd, lrStart, lrEnd)) return rc;
  if (workspace_bytes < ragged_ws_bytes(P, stride)) {
    set_last_error_msg("local search (ragged): workspace smaller than "
                       "sidon_local_search_ragged_workspace(P, stride)");
    return SIDON_ERR_ARG;
  }
  if (P == 0 || iters == 0) return SIDON_OK;
  LsArgs<float> A{pop, stride, ns, nmax, P, nullptr, nullptr, skip, mode, modeDefault, order,
                  nullptr, iters, pStart, pEnd, lrStart, lrEnd, beta1, beta2,
                  split, nullptr, nullptr, nullptr, nullptr, 0};
  return dispatch_ls<float>(A, nmax, pick_tile(nmax), workspace, stream);
}

extern "C" int sidon_lp_grad(const float* a, int n, int B, float p, float* J, float* grad,
                             hipStream_t stream) {
  if (!a || n < 1 || B < 0 || !(p >= 1.0f)) return SIDON_ERR_ARG;
  if (B == 0 || (!J && !grad)) return SIDON_OK;
  GradArgs<float> A{a, n, B, p, J, grad};
  return n >= 2048 ? launch_grad<float, 1024>(A, stream) : launch_grad<float, 256>(A, stream);
}

extern "C" int sidon_lp_grad_f64(const double* a, int n, int B, double p, double* J,
                                 double* grad, hipStream_t stream) {
  if (!a || n < 1 || B < 0 || !(p >= 1.0)) return SIDON_ERR_ARG;
  if (B == 0 || (!J && !grad)) return SIDON_OK;
  GradArgs<double> A{a, n, B, p, J, grad};
  return n >= 2048 ? launch_grad<double, 1024>(A, stream) : launch_grad<double, 256>(A, stream);
}

extern "C" long long sidon_ls_lds_bytes_impl(int kernel, int n, int block) {
  if (n < 1) return -1;
  switch (kernel) {
    case 0: case 4: {  // the layout a 160 KiB workgroup takes: wide where it fits
      constexpr long long cap = 160LL * 1024;
      const long long wide = ls_lds_bytes<float, kSegStrideWide>(n, block);
      if (wide <= cap) return wide;
      const long long narrow = ls_lds_bytes<float>(n, block);
      if (narrow <= cap) return narrow;
      // above the on-chip form: the correlation output in global memory
      // (the local search always runs 1024 threads there, plain layout; the
      // gradient the widest layout that fits)
      if (kernel == 0) return ls_lds_bytes<float, kSeg, true>(n, 1024);
      const long long g24 = ls_lds_bytes<float, kSegStrideWide, true>(n, block);
      if (g24 <= cap) return g24;
      const long long g20 = ls_lds_bytes<float, kSegStride, true>(n, block);
      return g20 <= cap ? g20 : ls_lds_bytes<float, kSeg, true>(n, block);
    }
    case 1: {  // G in LDS where it fits, else in the caller's rows (512 threads)
      const long long onchip = ls_lds_bytes<double>(n, block);
      return onchip <= 160LL * 1024 ? onchip : ls_lds_bytes<double, kSeg, true>(n, 512);
    }
    default: return -1;
  }
}

extern "C" int sidon_ls_phase_clocks(unsigned long long* out, int reset) {
  if (!out) return SIDON_ERR_ARG;
#ifdef SIDON_PHASE_CLOCKS
  int rc = SIDON_OK;
  if (!device_state(&rc)) return rc;
  static std::mutex mu;
  std::lock_guard<std::mutex> lk(mu);
  hipLaunchKernelGGL(k_take_phase_clocks, dim3(1), dim3(1), 0, 0, reset);
  SIDON_HIP_TRY(hipGetLastError());
  SIDON_HIP_TRY(hipMemcpyFromSymbol(out, HIP_SYMBOL(g_phase_clk_taken),
                                    sizeof(unsigned long long) * kPhaseSlots));
  return SIDON_OK;
#else
  (void)reset;
  set_last_error_msg("phase clocks: this library was built without SIDON_PHASE_CLOCKS");
  return SIDON_ERR_UNSUPPORTED;
#endif
}

extern "C" int sidon_ls_block_steps(unsigned long long* out, int reset) {
  if (!out) return SIDON_ERR_ARG;
  int rc = SIDON_OK;
  if (!device_state(&rc)) return rc;
  static std::mutex mu;  // one reader of g_block_steps_taken at a time
  std::lock_guard<std::mutex> lk(mu);
  hipLaunchKernelGGL(k_take_block_steps, dim3(1), dim3(1), 0, 0, reset);
  SIDON_HIP_TRY(hipGetLastError());
  SIDON_HIP_TRY(hipMemcpyFromSymbol(out, HIP_SYMBOL(g_block_steps_taken), sizeof(*out)));
  return SIDON_OK;
}
