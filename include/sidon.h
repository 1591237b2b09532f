/*
 * sidon.h -- C ABI of the MI355X (gfx950) Sidon-autocorrelation search path.
 *
 * Objective: F(a) = 2n * max_k b_k / (sum_i a_i)^2, b_k = sum_{i+j=k} a_i a_j,
 * a_i >= 0 (reference: src/kernels.cuh:1-13).
 *
 * Conventions (every entry point):
 *  - all array arguments are DEVICE pointers on the current HIP device,
 *    owned by the caller; population arrays are row-major [P x n];
 *  - work is enqueued on `stream` (0 = the null stream) and is asynchronous;
 *  - nothing here allocates device memory; the generation-graph constructor
 *    allocates only its graph objects.  Work-queue counters and workspaces
 *    are caller-owned arguments, so launches on different streams with
 *    distinct buffers never share state;
 *  - return value: SIDON_OK (0) or one of the sidon_status codes; the entry
 *    points never exit or throw.  sidon_last_error() describes the last error.
 *
 * The reference has no library API: its path sits behind static host helpers
 * in src/main.cu that launch templated kernels from src/kernels.cuh.  Each
 * entry point names the helper/kernel it replaces.
 */
#ifndef SIDON_H
#define SIDON_H

#include <hip/hip_runtime_api.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
  SIDON_OK = 0,
  SIDON_ERR_ARG = 1,         /* bad size / null pointer / bad parameter */
  SIDON_ERR_LDS = 2,         /* n needs more LDS than one workgroup may use
                                (the reference returns false / exits, main.cu:173, :391-399) */
  SIDON_ERR_HIP = 3,         /* a HIP runtime call failed */
  SIDON_ERR_UNSUPPORTED = 4  /* configuration not built (e.g. block size) */
} sidon_status;

/* Reads and caches the attributes of `device` (-1 = current device) that the
 * launches need (LDS per workgroup, CU count, the device address of the
 * no-alloc polish's static scratch array).  Allocates nothing; called
 * implicitly on first use.  No reference counterpart. */
int sidon_init(int device);
/* Human-readable text of the last error on this thread. */
const char* sidon_last_error(void);
const char* sidon_status_string(int status);
/* Dynamic LDS bytes that `kernel` needs at length n with `block` threads
 * (kernel: 0 local search fp32, 1 local search fp64, 2 eval, 3 Frank-Wolfe,
 * 4 lp_grad fp32), or -1 if unsupported.  Counterpart of localSearchSmem /
 * evalSmem / fwSmem (src/kernels.cuh:161-162, :379-380, :405-406). */
long long sidon_lds_bytes(int kernel, int n, int block);
/* Largest dynamic LDS per workgroup on the current device. */
int sidon_lds_limit(void);

/* ---- FP64 objective --------------------------------------------------- */

/* F in fp64 from fp32 heights plus the count of near-active peaks
 * b_k >= M (1 - activeTol); nActive may be NULL.
 * Replaces runEval (src/main.cu:161-166) / kEval<float,double>
 * (src/kernels.cuh:301-377).  F = 1e30 when sum a <= 0. */
int sidon_eval_f64(const float* pop, int n, int P, double* F, int* nActive,
                   double activeTol, hipStream_t stream);
/* Uneven lengths: candidate p has ns[p] <= nmax heights at pop + p*stride
 * (ns on the device).  No reference counterpart (one n per launch there). */
int sidon_eval_f64_ragged(const float* pop, int stride, const int* ns, int P, int nmax,
                          double* F, int* nActive, double activeTol, hipStream_t stream);
/* Same for fp64 heights: kEval<double,double> (src/main.cu:843). */
int sidon_eval_f64_d(const double* pop, int n, int P, double* F, int* nActive,
                     double activeTol, hipStream_t stream);
/* Correctly rounded F: every b_k and S rounded once from their exact values
 * (error-free products, double-double sums with a certified error bound).
 * ambiguous[p] counts sums whose rounding the bound could not certify (0 in
 * practice).  With correctly rounded b and S the n -> 2n cell split preserves
 * F bit for bit.  M, S, ambiguous may be NULL.  No reference counterpart
 * (the reference checks refinement to 1e-12, src/main.cu:208-219). */
int sidon_eval_exact(const float* pop, int n, int P, double* F, double* M, double* S,
                     int* ambiguous, hipStream_t stream);

/* ---- L^p-annealed projected Adam / mirror descent -------------------- */

/* `iters` fused steps per candidate with `a` resident in LDS:
 * b = a*a, M = max b, c = (b/M)^(p-1) (log1p form), R = sum u^p, G = corr(c,a),
 * direction g = G/R - M; mode 0 projected Adam (step (lr/n) mhat/(sqrt(vhat)+1e-30),
 * clamp >= 0), mode 1 mirror descent (a exp(clamp(-lr g/M, +-4))); then
 * renormalise sum a = 1.  p and lr are annealed geometrically from
 * (pStart, lrStart) to (pEnd, lrEnd).  Adam moments restart at zero every call;
 * mom/vel (may be NULL) receive the final moments.  skip[p] != 0 leaves
 * candidate p untouched; mode (may be NULL) overrides modeDefault per candidate.
 * When mom (or, when mom is NULL, vel) is given, the launch keeps the
 * correlation output in its rows until the final moments are written (the
 * same bits, measured faster); with both given (and apart) it uses the rows
 * of both, which moves about an eighth of the memory-side traffic of one
 * (the halves of a split tile are stored, not added atomically).  Lengths
 * whose correlation output does not fit one workgroup's LDS beside a and the
 * weights (n > 9296 on a 160 KiB workgroup) need one of them (SIDON_ERR_LDS
 * otherwise; the reference always passes both).  Neither may overlap pop
 * (SIDON_ERR_ARG).  The longest fp32 length is 13216
 * (sidon_lds_bytes(0, n, 1024) <= sidon_lds_limit()).
 * Replaces runLocalSearch (src/main.cu:152-159) / kLocalSearch<float>
 * (src/kernels.cuh:177-289). */
int sidon_local_search(float* pop, float* mom, float* vel, const int* skip, const int* mode,
                       int n, int P, int iters, float pStart, float pEnd, float lrStart,
                       float lrEnd, float beta1, float beta2, int modeDefault,
                       hipStream_t stream);
/* fp64 form: kLocalSearch<double> (src/main.cu:837-839).  Above n = 4544 the
 * correlation output goes to the rows of mom (else vel; of both when both
 * are given), which must then be given; the longest length is 6480. */
int sidon_local_search_f64(double* pop, double* mom, double* vel, const int* skip,
                           const int* mode, int n, int P, int iters, double pStart,
                           double pEnd, double lrStart, double lrEnd, double beta1,
                           double beta2, int modeDefault, hipStream_t stream);

/* fp64 local search of K candidates of any length n, spread over the whole
 * device (the reference's fp64 polish, src/main.cu:798-846, whose fused
 * kernel needs ~3n doubles of on-chip memory: n = 8400 does not fit one
 * MI355X workgroup).  Same step as sidon_local_search_f64 (mode 0 projected
 * Adam, 1 mirror descent; p, lr annealed geometrically over `iters`; moments
 * start at zero; mom/vel, may be NULL, receive the final moments), with both
 * products cut into work units for every CU and summed in a fixed order
 * (deterministic).  `workspace` is a device buffer of at least
 * sidon_polish_f64_workspace(n, K) bytes (about 3n^2/1024 + 7n doubles per
 * candidate: 17 MB for K = 8 at n = 8400); it holds no state between calls.
 * pop is [K x n] row-major. */
size_t sidon_polish_f64_workspace(int n, int K);
int sidon_polish_f64(double* pop, double* mom, double* vel, int n, int K, int iters,
                     double pStart, double pEnd, double lrStart, double lrEnd, double beta1,
                     double beta2, int mode, void* workspace, size_t workspace_bytes,
                     hipStream_t stream);
/* sidon_polish_f64 with no workspace argument and no allocation: the
 * workspace is a static device array of 32 MiB (33554432 bytes) that each GPU
 * the library runs on holds until the process ends, resolved once per device
 * by sidon_init.  Same kernels and results as sidon_polish_f64.  Returns
 * SIDON_ERR_ARG when sidon_polish_f64_workspace(n, K) exceeds 32 MiB (K <= 15
 * at n = 8400).  Every call on one GPU shares that array: calls that overlap
 * on one GPU (two streams, two threads) race; serialise them.  No counterpart
 * in the reference: for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc). */
int sidon_polish_f64_noalloc(double* pop, double* mom, double* vel, int n, int K, int iters,
                             double pStart, double pEnd, double lrStart, double lrEnd,
                             double beta1, double beta2, int mode, hipStream_t stream);
/* Uneven lengths in one persistent launch: candidate p occupies
 * pop[p*stride .. p*stride + ns[p]) (ns on the device, ns[p] <= nmax <= stride);
 * `order` (may be NULL) lists the processing order, longest first gives the
 * best balance.  As many workgroups as fit the device at once (occupancy x
 * CUs, at most P) pull tickets from a work queue.  With split = 1 and an
 * order, the longest candidates (n >= 4096) may each be shared by two
 * workgroups that exchange b and the new heights through global memory
 * twice a step: as many as a makespan model of the launch says shortens it
 * (none when every workgroup has a candidate of the longest length to run);
 * the result is bit-identical to one workgroup's.  `workspace` (caller-owned, device,
 * sidon_local_search_ragged_workspace(P, stride) bytes) holds the queue
 * counter, the pairs' progress flags and exchange buffers, and two rows per
 * candidate for its correlation output (the launch keeps it there, as
 * sidon_local_search does in the moments' rows; nmax up to 13216); the
 * launch clears
 * what must start at zero with a kernel on `stream` (no memset node).
 * Launches that may overlap need distinct workspaces.  After the launch the
 * int at byte 4 of the workspace counts pair waits that gave up (a partner
 * workgroup that never ran within 20 s; 0 in any normal run) and the int at
 * byte 8 holds the number of candidates split.  No reference counterpart
 * (the reference runs one n per launch). */
size_t sidon_local_search_ragged_workspace(int P, int stride);
int sidon_local_search_ragged(float* pop, int stride, const int* ns, int P, int nmax,
                              const int* order, const int* skip, const int* mode, int iters,
                              float pStart, float pEnd, float lrStart, float lrEnd,
                              float beta1, float beta2, int modeDefault, int split,
                              void* workspace, size_t workspace_bytes, hipStream_t stream);
/* Choose the launch configuration the fp32 local search uses at length n:
 * workgroup size `block` (256, 512 or 1024 threads: 4, 8 or 16 waves per
 * candidate), `depth` operand steps in flight in the matrix-core tile loops
 * (2 or 4) and `per_cu` candidates a CU (2 only with block 512 and
 * n <= 4096, 4 only with block 256 and n <= 2304: workgroups held to 128
 * registers a lane share a CU; else 1).  block 0 restores the built-in rule
 * (256/2/4 below n = 1024, 512/4/2 up to 4096, 1024/4/1 above).  Set by the
 * autotuner.  The rule is
 * process-wide host state, read when a launch is enqueued (a captured graph
 * keeps the configuration it was captured with). */
int sidon_set_ls_tile(int n, int block, int depth, int per_cu);
int sidon_get_ls_tile(int n, int* block, int* depth, int* per_cu);
/* The workgroup-size-only forms: set_ls_block takes the built-in
 * configuration when it has that workgroup size, else depth 4 with the
 * CU-sharing form (per_cu 2 or 4) where the built-in rule's lengths allow
 * it. */
int sidon_set_ls_block(int n, int block);
int sidon_get_ls_block(int n);

/* ---- gradient of the smoothed objective ------------------------------ */

/* J_p(a) = 2n ||a*a||_p / (sum a)^2 and its true gradient
 * dJ/da_i = (4 n R^{1/p}/S^2)(G_i/R - M/S), recomputing b (never stored).
 * a is [B x n]; J (length B) and grad ([B x n]) may be NULL independently.
 * Above the on-chip form the correlation accumulates in grad itself
 * (fp32 to n = 13216).
 * No reference counterpart (kLocalSearch forms only the scaled direction,
 * src/kernels.cuh:165-172, :244). */
int sidon_lp_grad(const float* a, int n, int B, float p, float* J, float* grad,
                  hipStream_t stream);
int sidon_lp_grad_f64(const double* a, int n, int B, double p, double* J, double* grad,
                      hipStream_t stream);

/* ---- Frank-Wolfe step solved by Chambolle-Pock ------------------------ */

/* `rounds` Frank-Wolfe rounds per candidate: normalise, M = max a*a, solve
 * max sum g s.t. (a*g)_k <= M, g >= 0 by `cpIters` PDHG iterations
 * (tau = 0.9/n, sigma = 0.9n; conv + dual update, corr + primal update,
 * extrapolation), normalise g, grid line search over t = tMax*s/nT,
 * accept the best strict improvement.  Replaces runFrankWolfe
 * (src/main.cu:170-178) / kFrankWolfe<float> (src/kernels.cuh:408-532);
 * returns SIDON_ERR_LDS where the reference returns false. */
int sidon_frank_wolfe(float* pop, const int* skip, int n, int P, int rounds, int cpIters,
                      int nT, float tMax, hipStream_t stream);

/* ---- cell split, variation, selection --------------------------------- */

/* n -> 2n cell split dst[p][2i] = dst[p][2i+1] = src[p][i]; F is preserved.
 * Replaces kRefine (src/kernels.cuh:538-547) and the host refine loop
 * (src/main.cu:491-502). */
int sidon_refine(const float* src, int n, int P, float* dst, hipStream_t stream);

/* Offspring j = operator op[j] applied to parents p1[j] and p2[j] (p2 may be
 * NULL: p2 = p1), the reference's fourteen operators (src/kernels.cuh:614-787):
 * 0 blend, 1 segment crossover, 2 log-normal noise exp(strength[j] z),
 * 3 smooth Fourier perturbation, 4 additive noise, 5 sparsify, 6 symmetrise /
 * reflect, 7 shift, 8 block splice, 9 random restart, 10 spike, 11 structured
 * seed, 12 gap surgery, 13 endpoint atom; any other op clones.  Each height is
 * formed in fp64 and rounded once to fp32; normals z come from Philox4x32-10
 * keyed by (seed, gen, j, i), the 64 operator parameters from Philox keyed by
 * (seed, gen, j).  Then renormalise sum = 1 (uniform 1/n if the sum is not
 * positive).  Replaces kVariation<float> (src/kernels.cuh:582-801, launch
 * src/main.cu:737-742). */
int sidon_variation(const float* pop, float* out, int n, int nOff, const int* p1,
                    const int* p2, const int* op, const float* strength,
                    unsigned long long seed, unsigned long long gen, hipStream_t stream);
/* Uneven-length form: parent p1[j] has ns[p1[j]] heights at pop + p1[j]*stride;
 * the offspring inherits that length (ns_out[j]).  `gen` is read from device
 * memory (*gen_dev) so the launch can be replayed from a graph. */
int sidon_variation_ragged(const float* pop, const int* ns, float* out, int* ns_out,
                           int stride, int nOff, const int* p1, const int* p2, const int* op,
                           const float* strength, unsigned long long seed,
                           const unsigned long long* gen_dev, hipStream_t stream);

/* Support signature of each row (P rows of n heights): the support is
 * {i : a_i > mean * tol} with the fp32 mean summed in the reference kernel's
 * order; sig[p] = XOR over the support of a 64-bit mix of i, nnz[p] = its
 * size.  Replaces kSupportSig<float> (src/kernels.cuh:803-849, launch
 * src/main.cu:180-184, tol 0.02), the key of the tabu and cell-best tables. */
int sidon_support_sig(const float* pop, int n, int P, float tol, unsigned long long* sig,
                      int* nnz, hipStream_t stream);

/* Island selection plan on the device: rank each island (P/NI members) by
 * fit (ties by index); every migrate_every generations (gen > 0) the best of
 * island i replaces the worst of island i+1 (vector, length and fit; pass
 * pop/ns/stride, ns may be NULL for uniform n); the first `elite` slots of each
 * island clone its ranked elites (op 20); the others take two binary
 * tournaments inside the island (op 2, strength sigma) drawn from Philox keyed
 * by (seed, gen, slot).  `gen` is read from *gen_dev.  fit is updated in place
 * for migrated slots.  Replaces the host selection of src/main.cu:650-735
 * (the operator is fixed to 2, as the configurations specify). */
int sidon_select(double* fit, float* pop, int* ns, int stride, int P, int NI, int elite,
                 int migrate_every, float sigma, unsigned long long seed,
                 const unsigned long long* gen_dev, int* p1, int* p2, int* op,
                 float* strength, hipStream_t stream);

/* Adds `delta` to the device generation counter *gen_dev (one-thread kernel,
 * graph-safe). */
int sidon_advance_gen(unsigned long long* gen_dev, unsigned long long delta,
                      hipStream_t stream);

/* Orders candidates by decreasing length (ties by index) for the ragged
 * local search: order[0..P) on the device. */
int sidon_order_by_length(const int* ns, int P, int* order, hipStream_t stream);

/* ---- one generation as a HIP graph ------------------------------------ */

/* Device buffers (caller-owned) and parameters of a generation of the island
 * search.  Populations are [P x stride] fp32 with per-slot lengths ns (<= nmax);
 * a generation reads population A (ns_a) and writes population B (ns_b), the
 * next one the other way round.  gen points to a device uint64 counter;
 * ls_workspace is the local search's workspace
 * (sidon_local_search_ragged_workspace(P, stride) bytes; one per generation
 * object that may run concurrently with another); split as in
 * sidon_local_search_ragged. */
typedef struct {
  float* pop_a; float* pop_b;
  int* ns_a; int* ns_b;
  double* fit; int* nact;
  int* p1; int* p2; int* op; float* strength;
  int* order;
  unsigned long long* gen;
  void* ls_workspace;
  unsigned long long ls_workspace_bytes;
  int P, stride, nmax, islands, elite, migrate_every, iters, split;
  float pStart, pEnd, lrStart, lrEnd, sigma;
  double activeTol;
  unsigned long long seed;
} sidon_generation_desc;

typedef struct sidon_generation sidon_generation;

/* Captures the two generation graphs (A -> B and B -> A):
 * order by length -> sidon_local_search_ragged (iters fused steps) ->
 * sidon_eval_f64_ragged -> sidon_select -> sidon_variation_ragged ->
 * sidon_advance_gen.  Allocates only the graph objects.  Replaces the host
 * generation loop body of src/main.cu:586-785, whose ~10 blocking transfers
 * per generation kept it off the device. */
int sidon_generation_create(const sidon_generation_desc* desc, sidon_generation** out,
                            hipStream_t stream);
/* Replays `generations` generations on `stream`, alternating the graphs. */
int sidon_generation_run(sidon_generation* g, int generations, hipStream_t stream);
/* Which population holds the current generation: 0 = A, 1 = B. */
int sidon_generation_current(const sidon_generation* g, int* parity);
/* The same sequence enqueued eagerly (no graph), reading A (parity 0) or B. */
int sidon_generation_enqueue(const sidon_generation_desc* desc, int parity, hipStream_t stream);
int sidon_generation_destroy(sidon_generation* g);

/* ---- engine ceilings ------------------------------------------------------ */

/* Measures, on the current device and under this library's compile flags, the
 * FP32 FMA rate of the vector ALUs (packed FP32 off) and of the matrix cores
 * (v_mfma_f32_16x16x4f32), in TFLOP/s, with register-only kernels on every CU.
 * `scratch` is a device buffer of at least (CUs * 8 * 256) floats.  The local
 * search runs its products on the matrix cores because the measured matrix
 * ceiling is above the vector one.  No reference counterpart. */
int sidon_measure_fp32_ceilings(float* scratch, int scratch_len, double* valu_tflops,
                                double* mfma_tflops, hipStream_t stream);

/* Matrix block steps (four v_mfma_f32_16x16x4f32 per wave: 8192 FLOP per
 * lane-group of 64, i.e. 4 * 2048 FLOP) that the local-search kernels and
 * sidon_lp_grad[_f64] have executed on the current device since the last call with reset != 0; the
 * count excludes correlation steps skipped because their weights are all
 * zero.  The counter is device-wide: launches on every stream add to it, and
 * the read (with reset, an atomic exchange with zero on the device) loses no
 * step that a concurrently running launch adds.  Synchronous.  For roofline
 * accounting; no reference counterpart. */
int sidon_ls_block_steps(unsigned long long* steps, int reset);
/* Profiling builds of the library only (`make clocks`: libsidon_amd_clocks.so,
 * the local search compiled with SIDON_PHASE_CLOCKS); the product library
 * returns SIDON_ERR_UNSUPPORTED.  out[8]: shader-clock cycles of the local
 * search's step phases as its first wave sees them, barrier to barrier,
 * summed over steps and workgroups ([0] autoconvolution with finish and max,
 * [1] weights and R, [2] correlation, [3] update, S and renormalisation), the
 * wave-cycles spent between a wave's last work item and the end of the
 * autoconvolution [4] / correlation [5] barrier, then the steps [6] and
 * candidates [7] counted.  Synchronous; reset != 0 zeroes the sums.  No
 * reference counterpart. */
int sidon_ls_phase_clocks(unsigned long long* out, int reset);
/* The same count for sidon_frank_wolfe: Toeplitz steps of the dual updates,
 * of the M and line-search autoconvolutions, and the correlation steps run
 * (steps over all-zero multipliers are skipped).  No reference counterpart. */
int sidon_fw_block_steps(unsigned long long* steps, int reset);

/* ---- autotuner ---------------------------------------------------------- */

/* The configurations the autotuner tries, i = 0 .. SIDON_LS_TILES-1:
 * 256/2/1, 256/4/1, 256/2/4, 256/4/4, 512/2/1, 512/4/1, 512/2/2, 512/4/2,
 * 1024/2/1, 1024/4/1 (block/depth/per_cu as in sidon_set_ls_tile). */
#define SIDON_LS_TILES 10
int sidon_ls_tile_config(int i, int* block, int* depth, int* per_cu);

/* Times the fused local search at length n in every configuration above on
 * `scratch` (P x n, overwritten; reset from `reference` before every trial
 * when non-NULL), `repeats` timed runs of `iters` steps after one warm-up,
 * and keeps the fastest as the rule for n (sidon_get_ls_tile); the built-in
 * configuration stays unless another beats it by more than 1%.  `rows`
 * (2 x P x n floats, may be NULL) is passed to the trials as their moments'
 * buffers (mom, then vel), so they time the form that launches with both
 * run (the correlation output in their rows); NULL times the form without, and
 * above n = 9296, which has no other form, reports the one configuration
 * untimed.  Reports the choice and its time, and, when `trial_ms` is
 * non-NULL, the time of every configuration (SIDON_LS_TILES floats, -1
 * where it does not fit n).  No reference counterpart (tile sizes are
 * compile-time constants there, src/kernels.cuh:30-33). */
int sidon_tune_local_search(float* scratch, const float* reference, float* rows, int n, int P,
                            int iters, int repeats, int* best_block, int* best_depth,
                            int* best_per_cu, float* best_ms, float* trial_ms,
                            hipStream_t stream);

/* ---- island search policy (host only) ---------------------------------- */

/* The island search policy of the reference's driver (src/main.cu:461-472,
 * :598-609, :650-769) as a host-side object: it decides, the driver moves the
 * vectors.  No function here touches the device; all arrays are HOST arrays
 * of length P.  Operators 0..13 are sidon_variation's, 20 = clone. */
#define SIDON_POLICY_OPS 14
typedef struct sidon_policy sidon_policy;

/* P candidates in `islands` equal islands, `elite` clones per island;
 * random draws from a std::mt19937_64 seeded with `seed`.  Operator weights
 * start at 1 except the structured seed, gap and atom (2.5, 1.8, 1.8). */
int sidon_policy_create(int P, int islands, int elite, unsigned long long seed,
                        sidon_policy** out);
void sidon_policy_destroy(sidon_policy* pol);
/* Adopts an incumbent found outside the population (a resumed vector, its F
 * by hostF): the tabu test compares cells against it (main.cu:414-428). */
int sidon_policy_adopt(sidon_policy* pol, double f);
/* Starts a ladder level: the level best and the generations-since-improvement
 * count restart. */
int sidon_policy_new_level(sidon_policy* pol);
/* One generation's fp64 fitness and support signatures: tracks the best
 * (*improved = 1 when it beats the incumbent by more than 1e-13, *best_slot
 * = first minimum) and the cell tables (a signature seen again without
 * improving its cell's best by 1e-12 gains a tabu hit). */
int sidon_policy_observe(sidon_policy* pol, const double* fit, const unsigned long long* sig,
                         int* best_slot, int* improved);
/* The offspring plan of generation `gen` (counted from 0 within a level):
 * islands ranked by (fit, slot); when gen > 0 and gen % 12 == 0 the best of
 * island i replaces the worst of island i+1 (mig_src[k] -> mig_dst[k],
 * k < *n_mig <= islands: the caller copies those vectors before the
 * variation); the first `elite` slots of each island clone its ranked elites
 * (op 20, strength 0); every other slot draws two binary tournaments in its
 * island, an operator from the weights, a diversifying operator instead when
 * the first parent's cell has more than 6 tabu hits and its best is worse
 * than best*(1+2e-4), a restart (op 9) with probability 0.1 once stagnant
 * (more than 17 generations without a level improvement), and strength
 * (0.03 + 0.35 U)(0.5 + min(1, gen/40))(1 + 1.5 min(1, stagnant/25)).
 * *reseed_slot (may be NULL) = the slot where the caller puts the incumbent
 * back (gen % 10 == 9), else -1. */
int sidon_policy_plan(sidon_policy* pol, int gen, int* p1, int* p2, int* op, float* strength,
                      int* mig_src, int* mig_dst, int* n_mig, int* reseed_slot);
/* Credit assignment (the caller runs it every 5th generation, gen > 0): an
 * operator wins when its child's fitness beats its first parent's; weights
 * move to 0.85 w + 0.15 (0.15 + 3 win rate) (rate 0.25 below 20 uses),
 * clamped to [0.05, 4]; use and win counts decay by 0.7. */
int sidon_policy_credit(sidon_policy* pol, const double* child_fit);
/* Operator weights (SIDON_POLICY_OPS), incumbent and level best, cell count,
 * generations since the level improved; any output may be NULL. */
int sidon_policy_stats(const sidon_policy* pol, double* weights, double* best_f,
                       double* level_best, int* cells, int* since_improved);
/* Anneal tier of each candidate from the operator that made it: 0 cold
 * (restart or structured seed, at most 18 % of P), 1 medium (structure
 * changed: ops 5, 7, 8, 9, 10, 11, 12, 13), 2 warm (everything else). */
int sidon_policy_tiers(const int* cur_op, int P, int* tier, int* n_cold, int* n_med);

#ifdef __cplusplus
}
#endif
#endif /* SIDON_H */
