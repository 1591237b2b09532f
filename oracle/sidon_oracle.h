/*
 * sidon_oracle.h -- CPU oracle for the Sidon-autocorrelation memetic-search path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker that the HIP library
 * is compared against.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it; the product path never calls it and never
 * falls back to it.
 *
 * Every function restates one piece of the reference's algorithm and cites the
 * reference file:line it follows.  All functions are plain C, OpenMP where the
 * reference uses OpenMP, and take plain pointers + sizes.
 *
 * Objective: F(a) = 2n * max_k b_k / (sum_i a_i)^2,  b_k = sum_{i+j=k} a_i a_j.
 */
#ifndef SIDON_ORACLE_H
#define SIDON_ORACLE_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

/* fp64 F in hostF order: i-major scatter, zero rows skipped, sequential S.
 * Follows src/main.cu:94-105 (hostF) and tools/probe.cpp:30-50 (objective). */
double so_F_hostorder(const double* a, int n);

/* fp64 F with a per-k ascending-i gather and a sequential S; also M and argmax k.
 * Follows verify.py:62-78 (ratio). */
double so_F_gather(const double* a, int n, double* M, int* kmax);

/* Long-double F with OpenMP over k, per-k ascending-i gather, sequential M/S,
 * active-peak count at 1e-9 and the 3*n*eps/(1-n*eps) round-off bound.
 * Follows tools/probe.cpp:163-183.  Returns 1 if any weight is negative
 * (probe.cpp:161, :194), else 0. */
int so_verify_ld(const long double* a, int n, long double* F, long double* M,
                 long double* S, int* kmax, int* nActive, long double* relBound);

/* Same, reading a saved vector text file ("# ..." comments, "n N", "F x", one
 * height per line) with strtold, as tools/probe.cpp:146-160 does.  Returns
 * 0 ok, 1 negative weight, 2 unreadable, 3 length mismatch. */
int so_verify_file(const char* path, long double* F, long double* M, long double* S,
                   int* kmax, int* nActive, long double* relBound, int* n_out);

/* so_verify_file's figures formatted as tools/probe.cpp:185-193 prints them:
 * "S M kmax nActive F relBound F_upper". */
int so_verify_file_str(const char* path, char* buf, int len);

/* Correctly rounded b_k (error-free TwoProd + exact Shewchuk/fsum summation)
 * and correctly rounded S; F = 2*n*M/(S*S) from those.  With correctly rounded
 * b and S the n -> 2n cell split preserves F bit for bit (b' = 2b, S' = 2S
 * exactly, and fl(2x) = 2 fl(x)).  b_out (length 2n-1) may be NULL. */
double so_F_exact(const double* a, int n, double* M, double* S, int* kmax, double* b_out);

/* Correctly rounded sum of m doubles (Shewchuk's non-overlapping partials with
 * the round-half-even fix-up; the algorithm Python's math.fsum publishes). */
double so_fsum(const double* x, int m);

/* Autoconvolution b_k, per-k ascending-i order (kernels.cuh:97-135 order). */
void so_autoconv(const double* a, int n, double* b);
/* Correlation G_i = sum_j c_{i+j} a_j, ascending j (kernels.cuh:142-158). */
void so_corr(const double* c, const double* a, int n, double* G);

/* L^p weights in the stable log1p form (kernels.cuh:35-44, :221-230):
 * d = (b_k - M)/M, c_k = exp((p-1) log1p(d)) if d > -1 else 0,
 * R = sum_k c_k (1+d) (sequential).  c may alias b.  Returns R. */
double so_lp_weights(const double* b, int nb, double M, double p, double* c);

/* Smoothed objective J_p(a) = 2n ||b||_p / (sum a)^2 and its true gradient
 *   dJ/da_i = (4 n R^{1/p} / S^2) (G_i / R - M / S),  G = corr(u^{p-1}, a).
 * (No reference kernel: the reference only forms the scaled direction
 * G_i/R - M of kernels.cuh:165-172, :244.)  grad may be NULL. */
double so_Jp_grad(const double* a, int n, double p, double* grad);
long double so_Jp_grad_ld(const long double* a, int n, long double p, long double* grad);

/* fp64 L^p-annealed projected Adam / mirror descent.
 * probe_semantics = 0: the device kernel's semantics (kernels.cuh:196-288):
 *   no entry normalisation, renormalise by multiplying with 1/S.
 * probe_semantics = 1: the CPU twin (tools/probe.cpp:61-116): normalise at
 *   entry, renormalise by dividing by S, return objective() at the end.
 * m_out/v_out (may be NULL) receive the final Adam moments.
 * Returns the fp64 F of the final vector (hostF order), and the number of
 * completed iterations in *iters_done (M <= 0 stops early, kernels.cuh:219). */
double so_local_search(double* a, int n, int iters, double pA, double pB,
                       double lrA, double lrB, double b1, double b2, int mode,
                       int probe_semantics, double* m_out, double* v_out, int* iters_done);

/* The device-kernel semantics of so_local_search (probe_semantics = 0) that
 * also records, per height i, min over the steps of |g_i| / M, g_i = G_i/R - M
 * the direction (gmin, length n).  A height whose direction came near zero
 * may take either Adam sign under another summation order; parity tests
 * exclude those by this record. */
double so_local_search_gmin(double* a, int n, int iters, double pA, double pB, double lrA,
                            double lrB, double b1, double b2, int mode, double* m_out,
                            double* v_out, double* gmin);

/* Batched form of the above over P candidates (row-major P x n), OpenMP over
 * candidates with nthreads threads (0 = OpenMP default).  Used as the timed
 * CPU baseline.  F_out (length P) may be NULL. */
void so_local_search_batch(double* pop, int P, int n, int iters, double pA, double pB,
                           double lrA, double lrB, double b1, double b2, int mode,
                           int probe_semantics, int nthreads, double* F_out);

/* Frank-Wolfe round(s) with the LP direction solved by Chambolle-Pock (PDHG),
 * then a grid line search; fp64.  Follows kernels.cuh:408-532 step by step.
 * Returns the number of accepted rounds.  M_out (may be NULL) receives the
 * final max_k (a*a)_k with sum a = 1. */
int so_frank_wolfe(double* a, int n, int rounds, int cpIters, int nT, double tMax,
                   double* M_out);

/* so_frank_wolfe that also reports, for the last round that reached its line
 * search, the normalised LP direction g (g_out, length n) and max_k (h*h)_k of
 * each trial h = (1-t) a + t g, t = tMax s/nT, s = 1..nT (lsM_out, length nT).
 * Either may be NULL.  With them a test can tell a near-tie of the line
 * search (two t within rounding of each other) from a wrong step. */
int so_frank_wolfe_trace(double* a, int n, int rounds, int cpIters, int nT, double tMax,
                         double* M_out, double* g_out, double* lsM_out);

/* One PDHG solve for the LP of a fixed a (sum a = 1, M = max a*a):
 * returns g (not normalised) after cpIters iterations (kernels.cuh:451-503). */
void so_pdhg(const double* a, int n, double M, int cpIters, double* g, double* lam);

/* n -> 2n cell split a'_{2i} = a'_{2i+1} = a_i (kernels.cuh:534-547). */
void so_refine(const double* src, int n, double* dst);

/* Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11). */
void so_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);

/* The build's own mutation stream (op 2, kernels.cuh:629-636 + renorm
 * :791-800): O_i = A_i * exp(sigma z_i), z from Philox4x32-10 + Box-Muller in
 * fp32, then clamp >= 0 and renormalise sum = 1 (uniform 1/n if the sum is not
 * positive).  op 20 (and any op other than 2) is the clone branch
 * (kernels.cuh:784-786) followed by the same renormalisation. */
void so_variation(const float* A, int n, int op, float sigma, uint64_t seed, uint64_t gen,
                  int slot, float* out);
/* The standard normals the variation stream uses for element i of slot j. */
float so_normal(uint64_t seed, uint64_t gen, int slot, int i);
/* All fourteen operators (kernels.cuh:614-787) on parents A, B (B = NULL:
 * B = A), each height in fp64 rounded once to fp32, then the renormalisation
 * above; operator parameters from so_op_params. */
void so_variation2(const float* A, const float* B, int n, int op, float sigma, uint64_t seed,
                   uint64_t gen, int slot, float* out);
void so_op_params(uint64_t seed, uint64_t gen, int slot, float prm[64]);
int so_op_uses_mean(int op);
/* kSupportSig (kernels.cuh:803-849): support {a_i > mean * tol} with the
 * kernel's fp32 mean, 64-bit XOR signature and support size. */
void so_support_sig(const float* a, int n, float tol, uint64_t* sig, int* nnz);
uint64_t so_sig_mix(int i);

/* Island selection plan (main.cu:676-735 with the uneven-islands workload's fixed operator):
 * per island rank by fp64 fit (ties by index), ring migration every
 * migrate_every generations (best of island i replaces the worst of island
 * i+1; fit and the slot's vector follow), elites cloned (op 20), the rest
 * filled by two binary tournaments (op 2, strength sigma).
 * fit (length P) is updated in place for migrated slots.  mig_src/mig_dst
 * (length NI, may be NULL) receive the migration copies (-1 when none). */
void so_select_plan(double* fit, int P, int NI, int elite, uint64_t seed, uint64_t gen,
                    int migrate_every, float sigma, int* p1, int* p2, int* op,
                    float* strength, int* mig_src, int* mig_dst);

/* Thread count OpenMP would use (1 when built without OpenMP). */
int so_max_threads(void);

#ifdef __cplusplus
}
#endif
#endif
