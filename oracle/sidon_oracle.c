/*
 * sidon_oracle.c -- CPU oracle (checker) for the Sidon memetic-search path.
 *
 * TEST INFRASTRUCTURE ONLY: loaded by tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py, never by the product path.  Each function
 * names the reference file:line whose semantics it restates.  Reference paths
 * are relative to the reference repository root (src/, tools/, verify.py).
 */
#include "sidon_oracle.h"

#include <float.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

int so_max_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

/* ------------------------------------------------------------------ F ---- */

/* src/main.cu:94-105 (hostF): b[i+j] += a_i a_j for i ascending (zero rows
 * skipped), M = max b, S sequential, F = 2*n*M/(S*S). */
double so_F_hostorder(const double* a, int n) {
  const int nb = 2 * n - 1;
  double* b = (double*)calloc((size_t)nb, sizeof(double));
  for (int i = 0; i < n; ++i) {
    const double ai = a[i];
    if (ai == 0.0) continue;
    double* bp = b + i;
    for (int j = 0; j < n; ++j) bp[j] += ai * a[j];
  }
  double M = 0.0;
  for (int k = 0; k < nb; ++k) M = b[k] > M ? b[k] : M;
  double S = 0.0;
  for (int i = 0; i < n; ++i) S += a[i];
  free(b);
  return 2.0 * n * M / (S * S);
}

/* verify.py:62-78 (ratio): per-k ascending-i gather, strict '>' argmax,
 * sequential sum. */
double so_F_gather(const double* a, int n, double* Mo, int* kmax) {
  double best = -1.0;
  int bk = -1;
  for (int k = 0; k < 2 * n - 1; ++k) {
    const int lo = k - n + 1 > 0 ? k - n + 1 : 0;
    const int hi = k < n - 1 ? k : n - 1;
    double b = 0.0;
    for (int i = lo; i <= hi; ++i) b += a[i] * a[k - i];
    if (b > best) { best = b; bk = k; }
  }
  double S = 0.0;
  for (int i = 0; i < n; ++i) S += a[i];
  if (Mo) *Mo = best;
  if (kmax) *kmax = bk;
  return 2 * n * best / (S * S);
}

/* tools/probe.cpp:161-183: long-double gather parallel over k (static
 * schedule), sequential M (strict '>'), sequential S, active count at 1e-9,
 * bound 3 * n*eps/(1-n*eps) with eps = 2^-63. */
int so_verify_ld(const long double* a, int n, long double* Fo, long double* Mo,
                 long double* So, int* kmaxo, int* nActiveo, long double* relBoundo) {
  int neg = 0;
  for (int i = 0; i < n; ++i)
    if (a[i] < 0) neg = 1;
  const int nb = 2 * n - 1;
  long double* b = (long double*)malloc(sizeof(long double) * (size_t)nb);
#pragma omp parallel for schedule(static)
  for (int k = 0; k < nb; ++k) {
    const int ilo = k - n + 1 > 0 ? k - n + 1 : 0;
    const int ihi = k < n - 1 ? k : n - 1;
    long double s = 0.0L;
    for (int i = ilo; i <= ihi; ++i) s += a[i] * a[k - i];
    b[k] = s;
  }
  long double M = 0, S = 0;
  int kmax = 0;
  for (int k = 0; k < nb; ++k)
    if (b[k] > M) { M = b[k]; kmax = k; }
  for (int i = 0; i < n; ++i) S += a[i];
  const long double F = 2.0L * (long double)n * M / (S * S);
  const long double eps = 1.084202172485504434e-19L;
  const long double gam = (long double)n * eps / (1 - (long double)n * eps);
  int nA = 0;
  for (int k = 0; k < nb; ++k)
    if (b[k] >= M * (1 - 1e-9L)) ++nA;
  free(b);
  if (Fo) *Fo = F;
  if (Mo) *Mo = M;
  if (So) *So = S;
  if (kmaxo) *kmaxo = kmax;
  if (nActiveo) *nActiveo = nA;
  if (relBoundo) *relBoundo = 3 * gam;
  return neg ? 1 : 0;
}

/* tools/probe.cpp:146-160: parse the saved-vector text format with strtold. */
int so_verify_file(const char* path, long double* F, long double* M, long double* S,
                   int* kmax, int* nActive, long double* relBound, int* n_out) {
  FILE* in = fopen(path, "r");
  if (!in) return 2;
  size_t cap = 1024, len = 0;
  long double* a = (long double*)malloc(cap * sizeof(long double));
  char line[512];
  int nDecl = 0;
  while (fgets(line, sizeof line, in)) {
    if (line[0] == '\0' || line[0] == '#' || line[0] == '\n' || line[0] == '\r') continue;
    if (line[0] == 'n' && line[1] == ' ') { nDecl = atoi(line + 2); continue; }
    if (line[0] == 'F' && line[1] == ' ') continue;
    if (len == cap) { cap *= 2; a = (long double*)realloc(a, cap * sizeof(long double)); }
    a[len++] = strtold(line, NULL);
  }
  fclose(in);
  const int n = (int)len;
  if (n_out) *n_out = n;
  if (nDecl && nDecl != n) { free(a); return 3; }
  const int rc = so_verify_ld(a, n, F, M, S, kmax, nActive, relBound);
  free(a);
  return rc;
}

/* The figures tools/probe.cpp:185-193 prints, in its formats, as one line:
 * "S M kmax nActive F relBound F_upper" (%.20Lg, %d, %.3Lg). */
int so_verify_file_str(const char* path, char* buf, int len) {
  long double F, M, S, rb;
  int kmax, nA, n;
  const int rc = so_verify_file(path, &F, &M, &S, &kmax, &nA, &rb, &n);
  snprintf(buf, (size_t)len, "%.20Lg %.20Lg %d %d %.20Lg %.3Lg %.20Lg", S, M, kmax, nA, F, rb,
           F * (1 + rb));
  return rc;
}

/* ----------------------------------------------------- exact summation ---- */

/* Correctly rounded sum of doubles: non-overlapping partials (Shewchuk, "Adaptive
 * Precision Floating-Point Arithmetic", 1997) plus the final half-even
 * adjustment, the method Python's math.fsum documents.  Finite inputs only. */
typedef struct { double* p; int len, cap; } partials_t;

static void fs_add(partials_t* P, double x) {
  int i = 0;
  for (int j = 0; j < P->len; ++j) {
    double y = P->p[j];
    if (fabs(x) < fabs(y)) { double t = x; x = y; y = t; }
    const double hi = x + y;
    const double lo = y - (hi - x);
    if (lo != 0.0) P->p[i++] = lo;
    x = hi;
  }
  if (i >= P->cap) { P->cap = 2 * P->cap + 8; P->p = (double*)realloc(P->p, sizeof(double) * (size_t)P->cap); }
  P->p[i++] = x;
  P->len = i;
}

static double fs_result(const partials_t* P) {
  int n = P->len;
  double hi = 0.0, lo = 0.0;
  if (n > 0) {
    hi = P->p[--n];
    while (n > 0) {
      const double x = hi;
      const double y = P->p[--n];
      hi = x + y;
      const double yr = hi - x;
      lo = y - yr;
      if (lo != 0.0) break;
    }
    if (n > 0 && ((lo < 0.0 && P->p[n - 1] < 0.0) || (lo > 0.0 && P->p[n - 1] > 0.0))) {
      const double y = lo * 2.0;
      const double x = hi + y;
      const double yr = x - hi;
      if (y == yr) hi = x;
    }
  }
  return hi;
}

double so_fsum(const double* x, int m) {
  partials_t P = {NULL, 0, 0};
  for (int i = 0; i < m; ++i) fs_add(&P, x[i]);
  const double r = fs_result(&P);
  free(P.p);
  return r;
}

/* Correctly rounded b_k and S.  a_i a_j = p + e exactly (TwoProd via fma). */
double so_F_exact(const double* a, int n, double* Mo, double* So, int* kmaxo, double* b_out) {
  const int nb = 2 * n - 1;
  double* b = b_out ? b_out : (double*)malloc(sizeof(double) * (size_t)nb);
#pragma omp parallel
  {
    partials_t P = {NULL, 0, 0};
#pragma omp for schedule(dynamic, 16)
    for (int k = 0; k < nb; ++k) {
      const int ilo = k - n + 1 > 0 ? k - n + 1 : 0;
      const int ihi = k < n - 1 ? k : n - 1;
      P.len = 0;
      for (int i = ilo; i <= ihi; ++i) {
        const double p = a[i] * a[k - i];
        const double e = fma(a[i], a[k - i], -p);
        fs_add(&P, p);
        if (e != 0.0) fs_add(&P, e);
      }
      b[k] = fs_result(&P);
    }
    free(P.p);
  }
  double M = 0.0;
  int kmax = 0;
  for (int k = 0; k < nb; ++k)
    if (b[k] > M) { M = b[k]; kmax = k; }
  const double S = so_fsum(a, n);
  if (!b_out) free(b);
  if (Mo) *Mo = M;
  if (So) *So = S;
  if (kmaxo) *kmaxo = kmax;
  return 2.0 * n * M / (S * S);
}

/* ------------------------------------------------ conv / corr / weights ---- */

/* b_k = sum_{i=max(0,k-n+1)}^{min(k,n-1)} a_i a_{k-i}, ascending i
 * (kernels.cuh:97-135 accumulates each k over ascending i). */
void so_autoconv(const double* a, int n, double* b) {
  const int nb = 2 * n - 1;
  for (int k = 0; k < nb; ++k) {
    const int ilo = k - n + 1 > 0 ? k - n + 1 : 0;
    const int ihi = k < n - 1 ? k : n - 1;
    double s = 0.0;
    for (int i = ilo; i <= ihi; ++i) s += a[i] * a[k - i];
    b[k] = s;
  }
}

/* G_i = sum_{j=0}^{n-1} c_{i+j} a_j, ascending j (kernels.cuh:142-158,
 * probe.cpp:87-92). */
void so_corr(const double* c, const double* a, int n, double* G) {
  for (int i = 0; i < n; ++i) {
    double s = 0.0;
    const double* cp = c + i;
    for (int j = 0; j < n; ++j) s += cp[j] * a[j];
    G[i] = s;
  }
}

/* kernels.cuh:221-230 / probe.cpp:81-86. */
double so_lp_weights(const double* b, int nb, double M, double p, double* c) {
  double R = 0.0;
  for (int k = 0; k < nb; ++k) {
    const double d = (b[k] - M) / M;
    const double q = d > -1.0 ? exp((p - 1) * log1p(d)) : 0.0;
    c[k] = q;
    R += q * (1.0 + d);
  }
  return R;
}

double so_Jp_grad(const double* a, int n, double p, double* grad) {
  const int nb = 2 * n - 1;
  double* b = (double*)malloc(sizeof(double) * (size_t)nb);
  so_autoconv(a, n, b);
  double M = 0.0;
  for (int k = 0; k < nb; ++k) M = b[k] > M ? b[k] : M;
  double S = 0.0;
  for (int i = 0; i < n; ++i) S += a[i];
  const double R = so_lp_weights(b, nb, M, p, b);
  const double Rp = pow(R, 1.0 / p);
  const double J = 2.0 * n * M * Rp / (S * S);
  if (grad) {
    double* G = (double*)malloc(sizeof(double) * (size_t)n);
    /* c needs zero padding up to index 2n-2+... : corr reads c[i+j] <= 2n-2 */
    so_corr(b, a, n, G);
    const double scale = 4.0 * n * Rp / (S * S);
    for (int i = 0; i < n; ++i) grad[i] = scale * (G[i] / R - M / S);
    free(G);
  }
  free(b);
  return J;
}

long double so_Jp_grad_ld(const long double* a, int n, long double p, long double* grad) {
  const int nb = 2 * n - 1;
  long double* b = (long double*)malloc(sizeof(long double) * (size_t)nb);
#pragma omp parallel for schedule(static)
  for (int k = 0; k < nb; ++k) {
    const int ilo = k - n + 1 > 0 ? k - n + 1 : 0;
    const int ihi = k < n - 1 ? k : n - 1;
    long double s = 0.0L;
    for (int i = ilo; i <= ihi; ++i) s += a[i] * a[k - i];
    b[k] = s;
  }
  long double M = 0.0L, S = 0.0L, R = 0.0L;
  for (int k = 0; k < nb; ++k) M = b[k] > M ? b[k] : M;
  for (int i = 0; i < n; ++i) S += a[i];
  for (int k = 0; k < nb; ++k) {
    const long double d = (b[k] - M) / M;
    const long double q = d > -1.0L ? expl((p - 1) * log1pl(d)) : 0.0L;
    b[k] = q;
    R += q * (1.0L + d);
  }
  const long double Rp = powl(R, 1.0L / p);
  const long double J = 2.0L * n * M * Rp / (S * S);
  if (grad) {
    const long double scale = 4.0L * n * Rp / (S * S);
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; ++i) {
      long double s = 0.0L;
      for (int j = 0; j < n; ++j) s += b[i + j] * a[j];
      grad[i] = scale * (s / R - M / S);
    }
  }
  free(b);
  return J;
}

/* ------------------------------------------------------ local search ---- */

/* kernels.cuh:196-288 (probe_semantics = 0) and tools/probe.cpp:61-116
 * (probe_semantics = 1).  The two differ only in the entry normalisation,
 * divide-vs-reciprocal renormalisation and the autoconvolution loop order
 * (probe.cpp:32-41 scatters i-major; both add terms for each k in ascending
 * i, so the sums are identical). */
static double local_search_impl(double* a, int n, int iters, double pA, double pB,
                                double lrA, double lrB, double b1, double b2, int mode,
                                int probe_sem, double* m_out, double* v_out, int* iters_done,
                                double* gmin) {
  const int nb = 2 * n - 1;
  double* b = (double*)calloc((size_t)nb, sizeof(double));
  double* cc = (double*)calloc((size_t)nb, sizeof(double));
  double* m = (double*)calloc((size_t)n, sizeof(double));
  double* v = (double*)calloc((size_t)n, sizeof(double));
  double* g = (double*)calloc((size_t)n, sizeof(double));
  double b1t = 1, b2t = 1;
  const double lnP = log(pB / pA), lnL = log(lrB / lrA);
  int done = 0;

  if (probe_sem) { /* probe.cpp:70 */
    double S = 0;
    for (int i = 0; i < n; ++i) S += a[i];
    for (int i = 0; i < n; ++i) a[i] /= S;
  }

  for (int it = 0; it < iters; ++it) {
    const double frac = iters > 1 ? (double)it / (double)(iters - 1) : 0.0;
    const double p = pA * exp(lnP * frac);
    const double lr = lrA * exp(lnL * frac);

    so_autoconv(a, n, b);
    double M = 0;
    for (int k = 0; k < nb; ++k) M = b[k] > M ? b[k] : M;
    if (!(M > 0)) break;

    double R = 0;
    for (int k = 0; k < nb; ++k) {
      const double d = (b[k] - M) / M;
      const double q = d > -1.0 ? exp((p - 1) * log1p(d)) : 0.0;
      cc[k] = q;
      R += q * (1.0 + d);
    }
    for (int i = 0; i < n; ++i) {
      double s = 0;
      const double* cp = cc + i;
      for (int j = 0; j < n; ++j) s += cp[j] * a[j];
      g[i] = s / R - M;
      if (gmin) {
        const double r = fabs(g[i]) / M;
        if (it == 0 || r < gmin[i]) gmin[i] = r;
      }
    }
    b1t *= b1;
    b2t *= b2;
    const double bc1 = 1.0 / (1.0 - b1t), bc2 = 1.0 / (1.0 - b2t);
    double S = 0;
    for (int i = 0; i < n; ++i) {
      double x;
      if (mode == 0) {
        m[i] = b1 * m[i] + (1 - b1) * g[i];
        v[i] = b2 * v[i] + (1 - b2) * g[i] * g[i];
        const double step = (lr / n) * (m[i] * bc1) / (sqrt(v[i] * bc2) + 1e-30);
        x = a[i] - step;
        if (x < 0) x = 0;
      } else {
        double e = -lr * (g[i] / M);
        e = e > 4.0 ? 4.0 : (e < -4.0 ? -4.0 : e);
        x = a[i] * exp(e);
      }
      a[i] = x;
      S += x;
    }
    if (probe_sem) {
      if (S > 0) for (int i = 0; i < n; ++i) a[i] /= S;
      else       for (int i = 0; i < n; ++i) a[i] = 1.0 / n;
    } else {
      if (!(S > 0)) { for (int i = 0; i < n; ++i) a[i] = 1.0 / n; S = 1.0; }
      const double inv = 1.0 / S;
      for (int i = 0; i < n; ++i) a[i] *= inv;
    }
    ++done;
  }
  if (m_out) memcpy(m_out, m, sizeof(double) * (size_t)n);
  if (v_out) memcpy(v_out, v, sizeof(double) * (size_t)n);
  if (iters_done) *iters_done = done;
  free(b); free(cc); free(m); free(v); free(g);
  return so_F_hostorder(a, n);
}

double so_local_search(double* a, int n, int iters, double pA, double pB, double lrA,
                       double lrB, double b1, double b2, int mode, int probe_sem,
                       double* m_out, double* v_out, int* iters_done) {
  return local_search_impl(a, n, iters, pA, pB, lrA, lrB, b1, b2, mode, probe_sem, m_out, v_out,
                           iters_done, NULL);
}

double so_local_search_gmin(double* a, int n, int iters, double pA, double pB, double lrA,
                            double lrB, double b1, double b2, int mode, double* m_out,
                            double* v_out, double* gmin) {
  return local_search_impl(a, n, iters, pA, pB, lrA, lrB, b1, b2, mode, 0, m_out, v_out, NULL,
                           gmin);
}

void so_local_search_batch(double* pop, int P, int n, int iters, double pA, double pB,
                           double lrA, double lrB, double b1, double b2, int mode,
                           int probe_sem, int nthreads, double* F_out) {
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#else
  (void)nthreads;
#endif
#pragma omp parallel for schedule(dynamic)
  for (int p = 0; p < P; ++p) {
    const double F = so_local_search(pop + (size_t)p * n, n, iters, pA, pB, lrA, lrB, b1, b2,
                                     mode, probe_sem, NULL, NULL, NULL);
    if (F_out) F_out[p] = F;
  }
}

/* ------------------------------------------------------- Frank-Wolfe ---- */

/* kernels.cuh:451-503: g = gbar = a, lambda = 0; per iteration the dual step
 * lambda <- [lambda + sigma (a*gbar - M)]_+ over k, then the primal step
 * g_new <- [g + tau (1 - A^T lambda)]_+ over i, then gbar = 2 g_new - g_old. */
void so_pdhg(const double* a, int n, double M, int cpIters, double* g, double* lam) {
  const int nb = 2 * n - 1;
  const double tau = 0.9 / n, sig = 0.9 * n;
  double* gb = (double*)malloc(sizeof(double) * (size_t)n);
  double* gn = (double*)malloc(sizeof(double) * (size_t)n);
  double* lp = (double*)calloc((size_t)(nb + n), sizeof(double)); /* zero padded */
  for (int i = 0; i < n; ++i) { g[i] = a[i]; gb[i] = a[i]; }
  for (int it = 0; it < cpIters; ++it) {
    for (int k = 0; k < nb; ++k) {
      const int ilo = k - n + 1 > 0 ? k - n + 1 : 0;
      const int ihi = k < n - 1 ? k : n - 1;
      double acc = 0.0;
      for (int i = ilo; i <= ihi; ++i) acc += a[i] * gb[k - i];
      const double l = lp[k] + sig * (acc - M);
      lp[k] = l > 0.0 ? l : 0.0;
    }
    for (int i = 0; i < n; ++i) {
      double c = 0.0;
      for (int j = 0; j < n; ++j) c += lp[i + j] * a[j];
      const double x = g[i] + tau * (1.0 - c);
      gn[i] = x > 0.0 ? x : 0.0;
    }
    for (int i = 0; i < n; ++i) {
      gb[i] = 2.0 * gn[i] - g[i];
      g[i] = gn[i];
    }
  }
  if (lam) memcpy(lam, lp, sizeof(double) * (size_t)nb);
  free(gb); free(gn); free(lp);
}

/* kernels.cuh:427-532. */
static int frank_wolfe_impl(double* a, int n, int rounds, int cpIters, int nT, double tMax,
                            double* M_out, double* g_out, double* lsM_out) {
  const int nb = 2 * n - 1;
  double* b = (double*)malloc(sizeof(double) * (size_t)nb);
  double* g = (double*)malloc(sizeof(double) * (size_t)n);
  double* h = (double*)malloc(sizeof(double) * (size_t)n);
  int accepted = 0;
  double Mlast = 0.0;
  double s = 0.0;
  for (int i = 0; i < n; ++i) s += a[i];
  if (!(s > 0.0)) goto out;
  {
    const double inv = 1.0 / s;
    for (int i = 0; i < n; ++i) a[i] *= inv;
  }
  for (int r = 0; r < rounds; ++r) {
    so_autoconv(a, n, b);
    double M = 0.0;
    for (int k = 0; k < nb; ++k) M = b[k] > M ? b[k] : M;
    Mlast = M;
    if (!(M > 0.0)) break;
    so_pdhg(a, n, M, cpIters, g, NULL);
    double sg = 0.0;
    for (int i = 0; i < n; ++i) sg += g[i];
    if (!(sg > 0.0)) break;
    const double ig = 1.0 / sg;
    for (int i = 0; i < n; ++i) g[i] *= ig;
    double bestT = 0.0, bestM = M;
    for (int st = 1; st <= nT; ++st) {
      const double t = tMax * st / nT;
      for (int i = 0; i < n; ++i) h[i] = (1.0 - t) * a[i] + t * g[i];
      so_autoconv(h, n, b);
      double mm = 0.0;
      for (int k = 0; k < nb; ++k) mm = b[k] > mm ? b[k] : mm;
      if (lsM_out) lsM_out[st - 1] = mm;
      if (mm < bestM) { bestM = mm; bestT = t; }
    }
    if (g_out) memcpy(g_out, g, sizeof(double) * (size_t)n);
    if (!(bestT > 0.0)) break;
    for (int i = 0; i < n; ++i) a[i] = (1.0 - bestT) * a[i] + bestT * g[i];
    Mlast = bestM;
    ++accepted;
  }
out:
  if (M_out) *M_out = Mlast;
  free(b); free(g); free(h);
  return accepted;
}

int so_frank_wolfe(double* a, int n, int rounds, int cpIters, int nT, double tMax, double* M_out) {
  return frank_wolfe_impl(a, n, rounds, cpIters, nT, tMax, M_out, NULL, NULL);
}

int so_frank_wolfe_trace(double* a, int n, int rounds, int cpIters, int nT, double tMax,
                         double* M_out, double* g_out, double* lsM_out) {
  return frank_wolfe_impl(a, n, rounds, cpIters, nT, tMax, M_out, g_out, lsM_out);
}

/* ------------------------------------------------------------ refine ---- */

void so_refine(const double* src, int n, double* dst) { /* kernels.cuh:538-547 */
  for (int i = 0; i < n; ++i) { dst[2 * i] = src[i]; dst[2 * i + 1] = src[i]; }
}

/* ------------------------------------------------------ Philox / RNG ---- */

static inline void philox_round(uint32_t c[4], const uint32_t k[2]) {
  const uint64_t p0 = (uint64_t)0xD2511F53u * c[0];
  const uint64_t p1 = (uint64_t)0xCD9E8D57u * c[2];
  const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
  const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
  const uint32_t o0 = hi1 ^ c[1] ^ k[0];
  const uint32_t o2 = hi0 ^ c[3] ^ k[1];
  c[0] = o0; c[1] = lo1; c[2] = o2; c[3] = lo0;
}

void so_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
  uint32_t c[4] = {ctr[0], ctr[1], ctr[2], ctr[3]};
  uint32_t k[2] = {key[0], key[1]};
  for (int r = 0; r < 10; ++r) {
    if (r) { k[0] += 0x9E3779B9u; k[1] += 0xBB67AE85u; }
    philox_round(c, k);
  }
  out[0] = c[0]; out[1] = c[1]; out[2] = c[2]; out[3] = c[3];
}

#define SO_STREAM_VARIATION 0x56415249u /* stream tag in counter word 3 */
#define SO_STREAM_SELECT 0x53454C45u

/* Box-Muller on two Philox words: u = (2*(x>>9)+1) * 2^-24 in (0,1), exact. */
static void bm_pair(uint32_t x0, uint32_t x1, double* z0, double* z1) {
  const double u1 = (double)(2u * (x0 >> 9) + 1u) * (1.0 / 16777216.0);
  const double u2 = (double)(2u * (x1 >> 9) + 1u) * (1.0 / 16777216.0);
  const double r = sqrt(-2.0 * log(u1));
  const double th = 6.283185307179586476925286766559 * u2;
  *z0 = r * cos(th);
  *z1 = r * sin(th);
}

static double normal_d(uint64_t seed, uint64_t gen, int slot, int i) {
  const uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
  const uint32_t ctr[4] = {(uint32_t)(i >> 2), (uint32_t)slot, (uint32_t)gen,
                           (uint32_t)(gen >> 32) ^ SO_STREAM_VARIATION};
  uint32_t x[4];
  so_philox4x32_10(ctr, key, x);
  double z[4];
  bm_pair(x[0], x[1], &z[0], &z[1]);
  bm_pair(x[2], x[3], &z[2], &z[3]);
  return z[i & 3];
}

float so_normal(uint64_t seed, uint64_t gen, int slot, int i) {
  return (float)normal_d(seed, gen, slot, i);
}

/* The 64 per-offspring operator parameters in (0, 1] (the role of sPar,
 * kernels.cuh:604-609): Philox4x32-10 counters {q, slot, gen, gen>>32 ^ tag},
 * q = 0..15, four words each, u = ((x >> 8) + 1) 2^-24 (exact in fp32). */
#define SO_STREAM_PARAMS 0x50524D53u
void so_op_params(uint64_t seed, uint64_t gen, int slot, float prm[64]) {
  const uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
  for (int q = 0; q < 16; ++q) {
    const uint32_t ctr[4] = {(uint32_t)q, (uint32_t)slot, (uint32_t)gen,
                             (uint32_t)(gen >> 32) ^ SO_STREAM_PARAMS};
    uint32_t x[4];
    so_philox4x32_10(ctr, key, x);
    for (int w = 0; w < 4; ++w) prm[4 * q + w] = (float)((x[w] >> 8) + 1u) * (1.0f / 16777216.0f);
  }
}

/* One height of offspring operator `op` (kernels.cuh:614-787), in fp64 and
 * rounded once to fp32 by the caller: A the first parent, B the second (the
 * crossover ops 0, 1, 8), u the parameters, mean = sum(A) / n (correctly
 * rounded sum), sg the strength.  The same expression, operation for
 * operation, is the device's vary(). */
static double vary_one(int op, int i, int n, const float* A, const float* B, const double* u,
                       double mean, double sg, uint64_t seed, uint64_t gen, int slot) {
  const double a = A[i];
  const double twopi = 6.283185307179586476925286766559;
  switch (op) {
    case 0: {  /* blend crossover, lambda in [-0.3, 1.3] */
      const double lam = -0.3 + 1.6 * u[0];
      const double x = lam * a + (1.0 - lam) * (double)B[i];
      return x > 0.0 ? x : 0.0;
    }
    case 1: {  /* two-point segment crossover */
      int c1 = (int)(u[1] * n), c2 = (int)(u[2] * n);
      if (c1 > c2) { const int t = c1; c1 = c2; c2 = t; }
      return (i >= c1 && i < c2) ? (double)B[i] : a;
    }
    case 2: {  /* log-normal multiplicative noise */
      const double x = a * exp(sg * normal_d(seed, gen, slot, i));
      return x > 0.0 ? x : 0.0;
    }
    case 3: {  /* smooth low-frequency perturbation */
      const int nf = 1 + (int)(u[3] * 5.0);
      double m = 0.0;
      for (int f = 0; f < nf; ++f) {
        const double freq = 1.0 + u[4 + f] * 12.0;
        m += cos(twopi * freq * (double)i / (double)n + twopi * u[12 + f]) / (double)nf;
      }
      const double x = a * (1.0 + sg * 3.0 * m);
      return x > 0.0 ? x : 0.0;
    }
    case 4: {  /* additive Gaussian noise scaled by the mean height */
      const double x = a + sg * mean * normal_d(seed, gen, slot, i);
      return x > 0.0 ? x : 0.0;
    }
    case 5: {  /* sparsify below a fraction of the mean */
      const double thr = mean * (0.05 + 0.9 * u[20]);
      return a < thr ? 0.0 : a;
    }
    case 6: {  /* symmetrise, or hard reflect */
      const double r = A[n - 1 - i];
      return u[21] < 0.5 ? r : 0.5 * (a + r);
    }
    case 7: {  /* shift by up to n/8 either way, zero fill */
      const int sh = (int)((u[22] - 0.5) * (double)n * 0.25);
      const int k = i - sh;
      return (k >= 0 && k < n) ? (double)A[k] : 0.0;
    }
    case 8: {  /* splice a random block of B into A */
      const int len = 1 + (int)(u[23] * 0.5 * (double)n);
      const int st0 = (int)(u[24] * (double)(n - len > 0 ? n - len : 1));
      return (i >= st0 && i < st0 + len) ? (double)B[i] : a;
    }
    case 9: {  /* random restart: smooth positive random function */
      const int nf = 2 + (int)(u[25] * 6.0);
      double m = 0.0;
      for (int f = 0; f < nf; ++f) {
        const double freq = 1.0 + u[4 + (f % 8)] * 10.0;
        m += cos(twopi * freq * (double)i / (double)n + twopi * u[12 + (f % 8)]);
      }
      const double v = 0.3 + 0.7 * u[26] + m / (double)nf + 0.4 * normal_d(seed, gen, slot, i);
      return v > 0.0 ? v : 0.0;
    }
    case 10: {  /* spike of width 2w+1 at a random location */
      const int loc = (int)(u[27] * (double)n);
      const int wdt = 1 + (int)(u[28] * (double)n * 0.03);
      const double amp = mean * (1.0 + 6.0 * u[29]);
      return (i >= loc - wdt && i <= loc + wdt) ? a + amp : a;
    }
    case 11: {  /* structured seed: plateau, depression, hump, band, gap, spike, atom */
      const int flip = u[30] < 0.5;
      const double atom = 30.0 + 250.0 * u[31];
      const double pl1 = 0.28 + 0.16 * u[32];
      const double pl2 = pl1 + 0.10 + 0.14 * u[33];
      const double hump = pl2 + 0.02 + 0.08 * u[34];
      const double gap0 = 0.70 + 0.16 * u[35];
      const double gap1 = gap0 + 0.05 + 0.15 * u[36];
      const double tail = 3.0 + 14.0 * u[37];
      const int nf = 3 + (int)(u[38] * 6.0);
      const int d = flip ? n - 1 - i : i;
      const double x = ((double)d + 0.5) / (double)n;
      double rough = 0.0;
      for (int f = 0; f < nf; ++f)
        rough += cos(twopi * (2.0 + u[40 + f] * 40.0) * x + twopi * u[48 + f]);
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
    case 12: {  /* carve out (or fill in) a hard support gap */
      const int len = 1 + (int)((0.015 + 0.13 * u[30]) * (double)n);
      const int st0 = (int)(u[31] * (double)(n > len ? n - len : 1));
      if (i < st0 || i >= st0 + len) return a;
      return u[32] < 0.7 ? 0.0 : a + mean * (0.2 + 1.8 * u[33]);
    }
    case 13: {  /* grow or shrink the endpoint atom */
      const int w = 1 + (int)(u[31] * 0.012 * (double)n);
      const int d = u[30] < 0.5 ? i : n - 1 - i;
      if (d >= w) return a;
      const double x = a + mean * (-20.0 + 260.0 * u[32]) / (double)(1 + d);
      return x > 0.0 ? x : 0.0;
    }
    default:   /* clone */
      return a;
  }
}

int so_op_uses_mean(int op) { return op == 4 || op == 5 || op == 10 || op == 12 || op == 13; }

/* kernels.cuh:582-801: offspring = operator `op` on parents A (and B), each
 * height formed in fp64 and rounded once to fp32, then renormalised to sum 1
 * by the correctly rounded fp64 sum of the fp32 heights (uniform 1/n if that
 * sum is not positive).  B = NULL means B = A. */
void so_variation2(const float* A, const float* B, int n, int op, float sigma, uint64_t seed,
                   uint64_t gen, int slot, float* out) {
  if (!B) B = A;
  float prm[64];
  double u[64];
  so_op_params(seed, gen, slot, prm);
  for (int k = 0; k < 64; ++k) u[k] = prm[k];
  double* tmp = (double*)malloc(sizeof(double) * (size_t)n);
  double mean = 0.0;
  if (so_op_uses_mean(op)) {
    for (int i = 0; i < n; ++i) tmp[i] = A[i];
    mean = so_fsum(tmp, n) / (double)n;
  }
  for (int i = 0; i < n; ++i)
    out[i] = (float)vary_one(op, i, n, A, B, u, mean, (double)sigma, seed, gen, slot);
  for (int i = 0; i < n; ++i) tmp[i] = out[i];
  const double S = so_fsum(tmp, n); /* correctly rounded sum of the fp32 heights */
  free(tmp);
  if (S > 0.0) {
    const double inv = 1.0 / S;
    for (int i = 0; i < n; ++i) out[i] = (float)((double)out[i] * inv);
  } else {
    for (int i = 0; i < n; ++i) out[i] = 1.0f / (float)n;
  }
}

void so_variation(const float* A, int n, int op, float sigma, uint64_t seed, uint64_t gen,
                  int slot, float* out) {
  so_variation2(A, NULL, n, op, sigma, seed, gen, slot, out);
}

/* -------------------------------------------------- support signature ---- */

/* kernels.cuh:803-849 (kSupportSig), BD = 256 (kernels.cuh:30), tol = 0.02
 * as runSig passes it (main.cu:180-184).  The fp32 mean is summed in the
 * kernel's order: thread t of 256 adds a[t], a[t+256], ... in sequence;
 * blockSum (kernels.cuh:80-90) reduces each 32-lane warp by the shuffle tree
 * v += v[lane + 16], +8, +4, +2, +1 and adds the 8 warp sums in order from 0.
 * The support is {i : a_i > mean * tol}; the signature is the XOR over the
 * support of mix(i) (x = i * 0x9E3779B97F4A7C15, x ^= x >> 29,
 * x *= 0xBF58476D1CE4E5B9, x ^= x >> 32): every thread starts its hash at the
 * FNV offset basis, and 256 copies of it cancel in the XOR reduction.  nnz is
 * the support size. */
static float warp_tree_sum32(const float* v) {
  float w[32];
  for (int l = 0; l < 32; ++l) w[l] = v[l];
  for (int off = 16; off > 0; off >>= 1)
    for (int l = 0; l < 32; ++l)  /* ascending l: w[l + off] is not yet updated */
      w[l] = w[l] + (l + off < 32 ? w[l + off] : w[l]);
  return w[0];
}

uint64_t so_sig_mix(int i) {
  uint64_t x = (uint64_t)i * 0x9E3779B97F4A7C15ULL;
  x ^= x >> 29;
  x *= 0xBF58476D1CE4E5B9ULL;
  x ^= x >> 32;
  return x;
}

void so_support_sig(const float* a, int n, float tol, uint64_t* sig, int* nnz) {
  float part[256], warp[8];
  for (int t = 0; t < 256; ++t) {
    float s = 0.0f;
    for (int i = t; i < n; i += 256) s += a[i];
    part[t] = s;
  }
  for (int w = 0; w < 8; ++w) warp[w] = warp_tree_sum32(part + 32 * w);
  float tot = 0.0f;
  for (int w = 0; w < 8; ++w) tot += warp[w];
  const float mean = tot / (float)n;
  const float thr = mean * tol;
  uint64_t h = 0;
  int c = 0;
  for (int i = 0; i < n; ++i)
    if (a[i] > thr) {
      ++c;
      h ^= so_sig_mix(i);
    }
  *sig = h;
  *nnz = c;
}

/* ---------------------------------------------------------- selection ---- */

static inline int rank_less(const double* fit, int x, int y) {
  return fit[x] < fit[y] || (fit[x] == fit[y] && x < y);
}

/* main.cu:676-735 with op fixed to 2 (the uneven-islands workload's Gaussian multiplicative
 * mutation).  Tournament draws come from Philox keyed by (seed, gen, slot). */
void so_select_plan(double* fit, int P, int NI, int elite, uint64_t seed, uint64_t gen,
                    int migrate_every, float sigma, int* p1, int* p2, int* op,
                    float* strength, int* mig_src, int* mig_dst) {
  const int PI = P / NI;
  int* ord = (int*)malloc(sizeof(int) * (size_t)P);
  for (int i = 0; i < NI; ++i) {
    /* rank of each member: number of members ranked before it */
    for (int r = 0; r < PI; ++r) {
      const int x = i * PI + r;
      int rk = 0;
      for (int q = 0; q < PI; ++q)
        if (rank_less(fit, i * PI + q, x)) ++rk;
      ord[i * PI + rk] = x;
    }
  }
  for (int i = 0; i < NI; ++i) {
    if (mig_src) mig_src[i] = -1;
    if (mig_dst) mig_dst[i] = -1;
  }
  if (migrate_every > 0 && gen > 0 && gen % (uint64_t)migrate_every == 0 && NI > 1) {
    for (int i = 0; i < NI; ++i) { /* main.cu:685-691, sequential over i */
      const int src = ord[i * PI + 0];
      const int dst = ord[((i + 1) % NI) * PI + PI - 1];
      fit[dst] = fit[src];
      if (mig_src) mig_src[i] = src;
      if (mig_dst) mig_dst[i] = dst;
    }
  }
  const uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
  for (int i = 0; i < NI; ++i) {
    for (int r = 0; r < PI; ++r) {
      const int slot = i * PI + r;
      if (r < elite) {
        p1[slot] = ord[i * PI + r];
        p2[slot] = ord[i * PI + r];
        op[slot] = 20;
        strength[slot] = 0.0f;
        continue;
      }
      const uint32_t ctr[4] = {(uint32_t)slot, 0u, (uint32_t)gen,
                               (uint32_t)(gen >> 32) ^ SO_STREAM_SELECT};
      uint32_t x[4];
      so_philox4x32_10(ctr, key, x);
      int pick[4];
      for (int q = 0; q < 4; ++q)
        pick[q] = ord[i * PI + (int)(((uint64_t)x[q] * (uint64_t)PI) >> 32)];
      /* main.cu:708-711: hfit[a] < hfit[b] ? a : b */
      p1[slot] = fit[pick[0]] < fit[pick[1]] ? pick[0] : pick[1];
      p2[slot] = fit[pick[2]] < fit[pick[3]] ? pick[2] : pick[3];
      op[slot] = 2;
      strength[slot] = sigma;
    }
  }
  free(ord);
}
