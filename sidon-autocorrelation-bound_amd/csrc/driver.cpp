// driver.cpp -- C++ host driver over the C ABI (include/sidon.h).
//
// The counterpart of the reference's sidon.exe host side (src/main.cu:40-876)
// for the paths this library implements.  It takes the reference's command
// line (src/main.cu:41-82, :357-381: the same flags, meanings and defaults;
// unknown options exit 1) and runs, without a mode flag, the reference's
// search.  Modes:
//   (none) / --search        the reference's search loop: three-tier local search,
//                            Frank-Wolfe rounds, fp64 evaluation, support signatures,
//                            the island policy (tabu, operator weights, migration), all
//                            fourteen operators, n -> 2n split per level, fp64 polish,
//                            best vector and certificate out
//   --selftest               known-answer tests of src/main.cu:186-291 against the device
//   --bench                  kernel micro-benchmark (src/main.cu:294-320)
//   --uneven                 island generations of uneven lengths as HIP graphs
//   --fw-ladder              Frank-Wolfe (Chambolle-Pock) rounds + n -> 2n cell split
//   --search-graph           ladder search on the device generation graph (operator 2 only)
// Results are JSON lines; the search also prints the reference's STAT
// progress records.  All device work goes through the C ABI; host code here
// only seeds inputs, reads results and writes files.
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "../../include/sidon.h"

namespace {

using Clock = std::chrono::steady_clock;
double seconds_since(Clock::time_point t0) {
  return std::chrono::duration<double>(Clock::now() - t0).count();
}

#define CHECK_HIP(x)                                                                  \
  do {                                                                                \
    hipError_t _e = (x);                                                              \
    if (_e != hipSuccess) {                                                           \
      std::fprintf(stderr, "HIP error %s at %s:%d: %s\n", #x, __FILE__, __LINE__,     \
                   hipGetErrorString(_e));                                            \
      std::exit(3);                                                                   \
    }                                                                                 \
  } while (0)
#define CHECK_SIDON(x)                                                                \
  do {                                                                                \
    int _r = (x);                                                                     \
    if (_r != SIDON_OK) {                                                             \
      std::fprintf(stderr, "%s failed (%s): %s\n", #x, sidon_status_string(_r),      \
                   sidon_last_error());                                               \
      std::exit(4);                                                                   \
    }                                                                                 \
  } while (0)

// fp64 objective in hostF order (src/main.cu:94-105): the host-side recheck
// the reference prints next to every reported bound.
double host_F(const std::vector<double>& a) {
  const int n = (int)a.size();
  std::vector<double> b(2 * n - 1, 0.0);
  for (int i = 0; i < n; ++i) {
    if (a[i] == 0.0) continue;
    for (int j = 0; j < n; ++j) b[i + j] += a[i] * a[j];
  }
  double M = 0.0;
  for (double x : b) M = std::max(M, x);
  double S = 0.0;
  for (double x : a) S += x;
  return 2.0 * n * M / (S * S);
}

// Structured prior: endpoint atom = 5 x body mean, zero cells [1, 0.1 n),
// body LogNormal(0,1) (1 + 0.5 i/n) with 33 % exact zeros, sum 1.
void structured_prior(std::mt19937_64& rng, int n, float* out) {
  std::lognormal_distribution<double> LN(0.0, 1.0);
  std::uniform_real_distribution<double> U(0.0, 1.0);
  std::vector<double> a(n, 0.0);
  const int gap = std::max(1, (int)(0.1 * n));
  double sum = 0.0;
  for (int i = gap; i < n; ++i) {
    const double v = LN(rng) * (1.0 + 0.5 * i / n);
    a[i] = U(rng) < 0.33 ? 0.0 : v;
    sum += a[i];
  }
  a[0] = 5.0 * sum / std::max(1, n - gap);
  double S = 0.0;
  for (double x : a) S += x;
  for (int i = 0; i < n; ++i) out[i] = (float)(a[i] / S);
}

// A fresh host seed of the reference's level start (src/main.cu:505-551): with
// probability 0.6 the structured shape every record shares, else a smooth
// random Fourier function with noise; 2..7 cosine modes either way; heights
// clamped at zero, then scaled to sum 1 in fp64 (uniform 1/n if all zero).
// The structured shape, over x = (d + 1/2)/n with d = i or n-1-i (a mirror
// image half the time): five bands [0, e1) [e1, e2) [e2, e3) [e3, e4)
// [e4, e5) of base heights 4.2, 0.7, 5.2, 2.0, 0 with a relative ripple of
// 0.45, 0.80, 0.30, 0.70, 0 times the mean of the modes, then a tail band of
// height t (ripple 0.20); an atom of mass A on cell d = 0 and A/(10 d) on
// d = 1..3.
void host_seed(std::mt19937_64& rng, int n, float* out) {
  std::uniform_real_distribution<double> U(0.0, 1.0);
  std::normal_distribution<double> N01(0.0, 1.0);
  const double kTwoPi = 6.283185307;
  const int modes = 2 + (int)(U(rng) * 6);
  std::vector<double> freq(modes), phase(modes);
  for (int f = 0; f < modes; ++f) {
    freq[f] = 1.0 + 10.0 * U(rng);
    phase[f] = kTwoPi * U(rng);
  }
  std::vector<double> v(n, 0.0);
  if (U(rng) < 0.6) {
    const bool mirror = U(rng) < 0.5;
    const double atom = 30.0 + 250.0 * U(rng);
    double edge[5];
    edge[0] = 0.28 + 0.16 * U(rng);
    edge[1] = edge[0] + 0.10 + 0.14 * U(rng);
    edge[2] = edge[1] + 0.02 + 0.08 * U(rng);
    edge[3] = 0.70 + 0.16 * U(rng);
    edge[4] = edge[3] + 0.05 + 0.15 * U(rng);
    const double tail = 3.0 + 14.0 * U(rng);
    const double height[6] = {4.2, 0.7, 5.2, 2.0, 0.0, tail};
    const double ripple[6] = {0.45, 0.80, 0.30, 0.70, 0.0, 0.20};
    for (int i = 0; i < n; ++i) {
      const int d = mirror ? n - 1 - i : i;
      const double x = (d + 0.5) / n;
      double wave = 0.0;
      for (int f = 0; f < modes; ++f) wave += std::cos(kTwoPi * (2.0 + 6.0 * freq[f]) * x + phase[f]);
      wave /= modes;
      int band = 0;
      while (band < 5 && x >= edge[band]) ++band;
      double h = height[band] * (1.0 + ripple[band] * wave);
      if (d == 0) h += atom;
      else if (d < 4) h += atom * 0.10 / d;
      v[i] = (double)(float)std::max(0.0, h);
    }
  } else {
    const double base = 0.3 + 0.7 * U(rng);
    const double amp = 0.2 + 1.2 * U(rng);
    for (int i = 0; i < n; ++i) {
      double wave = 0.0;
      for (int f = 0; f < modes; ++f) wave += std::cos(kTwoPi * freq[f] * i / n + phase[f]);
      v[i] = (double)(float)std::max(0.0, base + amp * wave / modes + 0.3 * N01(rng));
    }
  }
  double S = 0.0;
  for (double x : v) S += x;
  for (int i = 0; i < n; ++i) out[i] = S > 0.0 ? (float)(v[i] / S) : 1.0f / n;
}

template <typename T>
T* dalloc(size_t count) {
  T* p = nullptr;
  CHECK_HIP(hipMalloc(&p, count * sizeof(T)));
  CHECK_HIP(hipMemset(p, 0, count * sizeof(T)));
  return p;
}

// saved-vector text format (src/main.cu:329-338)
void save_vec(const std::string& path, const std::vector<double>& a, double F) {
  FILE* f = std::fopen(path.c_str(), "w");
  if (!f) return;
  std::fprintf(f, "# AlphaEvolve problem 2 -- autocorrelation / Sidon\n");
  std::fprintf(f, "# F = 2*n*max_k b_k / (sum a)^2  is an upper bound on C\n");
  std::fprintf(f, "n %zu\nF %.17e\n", a.size(), F);
  for (double x : a) std::fprintf(f, "%.17g\n", x);
  std::fclose(f);
}

// saved-vector reader (src/main.cu:340-352): '#' comments, "n N", "F x", then
// one height a line; false when unreadable, empty, or its count disagrees with n.
bool load_vec(const std::string& path, std::vector<double>& a) {
  FILE* f = std::fopen(path.c_str(), "r");
  if (!f) return false;
  a.clear();
  char line[512];
  int n = 0;
  while (std::fgets(line, sizeof line, f)) {
    if (line[0] == '#' || line[0] == '\n' || line[0] == '\0') continue;
    if (line[0] == 'n' && line[1] == ' ') { n = std::atoi(line + 2); continue; }
    if (line[0] == 'F' && line[1] == ' ') continue;
    a.push_back(std::strtod(line, nullptr));
  }
  std::fclose(f);
  if (n && (int)a.size() != n) return false;
  return !a.empty();
}

// A vector of m heights onto n cells by nearest-lower index, sum 1
// (the reference's injectDouble, src/main.cu:555-562).
void resample(const std::vector<double>& v, int n, float* dst) {
  const int m = (int)v.size();
  double S = 0.0;
  for (int i = 0; i < n; ++i) {
    const double x = v[std::min(m - 1, (int)((long long)i * m / n))];
    dst[i] = (float)x;
    S += x;
  }
  if (S > 0.0)
    for (int i = 0; i < n; ++i) dst[i] = (float)(dst[i] / S);
}

// certificate JSON (the schema of certificates/*.json)
void save_certificate(const std::string& path, const std::vector<double>& a, double F) {
  FILE* f = std::fopen(path.c_str(), "w");
  if (!f) return;
  int nz = 0;
  for (double x : a) nz += x != 0.0;
  std::fprintf(f,
               "{\"problem\": \"AlphaEvolve problem 2 - an autocorrelation problem related to "
               "Sidon sets\", \"claim\": \"C <= F, where C is the largest constant in the "
               "problem statement\", \"formula\": \"F = 2*n*max_k b_k / (sum_i a_i)^2,  b_k = "
               "sum_{i+j=k} a_i a_j,  a_i >= 0\", \"n\": %zu, \"F\": %.17g, "
               "\"non_zero_cells\": %d, \"found_by\": \"sidon-autocorrelation-bound_amd "
               "(MI355X)\", \"weights\": [",
               a.size(), F, nz);
  for (size_t i = 0; i < a.size(); ++i) std::fprintf(f, "%s%.17g", i ? ", " : "", a[i]);
  std::fprintf(f, "]}\n");
  std::fclose(f);
}

// Options.  The defaults are the reference's (src/main.cu:41-62), which the
// search uses; the other modes replace some of them (mode_defaults) unless
// they are given on the command line.
struct Args {
  std::string mode = "--search";
  std::vector<int> ladder{64, 128, 256, 512, 1024, 2048};  // --ladder
  int P = 256, islands = 8, elite = 2;                     // --pop --islands --elite
  int iters = 4000, cold_iters = 20000;                    // --ls-iters --cold-iters
  double pmax = 2e4;                                       // --pmax
  int fw_rounds = 6, fw_iters = 500;                       // --fw-rounds --fw-iters (0 rounds = off)
  int gens = 0;                                            // --gens: cap per level, 0 = time only
  double seconds = 600.0;                                  // --seconds
  unsigned long long seed = 20260905ULL;                   // --seed
  std::string out = "best";                                // --out
  std::string resume;                                      // --resume
  int polish_iters = 40000, polish_top = 8;                // --polish-iters --polish-top
  int verbose = 1;                                         // --quiet sets 0
  // this driver's own modes
  int nT = 12;
  float sigma = 0.05f, t_max = 0.30f;
  std::vector<int> lengths{600, 1200, 2100, 4096, 8400};
  bool tune = true;
  int split = 1;  // --no-split: every candidate on one workgroup
};

// Parses "64,128,..." into positive integers; false on anything else.
bool parse_list(const char* s, std::vector<int>& v) {
  v.clear();
  std::string cur;
  for (const char* p = s;; ++p) {
    if (*p == ',' || *p == '\0') {
      if (!cur.empty()) {
        char* end = nullptr;
        const long x = std::strtol(cur.c_str(), &end, 10);
        if (*end || x <= 0 || x > (1L << 24)) return false;
        v.push_back((int)x);
      }
      cur.clear();
      if (!*p) break;
    } else {
      cur.push_back(*p);
    }
  }
  return !v.empty();
}

// ------------------------------------------------------------ self-test --
int selftest() {
  int bad = 0;
  auto check = [&](const char* name, double got, double want, double tol) {
    const bool ok = std::fabs(got - want) <= tol * std::max(1.0, std::fabs(want));
    std::printf("  %-32s got %.15f want %.15f %s\n", name, got, want, ok ? "OK" : "FAIL");
    bad += !ok;
  };
  auto gpu_F = [&](const std::vector<float>& h, int n, int P) {
    float* d = dalloc<float>(h.size());
    double* F = dalloc<double>(P);
    CHECK_HIP(hipMemcpy(d, h.data(), h.size() * sizeof(float), hipMemcpyHostToDevice));
    CHECK_SIDON(sidon_eval_f64(d, n, P, F, nullptr, 1e-9, 0));
    std::vector<double> out(P);
    CHECK_HIP(hipMemcpy(out.data(), F, P * sizeof(double), hipMemcpyDeviceToHost));
    CHECK_HIP(hipFree(d));
    CHECK_HIP(hipFree(F));
    return out;
  };
  // (1)-(4) closed forms (src/main.cu:196-207), on the device
  for (int n : {4, 8, 37, 128}) {
    char nm[64];
    std::snprintf(nm, sizeof nm, "constant n=%d", n);
    check(nm, gpu_F(std::vector<float>(n, 1.0f), n, 1)[0], 2.0, 1e-15);
  }
  check("(1,2,2,1)", gpu_F({1, 2, 2, 1}, 4, 1)[0], 80.0 / 36.0, 1e-15);
  check("(1,0,0,1)", gpu_F({1, 0, 0, 1}, 4, 1)[0], 4.0, 1e-15);
  check("spike (0,1,0,0)", gpu_F({0, 1, 0, 0}, 4, 1)[0], 8.0, 1e-15);
  // (5) refinement: correctly rounded F is preserved bit for bit
  {
    std::mt19937 rng(7);
    std::uniform_real_distribution<float> U(0.0f, 1.0f);
    for (int n : {23, 34, 45}) {
      std::vector<float> h(n);
      for (auto& x : h) x = U(rng);
      float* d = dalloc<float>(n);
      float* r = dalloc<float>(2 * n);
      double* F = dalloc<double>(2);
      int* amb = dalloc<int>(2);
      CHECK_HIP(hipMemcpy(d, h.data(), n * sizeof(float), hipMemcpyHostToDevice));
      CHECK_SIDON(sidon_refine(d, n, 1, r, 0));
      CHECK_SIDON(sidon_eval_exact(d, n, 1, F, nullptr, nullptr, amb, 0));
      CHECK_SIDON(sidon_eval_exact(r, 2 * n, 1, F + 1, nullptr, nullptr, amb + 1, 0));
      double f[2];
      int a2[2];
      CHECK_HIP(hipMemcpy(f, F, sizeof f, hipMemcpyDeviceToHost));
      CHECK_HIP(hipMemcpy(a2, amb, sizeof a2, hipMemcpyDeviceToHost));
      char nm[64];
      std::snprintf(nm, sizeof nm, "refine bit-exact n=%d", n);
      check(nm, f[1], f[0], 0.0);
      bad += (a2[0] != 0 || a2[1] != 0);
      CHECK_HIP(hipFree(d)); CHECK_HIP(hipFree(r)); CHECK_HIP(hipFree(F)); CHECK_HIP(hipFree(amb));
    }
  }
  // (7) device F vs host F on the same fp32 values (src/main.cu:227-256)
  {
    std::mt19937 rng(3);
    std::uniform_real_distribution<double> U(0, 1);
    const int n = 257, P = 5;
    std::vector<float> h((size_t)n * P);
    for (int p = 0; p < P; ++p)
      for (int i = 0; i < n; ++i) h[(size_t)p * n + i] = p == 0 ? 1.0f : (float)U(rng);
    const auto got = gpu_F(h, n, P);
    for (int p = 0; p < P; ++p) {
      std::vector<double> a(h.begin() + (size_t)p * n, h.begin() + (size_t)(p + 1) * n);
      char nm[64];
      std::snprintf(nm, sizeof nm, "device F individual %d", p);
      check(nm, got[p], host_F(a), 0.0);  // identical order of operations: bit for bit
    }
  }
  // (8) local search strictly improves (src/main.cu:257-288)
  {
    const int n = 128, P = 4;
    std::mt19937 rng(5);
    std::uniform_real_distribution<float> U(0.1f, 1.0f);
    std::vector<float> h((size_t)n * P);
    for (auto& x : h) x = U(rng);
    const auto F0 = gpu_F(h, n, P);
    float* d = dalloc<float>(h.size());
    CHECK_HIP(hipMemcpy(d, h.data(), h.size() * sizeof(float), hipMemcpyHostToDevice));
    CHECK_SIDON(sidon_local_search(d, nullptr, nullptr, nullptr, nullptr, n, P, 800, 4.0f,
                                   300.0f, 0.06f, 0.002f, 0.9f, 0.999f, 0, 0));
    CHECK_HIP(hipMemcpy(h.data(), d, h.size() * sizeof(float), hipMemcpyDeviceToHost));
    CHECK_HIP(hipFree(d));
    const auto F1 = gpu_F(h, n, P);
    bool nonneg = true;
    for (float x : h) nonneg &= (x >= 0.0f && std::isfinite(x));
    for (int p = 0; p < P; ++p) {
      std::printf("  local search %d: %.9f -> %.9f %s\n", p, F0[p], F1[p],
                  F1[p] < F0[p] ? "OK" : "FAIL");
      bad += !(F1[p] < F0[p]);
    }
    std::printf("  non-negativity preserved: %s\n", nonneg ? "OK" : "FAIL");
    bad += !nonneg;
  }
  std::printf("{\"selftest\": \"%s\", \"failures\": %d}\n", bad ? "FAILED" : "passed", bad);
  return bad ? 1 : 0;
}

// --------------------------------------------------- uneven island graph --
int uneven(const Args& A) {
  const int P = A.P, nmax = *std::max_element(A.lengths.begin(), A.lengths.end());
  const int stride = nmax;
  std::mt19937_64 rng(A.seed);
  std::vector<int> ns(P);
  std::uniform_int_distribution<int> pick(0, (int)A.lengths.size() - 1);
  for (int p = 0; p < P; ++p) ns[p] = A.lengths[pick(rng)];
  std::vector<float> h((size_t)P * stride, 0.0f);
  for (int p = 0; p < P; ++p) structured_prior(rng, ns[p], &h[(size_t)p * stride]);

  hipStream_t s;
  CHECK_HIP(hipStreamCreate(&s));  // blocking: orders with the null-stream copies below
  if (A.tune) {  // autotune the launch configuration at the longest length (it sets the launch)
    float* scratch = dalloc<float>((size_t)64 * nmax);
    float* ref = dalloc<float>((size_t)64 * nmax);
    float* rows = dalloc<float>((size_t)2 * 64 * nmax);  // the launches run with rows (workspace)
    std::vector<float> r((size_t)64 * nmax);
    for (int p = 0; p < 64; ++p) structured_prior(rng, nmax, &r[(size_t)p * nmax]);
    CHECK_HIP(hipMemcpy(ref, r.data(), r.size() * sizeof(float), hipMemcpyHostToDevice));
    int blk = 0, depth = 0, per_cu = 0;
    float ms = 0.0f;
    CHECK_SIDON(sidon_tune_local_search(scratch, ref, rows, nmax, 64, 10, 2, &blk, &depth, &per_cu,
                                        &ms, nullptr, s));
    CHECK_HIP(hipFree(rows));
    std::printf("{\"autotune\": {\"n\": %d, \"block\": %d, \"depth\": %d, \"per_cu\": %d, "
                "\"ms\": %.3f}}\n",
                nmax, blk, depth, per_cu, ms);
    CHECK_HIP(hipFree(scratch));
    CHECK_HIP(hipFree(ref));
  }
  sidon_generation_desc d{};
  d.pop_a = dalloc<float>(h.size());
  d.pop_b = dalloc<float>(h.size());
  d.ns_a = dalloc<int>(P);
  d.ns_b = dalloc<int>(P);
  d.fit = dalloc<double>(P);
  d.nact = dalloc<int>(P);
  d.p1 = dalloc<int>(P);
  d.p2 = dalloc<int>(P);
  d.op = dalloc<int>(P);
  d.strength = dalloc<float>(P);
  d.order = dalloc<int>(P);
  d.gen = dalloc<unsigned long long>(1);
  d.ls_workspace_bytes = sidon_local_search_ragged_workspace(P, stride);
  d.ls_workspace = dalloc<unsigned char>(d.ls_workspace_bytes);
  d.split = A.split;
  d.P = P; d.stride = stride; d.nmax = nmax; d.islands = A.islands; d.elite = A.elite;
  d.migrate_every = 12; d.iters = A.iters;
  d.pStart = 400.0f; d.pEnd = 6e4f; d.lrStart = 0.012f; d.lrEnd = 0.0003f;  // warm tier, main.cu:629
  d.sigma = A.sigma; d.activeTol = 1e-4; d.seed = A.seed;
  CHECK_HIP(hipMemcpy(d.pop_a, h.data(), h.size() * sizeof(float), hipMemcpyHostToDevice));
  CHECK_HIP(hipMemcpy(d.ns_a, ns.data(), P * sizeof(int), hipMemcpyHostToDevice));
  CHECK_HIP(hipMemcpy(d.ns_b, ns.data(), P * sizeof(int), hipMemcpyHostToDevice));
  sidon_generation* g = nullptr;
  CHECK_SIDON(sidon_generation_create(&d, &g, s));
  CHECK_SIDON(sidon_generation_run(g, 1, s));  // warm-up generation
  CHECK_HIP(hipStreamSynchronize(s));
  long long steps = 0;
  for (int p = 0; p < P; ++p) steps += A.iters;  // candidate-steps per generation
  hipEvent_t e0, e1;
  CHECK_HIP(hipEventCreate(&e0));
  CHECK_HIP(hipEventCreate(&e1));
  unsigned long long block_steps = 0, total_steps = 0;
  CHECK_SIDON(sidon_ls_block_steps(&block_steps, 1));
  // one generation at a time: its time (HIP events around the graph replay), the matrix
  // work its local search executed, and the lengths of the population it searched
  float ms = 0.0f;
  std::string per_gen;
  std::vector<int> cur(P);
  for (int gi = 0; gi < A.gens; ++gi) {
    int parity = 0;
    CHECK_SIDON(sidon_generation_current(g, &parity));
    CHECK_HIP(hipMemcpy(cur.data(), parity ? d.ns_b : d.ns_a, P * sizeof(int),
                        hipMemcpyDeviceToHost));
    CHECK_HIP(hipEventRecord(e0, s));
    CHECK_SIDON(sidon_generation_run(g, 1, s));
    CHECK_HIP(hipEventRecord(e1, s));
    CHECK_HIP(hipEventSynchronize(e1));
    float gms = 0.0f;
    CHECK_HIP(hipEventElapsedTime(&gms, e0, e1));
    CHECK_SIDON(sidon_ls_block_steps(&block_steps, 1));
    total_steps += block_steps;
    ms += gms;
    int split_n = 0;  // candidates this generation's local search split (byte 8 of its workspace)
    CHECK_HIP(hipMemcpy(&split_n, static_cast<char*>(d.ls_workspace) + 8, sizeof split_n,
                        hipMemcpyDeviceToHost));
    char buf[512];
    int off = std::snprintf(buf, sizeof buf, "%s{\"ms\": %.3f, \"executed_frac_fp32_peak\": %.4f, "
                            "\"split\": %d, \"lengths\": {", gi ? ", " : "", gms,
                            8192.0 * (double)block_steps / (gms / 1e3) / 1e12 / 157.3, split_n);
    for (size_t li = 0; li < A.lengths.size(); ++li) {
      const int L = A.lengths[li];
      const int c = (int)std::count(cur.begin(), cur.end(), L);
      off += std::snprintf(buf + off, sizeof buf - off, "%s\"%d\": %d", li ? ", " : "", L, c);
    }
    std::snprintf(buf + off, sizeof buf - off, "}}");
    per_gen += buf;
  }
  CHECK_HIP(hipStreamSynchronize(s));
  // matrix work the local searches executed (block steps of 4 x 2048 FLOP)
  const double flop = 8192.0 * (double)total_steps / A.gens;
  std::vector<double> fit(P);
  CHECK_HIP(hipMemcpy(fit.data(), d.fit, P * sizeof(double), hipMemcpyDeviceToHost));
  const double best = *std::min_element(fit.begin(), fit.end());
  int timeouts = 0;  // pair waits that gave up (byte 4 of the local-search workspace)
  CHECK_HIP(hipMemcpy(&timeouts, static_cast<char*>(d.ls_workspace) + 4, sizeof timeouts,
                      hipMemcpyDeviceToHost));
  std::printf("{\"workload\": \"uneven islands: %d candidates, n in {600,1200,2100,4096,8400} "
              "(initial draw), %d islands, %d Adam steps + mutation sigma=%.2f + tournament per "
              "generation, HIP graph\", \"generations\": %d, \"ms_per_generation\": %.3f, "
              "\"candidate_steps_per_s\": %.1f, \"local_search_executed_tflops\": %.2f, "
              "\"executed_frac_fp32_peak\": %.4f, \"split\": %d, \"pair_timeouts\": %d, "
              "\"best_F\": %.12f, \"per_generation\": [%s]}\n",
              P, A.islands, A.iters, A.sigma, A.gens, ms / A.gens,
              (double)steps * A.gens / (ms / 1e3), flop * A.gens / (ms / 1e3) / 1e12,
              flop * A.gens / (ms / 1e3) / 1e12 / 157.3, A.split, timeouts, best, per_gen.c_str());
  CHECK_SIDON(sidon_generation_destroy(g));
  return 0;
}

// ---------------------------------------------- Frank-Wolfe + split ladder --
int fw_ladder(const Args& A) {
  const std::vector<int>& L = A.ladder;
  const int P = A.P;
  std::mt19937_64 rng(A.seed);
  int n = L[0];
  std::vector<float> h((size_t)P * n);
  for (int p = 0; p < P; ++p) structured_prior(rng, n, &h[(size_t)p * n]);
  float* pop = dalloc<float>((size_t)P * L.back());
  float* nxt = dalloc<float>((size_t)P * L.back());
  double* F = dalloc<double>(P);
  double* F2 = dalloc<double>(P);
  int* amb = dalloc<int>(P);
  CHECK_HIP(hipMemcpy(pop, h.data(), h.size() * sizeof(float), hipMemcpyHostToDevice));
  hipStream_t s;
  CHECK_HIP(hipStreamCreate(&s));  // blocking: orders with the null-stream copies below
  for (size_t lvl = 0; lvl < L.size(); ++lvl) {
    n = L[lvl];
    hipEvent_t e0, e1;
    CHECK_HIP(hipEventCreate(&e0));
    CHECK_HIP(hipEventCreate(&e1));
    CHECK_SIDON(sidon_eval_f64(pop, n, P, F, nullptr, 1e-9, s));
    CHECK_HIP(hipStreamSynchronize(s));
    unsigned long long steps = 0;
    CHECK_SIDON(sidon_fw_block_steps(&steps, 1));
    CHECK_HIP(hipEventRecord(e0, s));
    const int rc = sidon_frank_wolfe(pop, nullptr, n, P, A.fw_rounds, A.fw_iters, A.nT, A.t_max, s);
    CHECK_HIP(hipEventRecord(e1, s));
    if (rc == SIDON_ERR_LDS) {
      std::printf("{\"level\": %d, \"frank_wolfe\": \"does not fit in LDS\"}\n", n);
    } else {
      CHECK_SIDON(rc);
    }
    CHECK_HIP(hipEventSynchronize(e1));
    float ms = 0.0f;
    CHECK_HIP(hipEventElapsedTime(&ms, e0, e1));
    CHECK_SIDON(sidon_eval_f64(pop, n, P, F2, nullptr, 1e-9, s));
    CHECK_HIP(hipStreamSynchronize(s));
    CHECK_SIDON(sidon_fw_block_steps(&steps, 1));
    const double exec_tflops = 8192.0 * (double)steps / (ms / 1e3) / 1e12;
    std::vector<double> f0(P), f1(P);
    CHECK_HIP(hipMemcpy(f0.data(), F, P * sizeof(double), hipMemcpyDeviceToHost));
    CHECK_HIP(hipMemcpy(f1.data(), F2, P * sizeof(double), hipMemcpyDeviceToHost));
    double m0 = 0, m1 = 0, b1 = 1e30;
    for (int p = 0; p < P; ++p) { m0 += f0[p] / P; m1 += f1[p] / P; b1 = std::min(b1, f1[p]); }
    // dense count: 4 n^2 FLOP per CP iteration; the dual correlation skips steps whose
    // multipliers are all zero, so this equivalent rate can exceed the matrix peak
    const double cp_flop = 4.0 * n * (double)n * A.fw_iters * A.fw_rounds * P;
    std::printf("{\"level\": %d, \"fw_rounds\": %d, \"cp_iters\": %d, \"ms\": %.1f, "
                "\"cp_iterations_per_s\": %.1f, \"cp_dense_equivalent_tflops\": %.2f, "
                "\"executed_block_steps\": %llu, \"executed_tflops\": %.2f, "
                "\"executed_frac_fp32_peak\": %.4f, \"mean_F_before\": %.9f, "
                "\"mean_F_after\": %.9f, \"best_F\": %.12f}\n",
                n, A.fw_rounds, A.fw_iters, ms, (double)A.fw_iters * A.fw_rounds * P / (ms / 1e3),
                cp_flop / (ms / 1e3) / 1e12, steps, exec_tflops, exec_tflops / 157.3, m0, m1, b1);
    if (lvl + 1 < L.size()) {
      if (L[lvl + 1] != 2 * n) { std::fprintf(stderr, "ladder must double\n"); return 2; }
      CHECK_SIDON(sidon_eval_exact(pop, n, P, F, nullptr, nullptr, nullptr, s));
      CHECK_SIDON(sidon_refine(pop, n, P, nxt, s));
      CHECK_SIDON(sidon_eval_exact(nxt, 2 * n, P, F2, nullptr, nullptr, amb, s));
      CHECK_HIP(hipStreamSynchronize(s));
      std::vector<double> a(P), b(P);
      std::vector<int> am(P);
      CHECK_HIP(hipMemcpy(a.data(), F, P * sizeof(double), hipMemcpyDeviceToHost));
      CHECK_HIP(hipMemcpy(b.data(), F2, P * sizeof(double), hipMemcpyDeviceToHost));
      CHECK_HIP(hipMemcpy(am.data(), amb, P * sizeof(int), hipMemcpyDeviceToHost));
      int same = 0, ambig = 0;
      for (int p = 0; p < P; ++p) { same += (a[p] == b[p]); ambig += am[p]; }
      std::printf("{\"split\": \"%d -> %d\", \"F_bit_exact\": %d, \"of\": %d, \"ambiguous\": %d}\n",
                  n, 2 * n, same, P, ambig);
      std::swap(pop, nxt);
    }
  }
  return 0;
}

// ---------------------------------------------------------- fp64 polish --
// The reference's polish of the incumbent (src/main.cu:798-846): K candidates
// (the incumbent and K-1 copies with bounded log-normal kicks, 0.003 (1 + k%8)),
// six fp64 passes of rising exponent p 200 4^q -> 30x that and falling step
// 0.01 0.45^q -> 2 % of that, the best re-evaluated in fp64 after each pass.
// Here the passes run through sidon_polish_f64, which spreads the K fp64
// searches over the whole device at any n.
void polish(const Args& A, std::vector<double>& bestA, double& bestF) {
  const int n = (int)bestA.size(), K = std::max(1, A.polish_top);
  std::mt19937_64 r2(A.seed ^ 0xABCDEFull);
  std::normal_distribution<double> N01(0.0, 1.0);
  std::vector<double> h((size_t)n * K);
  for (int k = 0; k < K; ++k) {
    double S = 0.0;
    for (int i = 0; i < n; ++i) {
      double x = bestA[i];
      if (k) x *= std::exp(0.003 * (1 + (k % 8)) * N01(r2));
      h[(size_t)k * n + i] = std::max(0.0, x);
      S += std::max(0.0, x);
    }
    for (int i = 0; i < n; ++i) h[(size_t)k * n + i] /= S;
  }
  double* pop = dalloc<double>(h.size());
  double* F = dalloc<double>(K);
  const size_t wsb = sidon_polish_f64_workspace(n, K);
  void* ws = nullptr;
  CHECK_HIP(hipMalloc(&ws, wsb));
  CHECK_HIP(hipMemcpy(pop, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice));
  hipStream_t s;
  CHECK_HIP(hipStreamCreate(&s));
  const int passes = 6, per = std::max(1, A.polish_iters / passes);
  const auto t0 = Clock::now();
  std::vector<double> f(K);
  for (int q = 0; q < passes; ++q) {
    const double pA = 200.0 * std::pow(4.0, q), pB = 30.0 * pA;
    const double lA = 0.010 * std::pow(0.45, q), lB = 0.02 * lA;
    CHECK_SIDON(sidon_polish_f64(pop, nullptr, nullptr, n, K, per, pA, pB, lA, lB, 0.9, 0.999, 0,
                                 ws, wsb, s));
    CHECK_SIDON(sidon_eval_f64_d(pop, n, K, F, nullptr, 1e-6, s));
    CHECK_HIP(hipMemcpyAsync(f.data(), F, K * sizeof(double), hipMemcpyDeviceToHost, s));
    CHECK_HIP(hipStreamSynchronize(s));
    const int bk = (int)(std::min_element(f.begin(), f.end()) - f.begin());
    if (f[bk] < bestF) {
      bestF = f[bk];
      CHECK_HIP(hipMemcpy(bestA.data(), pop + (size_t)bk * n, n * sizeof(double),
                          hipMemcpyDeviceToHost));
      save_vec(A.out + ".txt", bestA, bestF);
    }
    std::printf("{\"polish_pass\": %d, \"p\": [%.0f, %.0f], \"best_F\": %.15f}\n", q, pA, pB,
                f[bk]);
  }
  const double sec = seconds_since(t0);
  std::printf("{\"polish\": {\"n\": %d, \"candidates\": %d, \"iters\": %d, \"seconds\": %.2f, "
              "\"candidate_steps_per_s\": %.1f, \"best_F\": %.15f}}\n",
              n, K, per * passes, sec, (double)K * per * passes / sec, bestF);
  CHECK_HIP(hipFree(ws));
  CHECK_HIP(hipFree(pop));
  CHECK_HIP(hipFree(F));
  CHECK_HIP(hipStreamDestroy(s));
}

// --------------------------------------------------------- policy search --
// The reference's search loop (src/main.cu:445-451 level budgets, :586-769 the
// generation) over the C ABI: each generation runs the three anneal tiers of
// the local search (cold / medium / warm by the operator that made each
// candidate; every fourth island uses mirror descent), optional Frank-Wolfe
// rounds, the fp64 evaluation and support signatures; the host policy then
// plans the offspring (tabu, operator weights, tournaments, migration), the
// variation applies it, the incumbent is re-seeded every 10 generations and
// the operator weights are credited every 5.  Levels get 1.85^level shares of
// 82 % of the time budget (the rest is the polish's); a level ends when its
// share is spent or after --gens generations (0 = no cap); between levels
// every candidate is split n -> 2n (F preserved exactly).
int search(const Args& A) {
  std::vector<int> L = A.ladder;
  // --resume (main.cu:414-443): the saved vector becomes the incumbent at once,
  // and levels coarser than it are dropped (a split preserves F exactly, so
  // they cannot beat it)
  std::vector<double> seedVec;
  if (!A.resume.empty() && !load_vec(A.resume, seedVec)) {
    // the reference goes on without the vector (main.cu:424-426)
    std::fprintf(stderr, "could not read %s, ignoring\n", A.resume.c_str());
    seedVec.clear();
  }
  if (!seedVec.empty()) {
    size_t drop = 0;
    while (drop + 1 < L.size() && L[drop] < (int)seedVec.size()) ++drop;
    L.erase(L.begin(), L.begin() + drop);
    std::printf("{\"resumed\": \"%s\", \"n\": %zu, \"F\": %.15f, \"levels_dropped\": %zu}\n",
                A.resume.c_str(), seedVec.size(), host_F(seedVec), drop);
  }
  const int P = A.P, NI = A.islands, nmax = *std::max_element(L.begin(), L.end());
  if (P % NI != 0) { std::fprintf(stderr, "P must be divisible by islands\n"); return 2; }
  const int PI = P / NI;
  std::mt19937_64 rng(A.seed);
  hipStream_t s;
  CHECK_HIP(hipStreamCreate(&s));  // blocking: orders with the null-stream copies below
  float* pop = dalloc<float>((size_t)P * nmax);
  float* nxt = dalloc<float>((size_t)P * nmax);
  double* fit = dalloc<double>(P);
  int* nact = dalloc<int>(P);
  unsigned long long* sig = dalloc<unsigned long long>(P);
  int* nnz = dalloc<int>(P);
  int* dp1 = dalloc<int>(P); int* dp2 = dalloc<int>(P); int* dop = dalloc<int>(P);
  float* dst = dalloc<float>(P);
  int* skip = dalloc<int>(P);
  int* mode = dalloc<int>(P);
  // Adam moments out of each launch, which also serve as the launch's scratch
  // for the correlation output (the reference allocates g.mom and g.vel for
  // every level, main.cu:107-144)
  float* mom = dalloc<float>((size_t)P * nmax);
  float* vel = dalloc<float>((size_t)P * nmax);
  std::vector<int> hmode(P);
  for (int p = 0; p < P; ++p) hmode[p] = (p / PI) % 4 == 3 ? 1 : 0;  // mirror descent island
  CHECK_HIP(hipMemcpy(mode, hmode.data(), P * sizeof(int), hipMemcpyHostToDevice));
  {  // every slot of the first level starts from a fresh host seed (main.cu:502-551)
    std::vector<float> h((size_t)P * L[0]);
    for (int p = 0; p < P; ++p) host_seed(rng, L[0], &h[(size_t)p * L[0]]);
    CHECK_HIP(hipMemcpy(pop, h.data(), h.size() * sizeof(float), hipMemcpyHostToDevice));
  }
  sidon_policy* pol = nullptr;
  CHECK_SIDON(sidon_policy_create(P, NI, A.elite, A.seed, &pol));
  std::vector<int> curOp(P, 9);  // fresh seeds: the cold tier (up to 18 % of P), then medium
  std::vector<double> hfit(P), cfit(P);
  std::vector<unsigned long long> hsig(P);
  std::vector<int> hnnz(P), hnact(P), p1(P), p2(P), op(P), tier(P), hskip(P), msrc(NI), mdst(NI);
  std::vector<float> hst(P);
  std::vector<double> lvlBudget(L.size());
  double tot = 0.0;
  for (size_t l = 0; l < L.size(); ++l) tot += (lvlBudget[l] = std::pow(1.85, (double)l));
  for (auto& x : lvlBudget) x = x / tot * A.seconds * 0.82;
  const auto t0 = Clock::now();
  double bestF = 1e30;
  std::vector<double> bestA;
  int bestN = 0;
  if (!seedVec.empty()) {
    bestA = seedVec;
    bestF = host_F(seedVec);
    bestN = (int)seedVec.size();
    CHECK_SIDON(sidon_policy_adopt(pol, bestF));
  }
  long long lsSteps = 0, evals = 0, fwRuns = 0;
  double flop = 0.0;  // the reference's effort count (main.cu:624-642): 4 n^2 a local-search step
  for (size_t lvl = 0; lvl < L.size(); ++lvl) {
    const int n = L[lvl];
    CHECK_SIDON(sidon_policy_new_level(pol));
    // the incumbent in slot 0 and the resumed vector in slot 1, resampled to n
    // (main.cu:553-565), both cloned into the next generation's warm tier
    std::vector<float> row(n);
    const std::vector<double>* inject[2] = {&bestA, &seedVec};
    for (int k = 0; k < 2 && k < P; ++k) {
      if (inject[k]->empty()) continue;
      resample(*inject[k], n, row.data());
      CHECK_HIP(hipMemcpy(pop + (size_t)k * n, row.data(), n * sizeof(float),
                          hipMemcpyHostToDevice));
      curOp[k] = 20;
    }
    const auto tl = Clock::now();
    int gen = 0;
    const double n2 = (double)n * n;
    bool fwOk = true;
    while (seconds_since(tl) < lvlBudget[lvl] && (A.gens == 0 || gen < A.gens)) {
      // ---- three anneal tiers (main.cu:598-629) ----
      int nCold = 0, nMed = 0;
      CHECK_SIDON(sidon_policy_tiers(curOp.data(), P, tier.data(), &nCold, &nMed));
      const float pTop = (float)A.pmax;
      auto tier_run = [&](int t, int iters, float pa, float pb, float la, float lb) {
        for (int p = 0; p < P; ++p) hskip[p] = tier[p] != t;
        CHECK_HIP(hipMemcpy(skip, hskip.data(), P * sizeof(int), hipMemcpyHostToDevice));
        CHECK_SIDON(sidon_local_search(pop, mom, vel, skip, mode, n, P, iters, pa, pb,
                                       la, lb, 0.9f, 0.999f, 0, s));
        for (int p = 0; p < P; ++p) lsSteps += (long long)(hskip[p] == 0) * iters;
        flop += 4.0 * n2 * (double)iters * (double)(t == 0 ? nCold : t == 1 ? nMed : P - nCold - nMed);
      };
      if (nCold) tier_run(0, A.cold_iters, 6.0f, pTop, 0.085f, 0.0015f);
      if (nMed) tier_run(1, std::max(1, A.cold_iters / 3), 40.0f, pTop, 0.040f, 0.0010f);
      if (nCold + nMed < P) tier_run(2, A.iters, 400.0f, 3.0f * pTop, 0.012f, 0.0003f);
      if (A.fw_rounds > 0 && fwOk) {  // skipped for good where it does not fit (main.cu:633-639)
        const int rc = sidon_frank_wolfe(pop, nullptr, n, P, A.fw_rounds, A.fw_iters, A.nT,
                                         A.t_max, s);
        if (rc == SIDON_ERR_LDS) {
          fwOk = false;
        } else {
          CHECK_SIDON(rc);
          fwRuns += (long long)A.fw_rounds * P;
          flop += (double)P * A.fw_rounds * (4.0 * n2 * A.fw_iters + 26.0 * n2);
        }
      }
      // ---- evaluate, signatures, policy (main.cu:640-735) ----
      CHECK_SIDON(sidon_eval_f64(pop, n, P, fit, nact, 1e-4, s));
      CHECK_SIDON(sidon_support_sig(pop, n, P, 0.02f, sig, nnz, s));
      CHECK_HIP(hipMemcpyAsync(hfit.data(), fit, P * sizeof(double), hipMemcpyDeviceToHost, s));
      CHECK_HIP(hipMemcpyAsync(hsig.data(), sig, P * sizeof(unsigned long long),
                               hipMemcpyDeviceToHost, s));
      CHECK_HIP(hipMemcpyAsync(hnnz.data(), nnz, P * sizeof(int), hipMemcpyDeviceToHost, s));
      CHECK_HIP(hipMemcpyAsync(hnact.data(), nact, P * sizeof(int), hipMemcpyDeviceToHost, s));
      CHECK_HIP(hipStreamSynchronize(s));
      evals += P;
      flop += 2.0 * n2 * P;
      int bi = 0, improved = 0;
      CHECK_SIDON(sidon_policy_observe(pol, hfit.data(), hsig.data(), &bi, &improved));
      if (improved) {
        std::vector<float> row(n);
        CHECK_HIP(hipMemcpy(row.data(), pop + (size_t)bi * n, n * sizeof(float),
                            hipMemcpyDeviceToHost));
        bestA.assign(row.begin(), row.end());
        bestF = hfit[bi];
        bestN = n;
        save_vec(A.out + ".txt", bestA, bestF);
      }
      int nmig = 0, reseed = -1;
      CHECK_SIDON(sidon_policy_plan(pol, gen, p1.data(), p2.data(), op.data(), hst.data(),
                                    msrc.data(), mdst.data(), &nmig, &reseed));
      for (int k = 0; k < nmig; ++k)  // ring migration: copy the vector before the variation
        CHECK_HIP(hipMemcpyAsync(pop + (size_t)mdst[k] * n, pop + (size_t)msrc[k] * n,
                                 n * sizeof(float), hipMemcpyDeviceToDevice, s));
      CHECK_HIP(hipMemcpyAsync(dp1, p1.data(), P * sizeof(int), hipMemcpyHostToDevice, s));
      CHECK_HIP(hipMemcpyAsync(dp2, p2.data(), P * sizeof(int), hipMemcpyHostToDevice, s));
      CHECK_HIP(hipMemcpyAsync(dop, op.data(), P * sizeof(int), hipMemcpyHostToDevice, s));
      CHECK_HIP(hipMemcpyAsync(dst, hst.data(), P * sizeof(float), hipMemcpyHostToDevice, s));
      CHECK_SIDON(sidon_variation(pop, nxt, n, P, dp1, dp2, dop, dst, A.seed,
                                  (unsigned long long)(gen + 1000 * (int)lvl), s));
      std::swap(pop, nxt);
      curOp = op;
      if (reseed >= 0 && bestN == n) {  // keep the incumbent in the gene pool (main.cu:747-755)
        std::vector<float> row(bestA.begin(), bestA.end());
        CHECK_HIP(hipMemcpyAsync(pop + (size_t)reseed * n, row.data(), n * sizeof(float),
                                 hipMemcpyHostToDevice, s));
        CHECK_HIP(hipStreamSynchronize(s));
        curOp[reseed] = 20;
      }
      if (gen > 0 && gen % 5 == 0) {  // credit assignment (main.cu:757-769)
        CHECK_SIDON(sidon_eval_f64(pop, n, P, fit, nact, 1e-4, s));
        CHECK_HIP(hipMemcpyAsync(cfit.data(), fit, P * sizeof(double), hipMemcpyDeviceToHost, s));
        CHECK_HIP(hipStreamSynchronize(s));
        evals += P;
        flop += 2.0 * n2 * P;
        CHECK_SIDON(sidon_policy_credit(pol, cfit.data()));
      }
      if (A.verbose && (gen % 10 == 0 || improved)) {  // progress records (main.cu:771-783)
        double mean = 0.0, lb = 0.0;
        for (double x : hfit) mean += x / P;
        int cells = 0, since = 0;
        CHECK_SIDON(sidon_policy_stats(pol, nullptr, nullptr, &lb, &cells, &since));
        std::printf("{\"stat\": {\"t\": %.1f, \"n\": %d, \"gen\": %d, \"best\": %.15f, "
                    "\"level_best\": %.15f, \"mean\": %.9f, \"evals\": %lld, "
                    "\"ls_steps\": %lld, \"fw\": %lld, \"cells\": %d, \"peaks\": %d, "
                    "\"nnz\": %d, \"cold\": %d, \"medium\": %d, \"since_improved\": %d%s}}\n",
                    seconds_since(t0), n, gen, bestF, lb, mean, evals, lsSteps, fwRuns, cells,
                    hnact[bi], hnnz[bi], nCold, nMed, since, improved ? ", \"improved\": true" : "");
        // the reference's key=value record, for monitors written against it
        std::printf("STAT t=%.1f n=%d gen=%d best=%.15f lvl=%.15f mean=%.9f evals=%lld "
                    "lsiters=%lld fw=%lld cells=%d pflop=%.6f peaks=%d nnz=%d\n",
                    seconds_since(t0), n, gen, bestF, lb, mean, evals, lsSteps, fwRuns, cells,
                    flop / 1e15, hnact[bi], hnnz[bi]);
      }
      ++gen;
    }
    double w[SIDON_POLICY_OPS];
    int cells = 0;
    CHECK_SIDON(sidon_policy_stats(pol, w, nullptr, nullptr, &cells, nullptr));
    std::printf("{\"level\": %d, \"generations\": %d, \"seconds\": %.1f, \"best_F\": %.15f, "
                "\"cells\": %d, \"op_weights\": [", n, gen, seconds_since(t0), bestF, cells);
    for (int o = 0; o < SIDON_POLICY_OPS; ++o) std::printf("%s%.3f", o ? ", " : "", w[o]);
    std::printf("]}\n");
    if (lvl + 1 < L.size()) {  // carry the population to the next length (main.cu:489-501)
      const int m = L[lvl + 1];
      if (m == 2 * n) {  // cell split: F preserved exactly
        CHECK_SIDON(sidon_refine(pop, n, P, nxt, s));
        CHECK_HIP(hipStreamSynchronize(s));
      } else {  // any other ratio: nearest-neighbour resample a'_i = a_{min(n-1, i n / m)}
        std::vector<float> h((size_t)P * n), r((size_t)P * m);
        CHECK_HIP(hipMemcpy(h.data(), pop, h.size() * sizeof(float), hipMemcpyDeviceToHost));
        for (int p = 0; p < P; ++p)
          for (int i = 0; i < m; ++i)
            r[(size_t)p * m + i] =
                h[(size_t)p * n + std::min<long long>(n - 1, (long long)i * n / m)];
        CHECK_HIP(hipMemcpy(nxt, r.data(), r.size() * sizeof(float), hipMemcpyHostToDevice));
      }
      // the last 20 % of the slots start afresh, to keep the population diverse
      std::vector<float> fresh(m);
      for (int p = 0; p < P; ++p) {
        if (p >= 0.8 * P) {
          host_seed(rng, m, fresh.data());
          CHECK_HIP(hipMemcpy(nxt + (size_t)p * m, fresh.data(), m * sizeof(float),
                              hipMemcpyHostToDevice));
          curOp[p] = 9;   // cold / medium anneal
        } else {
          curOp[p] = 20;  // carried parents: warm refinement
        }
      }
      std::swap(pop, nxt);
      std::printf("{\"carry\": \"%d -> %d\", \"form\": \"%s\", \"fresh\": %d}\n", n, m,
                  m == 2 * n ? "cell split" : "nearest-neighbour resample",
                  P - (int)std::ceil(0.8 * P));
    }
  }
  sidon_policy_destroy(pol);
  CHECK_HIP(hipFree(pop)); CHECK_HIP(hipFree(nxt)); CHECK_HIP(hipFree(mom)); CHECK_HIP(hipFree(vel));
  if (A.polish_iters > 0 && !bestA.empty()) polish(A, bestA, bestF);
  const double hf = host_F(bestA);
  save_vec(A.out + ".txt", bestA, std::max(bestF, hf));
  save_certificate(A.out + ".json", bestA, hf);
  std::printf("{\"best_F\": %.15f, \"host_recheck\": %.15f, \"n\": %zu, \"bound\": \"C <= %.9f\", "
              "\"files\": [\"%s.txt\", \"%s.json\"], \"seconds\": %.1f, \"ls_steps\": %lld, "
              "\"evals\": %lld}\n",
              bestF, hf, bestA.size(), std::max(bestF, hf), A.out.c_str(), A.out.c_str(),
              seconds_since(t0), lsSteps, evals);
  return 0;
}

// ---------------------------------------------------------- graph search --
int search_graph(const Args& A) {
  const std::vector<int>& L = A.ladder;
  const int P = A.P, nmax = L.back();
  std::mt19937_64 rng(A.seed);
  hipStream_t s;
  CHECK_HIP(hipStreamCreate(&s));  // blocking: orders with the null-stream copies below
  std::vector<float> h((size_t)P * nmax, 0.0f);
  std::vector<int> ns(P, L[0]);
  for (int p = 0; p < P; ++p) structured_prior(rng, L[0], &h[(size_t)p * nmax]);
  sidon_generation_desc d{};
  d.pop_a = dalloc<float>(h.size()); d.pop_b = dalloc<float>(h.size());
  d.ns_a = dalloc<int>(P); d.ns_b = dalloc<int>(P);
  d.fit = dalloc<double>(P); d.nact = dalloc<int>(P);
  d.p1 = dalloc<int>(P); d.p2 = dalloc<int>(P); d.op = dalloc<int>(P);
  d.strength = dalloc<float>(P); d.order = dalloc<int>(P);
  d.gen = dalloc<unsigned long long>(1);
  d.ls_workspace_bytes = sidon_local_search_ragged_workspace(P, nmax);
  d.ls_workspace = dalloc<unsigned char>(d.ls_workspace_bytes);
  d.split = A.split;
  d.P = P; d.stride = nmax; d.islands = A.islands; d.elite = A.elite; d.migrate_every = 12;
  d.iters = A.iters; d.sigma = A.sigma; d.activeTol = 1e-4; d.seed = A.seed;
  const auto t0 = Clock::now();
  double bestF = 1e30;
  std::vector<double> bestA;
  float* tmp = dalloc<float>(h.size());
  for (size_t lvl = 0; lvl < L.size(); ++lvl) {
    const int n = L[lvl];
    if (lvl == 0) {
      CHECK_HIP(hipMemcpy(d.pop_a, h.data(), h.size() * sizeof(float), hipMemcpyHostToDevice));
    }
    std::fill(ns.begin(), ns.end(), n);
    CHECK_HIP(hipMemcpy(d.ns_a, ns.data(), P * sizeof(int), hipMemcpyHostToDevice));
    CHECK_HIP(hipMemcpy(d.ns_b, ns.data(), P * sizeof(int), hipMemcpyHostToDevice));
    d.nmax = n;
    // cold anneal at the first level, warm refinement after a split (main.cu:627-629)
    d.pStart = lvl == 0 ? 6.0f : 400.0f; d.pEnd = 2e4f;
    d.lrStart = lvl == 0 ? 0.085f : 0.012f; d.lrEnd = lvl == 0 ? 0.0015f : 0.0003f;
    sidon_generation* g = nullptr;
    CHECK_SIDON(sidon_generation_create(&d, &g, s));
    const double budget = A.seconds * (lvl + 1) / (L.size() * (L.size() + 1) / 2.0);
    const auto tl = Clock::now();
    int gens = 0;
    while (seconds_since(tl) < budget && (A.gens == 0 || gens < A.gens)) {
      CHECK_SIDON(sidon_generation_run(g, 1, s));
      ++gens;
      CHECK_HIP(hipStreamSynchronize(s));  // the clock sees the work
    }
    CHECK_HIP(hipStreamSynchronize(s));
    int parity = 0;
    CHECK_SIDON(sidon_generation_current(g, &parity));
    float* cur = parity ? d.pop_b : d.pop_a;
    // fp64 F of the current population (the fit buffer holds the pre-mutation values)
    CHECK_SIDON(sidon_eval_f64_ragged(cur, nmax, parity ? d.ns_b : d.ns_a, P, n, d.fit, d.nact,
                                      1e-4, s));
    CHECK_HIP(hipStreamSynchronize(s));
    std::vector<double> fit(P);
    CHECK_HIP(hipMemcpy(fit.data(), d.fit, P * sizeof(double), hipMemcpyDeviceToHost));
    const int bi = (int)(std::min_element(fit.begin(), fit.end()) - fit.begin());
    std::vector<float> row(n);
    CHECK_HIP(hipMemcpy(row.data(), cur + (size_t)bi * nmax, n * sizeof(float),
                        hipMemcpyDeviceToHost));
    std::vector<double> a(row.begin(), row.end());
    const double hf = host_F(a);
    if (fit[bi] < bestF) { bestF = fit[bi]; bestA = a; }
    std::printf("{\"level\": %d, \"generations\": %d, \"seconds\": %.1f, \"best_F\": %.15f, "
                "\"host_recheck\": %.15f}\n", n, gens, seconds_since(t0), fit[bi], hf);
    CHECK_SIDON(sidon_generation_destroy(g));
    if (lvl + 1 < L.size()) {  // cell split every candidate into population A
      if (L[lvl + 1] != 2 * n) { std::fprintf(stderr, "ladder must double\n"); return 2; }
      CHECK_HIP(hipMemcpy2DAsync(tmp, n * sizeof(float), cur, nmax * sizeof(float),
                                 n * sizeof(float), P, hipMemcpyDeviceToDevice, s));
      float* packed = dalloc<float>((size_t)P * 2 * n);
      CHECK_SIDON(sidon_refine(tmp, n, P, packed, s));
      CHECK_HIP(hipMemcpy2DAsync(d.pop_a, nmax * sizeof(float), packed, 2 * n * sizeof(float),
                                 2 * n * sizeof(float), P, hipMemcpyDeviceToDevice, s));
      CHECK_HIP(hipStreamSynchronize(s));
      CHECK_HIP(hipFree(packed));
      // mutation streams are keyed by gen: level l counts from 1000 l, as the
      // reference keys its variation by gen + 1000 lvl (main.cu:741)
      const unsigned long long gen0 = 1000ull * (lvl + 1);
      CHECK_HIP(hipMemcpy(d.gen, &gen0, sizeof gen0, hipMemcpyHostToDevice));
    }
  }
  const double hf = host_F(bestA);
  save_vec(A.out + ".txt", bestA, std::max(bestF, hf));
  save_certificate(A.out + ".json", bestA, hf);
  std::printf("{\"best_F\": %.15f, \"host_recheck\": %.15f, \"n\": %zu, \"bound\": \"C <= %.9f\", "
              "\"files\": [\"%s.txt\", \"%s.json\"], \"seconds\": %.1f}\n",
              bestF, hf, bestA.size(), std::max(bestF, hf), A.out.c_str(), A.out.c_str(),
              seconds_since(t0));
  return 0;
}

// ------------------------------------------------------------ benchmark --
// The reference's kernel micro-benchmark (src/main.cu:294-320): per length, a
// population of max(32, min(512, 2^22 / n)) candidates with heights U(0.2, 1),
// 10 untimed local-search steps, then 100 timed ones (p 8 -> 200, step
// 0.05 -> 0.005) and one fp64 evaluation, timed with events.  TFLOP/s on the
// reference's count (4 n^2 FLOP a candidate-step) and on the matrix work the
// kernel executed; n = 8400 follows the reference's lengths.
int bench_kernels() {
  hipStream_t s;
  CHECK_HIP(hipStreamCreate(&s));
  hipEvent_t e0, e1, e2;
  CHECK_HIP(hipEventCreate(&e0));
  CHECK_HIP(hipEventCreate(&e1));
  CHECK_HIP(hipEventCreate(&e2));
  for (int n : {128, 256, 512, 1024, 2048, 4096, 8400}) {
    const int P = std::max(32, std::min(512, (4 << 20) / n));
    std::vector<float> h((size_t)n * P);
    std::mt19937 rng(1);
    std::uniform_real_distribution<float> U(0.2f, 1.0f);
    for (auto& x : h) x = U(rng);
    float* pop = dalloc<float>(h.size());
    float* mom = dalloc<float>(h.size());  // moments' buffers, as the search passes them
    float* vel = dalloc<float>(h.size());
    double* F = dalloc<double>(P);
    CHECK_HIP(hipMemcpy(pop, h.data(), h.size() * sizeof(float), hipMemcpyHostToDevice));
    const int it = 100;
    CHECK_SIDON(sidon_local_search(pop, mom, vel, nullptr, nullptr, n, P, 10, 8.0f, 8.0f,
                                   0.01f, 0.01f, 0.9f, 0.999f, 0, s));
    CHECK_HIP(hipStreamSynchronize(s));
    unsigned long long steps = 0;
    CHECK_SIDON(sidon_ls_block_steps(&steps, 1));
    CHECK_HIP(hipEventRecord(e0, s));
    CHECK_SIDON(sidon_local_search(pop, mom, vel, nullptr, nullptr, n, P, it, 8.0f,
                                   200.0f, 0.05f, 0.005f, 0.9f, 0.999f, 0, s));
    CHECK_HIP(hipEventRecord(e1, s));
    CHECK_SIDON(sidon_eval_f64(pop, n, P, F, nullptr, 1e-4, s));
    CHECK_HIP(hipEventRecord(e2, s));
    CHECK_HIP(hipEventSynchronize(e2));
    CHECK_SIDON(sidon_ls_block_steps(&steps, 1));
    float ls = 0.0f, ev = 0.0f;
    CHECK_HIP(hipEventElapsedTime(&ls, e0, e1));
    CHECK_HIP(hipEventElapsedTime(&ev, e1, e2));
    const double ref_flop = 2.0 * P * it * 2.0 * (double)n * n;
    std::printf("{\"bench\": {\"n\": %d, \"P\": %d, \"lds_kb\": %lld, \"ls_ms\": %.3f, "
                "\"ms_per_iter\": %.4f, \"candidate_steps_per_s\": %.1f, "
                "\"tflops_reference_count\": %.2f, \"tflops_executed\": %.2f, "
                "\"eval_f64_ms\": %.3f}}\n",
                n, P, sidon_lds_bytes(0, n, 1024) / 1024, ls, ls / it,
                (double)P * it / (ls / 1e3), ref_flop / (ls / 1e3) / 1e12,
                8192.0 * (double)steps / (ls / 1e3) / 1e12, ev);
    CHECK_HIP(hipFree(pop));
    CHECK_HIP(hipFree(mom));
    CHECK_HIP(hipFree(vel));
    CHECK_HIP(hipFree(F));
  }
  CHECK_HIP(hipEventDestroy(e0));
  CHECK_HIP(hipEventDestroy(e1));
  CHECK_HIP(hipEventDestroy(e2));
  CHECK_HIP(hipStreamDestroy(s));
  return 0;
}

void usage() {
  std::printf(
      "sidon_search [mode] [options]     (no mode: the search, with the reference's defaults)\n"
      "  --ladder 64,128,...   multiresolution ladder (default 64..2048); a level twice\n"
      "                        the last is carried by the cell split, any other ratio by\n"
      "                        nearest-neighbour resampling (--fw-ladder, --search-graph:\n"
      "                        doubling only)\n"
      "  --pop N               population size (default 256)\n"
      "  --islands N           number of islands (default 8)\n"
      "  --elite N             elites kept per island (default 2)\n"
      "  --ls-iters N          warm local-search steps per generation (default 4000)\n"
      "  --cold-iters N        cold local-search steps (default 20000; medium tier: N/3)\n"
      "  --pmax P              top of the fp32 exponent anneal (default 20000)\n"
      "  --fw-rounds N         Frank-Wolfe rounds per generation (default 6, 0 = off)\n"
      "  --fw-iters N          Chambolle-Pock iterations per round (default 500)\n"
      "  --gens N              generations per level at most (0 = time driven, default)\n"
      "  --seconds S           total wall-clock budget (default 600; 18 %% for the polish)\n"
      "  --seed S              RNG seed (default 20260905)\n"
      "  --out PREFIX          output prefix: PREFIX.txt and PREFIX.json (default 'best')\n"
      "  --resume FILE         seed the run from a saved vector\n"
      "  --polish-iters N      fp64 polish steps, over six passes (default 40000, 0 = off)\n"
      "  --polish-top N        candidates in the polish (default 8)\n"
      "  --selftest            verify the objective and kernels against known values\n"
      "  --bench               micro-benchmark the kernels and exit\n"
      "  --quiet               no per-generation progress records\n"
      "  -h, --help\n"
      "modes of this build (the options above apply where they make sense):\n"
      "  --search              the search (the default)\n"
      "  --uneven              island generations of uneven n as HIP graphs\n"
      "                        (defaults --pop 512 --ls-iters 200 --gens 3;\n"
      "                        --lengths 600,1200,2100,4096,8400 --sigma 0.05 --no-tune\n"
      "                        --no-split: no candidate shared by two workgroups)\n"
      "  --fw-ladder           Frank-Wolfe rounds + exact n -> 2n splits per level\n"
      "                        (defaults --ladder 1050,2100,4200,8400 --fw-rounds 50\n"
      "                        --fw-iters 2000)\n"
      "  --search-graph        search on the device generation graph (operator 2)\n"
      "                        (defaults --ladder 256,512,1024,2048 --ls-iters 200 --seconds 60)\n"
      "aliases: --P = --pop, --levels = --ladder, --iters = --ls-iters, --cp-iters = --fw-iters\n");
}

// Per-mode defaults for the options not given on the command line.
void mode_defaults(Args& A, const std::vector<std::string>& given) {
  auto unset = [&](const char* o) {
    return std::find(given.begin(), given.end(), std::string(o)) == given.end();
  };
  if (A.mode == "--uneven") {
    if (unset("--pop")) A.P = 512;
    if (unset("--ls-iters")) A.iters = 200;
    if (unset("--gens")) A.gens = 3;
  } else if (A.mode == "--fw-ladder") {
    if (unset("--ladder")) A.ladder = {1050, 2100, 4200, 8400};
    if (unset("--fw-rounds")) A.fw_rounds = 50;
    if (unset("--fw-iters")) A.fw_iters = 2000;
  } else if (A.mode == "--search-graph") {
    if (unset("--ladder")) A.ladder = {256, 512, 1024, 2048};
    if (unset("--ls-iters")) A.iters = 200;
    if (unset("--seconds")) A.seconds = 60.0;
  }
}

// Checks that need no device; a message and false on the first failure.
bool check_args(const Args& A) {
  auto fail = [](const char* m) { std::fprintf(stderr, "%s\n", m); return false; };
  if (A.P < 1 || A.islands < 1) return fail("pop and islands must be positive");
  if (A.P % A.islands != 0) return fail("pop must be divisible by islands");
  if (A.elite < 0 || A.iters < 0 || A.cold_iters < 0 || A.gens < 0 || A.fw_rounds < 0 ||
      A.fw_iters < 0 || A.polish_iters < 0 || A.polish_top < 0)
    return fail("counts must not be negative");
  if (!(A.seconds >= 0.0)) return fail("seconds must not be negative");
  // this build's own ladder modes split cells between levels; the search
  // carries any ratio (nearest-neighbour resample), as the reference does
  if (A.mode == "--search-graph" || A.mode == "--fw-ladder")
    for (size_t l = 1; l < A.ladder.size(); ++l)
      if (A.ladder[l] != 2 * A.ladder[l - 1])
        return fail("--search-graph and --fw-ladder need each ladder level twice the one "
                    "before (n -> 2n cell split)");
  for (int n : A.ladder)
    if (n < 1) return fail("ladder lengths must be positive");
  return true;
}

}  // namespace

int main(int argc, char** argv) {
  setvbuf(stdout, nullptr, _IONBF, 0);
  Args A;
  std::vector<std::string> given;
  for (int i = 1; i < argc; ++i) {
    std::string s = argv[i];
    if (s == "--P") s = "--pop";  // this driver's earlier spellings
    else if (s == "--levels") s = "--ladder";
    else if (s == "--iters") s = "--ls-iters";
    else if (s == "--cp-iters") s = "--fw-iters";
    given.push_back(s);
    auto next = [&]() -> const char* {
      if (i + 1 >= argc) {
        std::fprintf(stderr, "option %s needs a value\n", s.c_str());
        usage();
        std::exit(1);
      }
      return argv[++i];
    };
    auto list = [&](std::vector<int>& v) {
      const char* x = next();
      if (!parse_list(x, v)) {
        std::fprintf(stderr, "bad list for %s: %s\n", s.c_str(), x);
        std::exit(1);
      }
    };
    if (s == "--search" || s == "--selftest" || s == "--bench" || s == "--uneven" ||
        s == "--fw-ladder" || s == "--search-graph")
      A.mode = s;
    else if (s == "--ladder") list(A.ladder);
    else if (s == "--pop") A.P = std::atoi(next());
    else if (s == "--islands") A.islands = std::atoi(next());
    else if (s == "--elite") A.elite = std::atoi(next());
    else if (s == "--ls-iters") A.iters = std::atoi(next());
    else if (s == "--cold-iters") A.cold_iters = std::atoi(next());
    else if (s == "--pmax") A.pmax = std::atof(next());
    else if (s == "--fw-rounds") A.fw_rounds = std::atoi(next());
    else if (s == "--fw-iters") A.fw_iters = std::atoi(next());
    else if (s == "--gens") A.gens = std::atoi(next());
    else if (s == "--seconds") A.seconds = std::atof(next());
    else if (s == "--seed") A.seed = std::strtoull(next(), nullptr, 10);
    else if (s == "--out") A.out = next();
    else if (s == "--resume") A.resume = next();
    else if (s == "--polish-iters") A.polish_iters = std::atoi(next());
    else if (s == "--polish-top") A.polish_top = std::atoi(next());
    else if (s == "--quiet") A.verbose = 0;
    else if (s == "--sigma") A.sigma = (float)std::atof(next());
    else if (s == "--lengths") list(A.lengths);
    else if (s == "--no-tune") A.tune = false;
    else if (s == "--no-split") A.split = 0;
    else if (s == "-h" || s == "--help") { usage(); return 0; }
    else {
      std::fprintf(stderr, "unknown option %s\n", s.c_str());
      usage();
      return 1;
    }
  }
  mode_defaults(A, given);
  if (!check_args(A)) return 1;

  CHECK_SIDON(sidon_init(-1));
  int dev = 0;
  hipDeviceProp_t prop;
  CHECK_HIP(hipGetDevice(&dev));
  CHECK_HIP(hipGetDeviceProperties(&prop, dev));
  std::printf("GPU: %s  %s  %d CUs  LDS per workgroup %d B\n",
              prop.name[0] ? prop.name : "(unnamed)", prop.gcnArchName, prop.multiProcessorCount,
              sidon_lds_limit());
  if (A.mode == "--selftest") return selftest();
  if (A.mode == "--bench") return bench_kernels();
  if (A.mode == "--search") {  // every level must fit the fused search (main.cu:390-399)
    for (int n : A.ladder) {
      const long long need = sidon_lds_bytes(0, n, 1024);
      if (need < 0 || need > sidon_lds_limit()) {
        std::fprintf(stderr, "n=%d needs %lld B of LDS, the device offers %d\n", n, need,
                     sidon_lds_limit());
        return 1;
      }
    }
  }
  if (A.mode == "--uneven") return uneven(A);
  if (A.mode == "--fw-ladder") return fw_ladder(A);
  if (A.mode == "--search-graph") return search_graph(A);
  return search(A);
}
