// generation.hip -- host side: one memetic generation captured as a HIP graph,
// and the local-search autotuner.
//
// A generation of the island search (reference: the body of the generation
// loop, src/main.cu:586-785, reduced to the operators the configurations use)
// is, on the device and with no host round trip:
//   order candidates longest-first -> fused local search (persistent, uneven
//   lengths) -> fp64 F -> island ranking / migration / elites / tournaments ->
//   Gaussian multiplicative mutation into the other buffer -> gen += 1.
// Two graphs are captured (population A -> B and B -> A) and replayed in turn,
// so a run of G generations is G graph launches.  Every kernel in the
// sequence reads its generation number from device memory and clears its
// work-queue counter and pair flags with a kernel, so replays need no host
// updates and no memset nodes.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstring>
#include <new>

#include "runtime.hpp"

using namespace sidon;

struct sidon_generation {
  sidon_generation_desc d;
  hipGraph_t graph[2] = {nullptr, nullptr};
  hipGraphExec_t exec[2] = {nullptr, nullptr};
  unsigned long long replays = 0;
};

namespace {

int enqueue_generation(const sidon_generation_desc& d, int parity, hipStream_t s) {
  float* src = parity ? d.pop_b : d.pop_a;
  float* dst = parity ? d.pop_a : d.pop_b;
  int* ns_src = parity ? d.ns_b : d.ns_a;
  int* ns_dst = parity ? d.ns_a : d.ns_b;
  int rc;
  if ((rc = sidon_order_by_length(ns_src, d.P, d.order, s))) return rc;
  if ((rc = sidon_local_search_ragged(src, d.stride, ns_src, d.P, d.nmax, d.order, nullptr,
                                      nullptr, d.iters, d.pStart, d.pEnd, d.lrStart, d.lrEnd,
                                      0.9f, 0.999f, 0, d.split, d.ls_workspace,
                                      (size_t)d.ls_workspace_bytes, s)))
    return rc;
  if ((rc = sidon_eval_f64_ragged(src, d.stride, ns_src, d.P, d.nmax, d.fit, d.nact,
                                  d.activeTol, s)))
    return rc;
  if ((rc = sidon_select(d.fit, src, ns_src, d.stride, d.P, d.islands, d.elite, d.migrate_every,
                         d.sigma, d.seed, d.gen, d.p1, d.p2, d.op, d.strength, s)))
    return rc;
  if ((rc = sidon_variation_ragged(src, ns_src, dst, ns_dst, d.stride, d.P, d.p1, d.p2, d.op,
                                   d.strength, d.seed, d.gen, s)))
    return rc;
  return sidon_advance_gen(d.gen, 1, s);
}

bool desc_ok(const sidon_generation_desc& d) {
  return d.pop_a && d.pop_b && d.ns_a && d.ns_b && d.fit && d.nact && d.p1 && d.p2 && d.op &&
         d.strength && d.order && d.gen && d.ls_workspace && d.P > 0 && d.islands > 0 && d.P % d.islands == 0 &&
         d.nmax > 0 && d.stride >= d.nmax && d.iters >= 0 && d.elite >= 0 &&
         d.elite <= d.P / d.islands && (d.split == 0 || d.split == 1) &&
         d.ls_workspace_bytes >= sidon_local_search_ragged_workspace(d.P, d.stride);
}

}  // namespace

extern "C" int sidon_generation_create(const sidon_generation_desc* desc,
                                       sidon_generation** out, hipStream_t stream) {
  if (!desc || !out || !desc_ok(*desc)) return SIDON_ERR_ARG;
  *out = nullptr;
  int rc = SIDON_OK;
  if (!device_state(&rc)) return rc;
  sidon_generation* g = new (std::nothrow) sidon_generation();
  if (!g) return SIDON_ERR_HIP;
  g->d = *desc;
  hipStream_t cap = stream;
  bool own = false;
  if (!cap) {  // the null stream cannot be captured: use a private one
    if (hipStreamCreateWithFlags(&cap, hipStreamNonBlocking) != hipSuccess) {
      delete g;
      return SIDON_ERR_HIP;
    }
    own = true;
  }
  for (int parity = 0; parity < 2 && rc == SIDON_OK; ++parity) {
    hipError_t e = hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) { set_last_error("hipStreamBeginCapture", e); rc = SIDON_ERR_HIP; break; }
    rc = enqueue_generation(g->d, parity, cap);
    hipGraph_t graph = nullptr;
    e = hipStreamEndCapture(cap, &graph);
    if (rc == SIDON_OK && e != hipSuccess) { set_last_error("hipStreamEndCapture", e); rc = SIDON_ERR_HIP; }
    if (rc != SIDON_OK) { if (graph) (void)hipGraphDestroy(graph); break; }
    g->graph[parity] = graph;
    e = hipGraphInstantiate(&g->exec[parity], graph, nullptr, nullptr, 0);
    if (e != hipSuccess) { set_last_error("hipGraphInstantiate", e); rc = SIDON_ERR_HIP; }
  }
  if (own) (void)hipStreamDestroy(cap);
  if (rc != SIDON_OK) {
    sidon_generation_destroy(g);
    return rc;
  }
  *out = g;
  return SIDON_OK;
}

extern "C" int sidon_generation_run(sidon_generation* g, int generations, hipStream_t stream) {
  if (!g || generations < 0) return SIDON_ERR_ARG;
  for (int k = 0; k < generations; ++k) {
    const int parity = (int)(g->replays & 1ULL);
    SIDON_HIP_TRY(hipGraphLaunch(g->exec[parity], stream));
    ++g->replays;
  }
  return SIDON_OK;
}

extern "C" int sidon_generation_current(const sidon_generation* g, int* parity) {
  if (!g || !parity) return SIDON_ERR_ARG;
  *parity = (int)(g->replays & 1ULL);  // 0: population in pop_a / ns_a, 1: pop_b / ns_b
  return SIDON_OK;
}

extern "C" int sidon_generation_enqueue(const sidon_generation_desc* desc, int parity,
                                        hipStream_t stream) {
  if (!desc || !desc_ok(*desc) || (parity != 0 && parity != 1)) return SIDON_ERR_ARG;
  return enqueue_generation(*desc, parity, stream);
}

extern "C" int sidon_generation_destroy(sidon_generation* g) {
  if (!g) return SIDON_OK;
  for (int p = 0; p < 2; ++p) {
    if (g->exec[p]) (void)hipGraphExecDestroy(g->exec[p]);
    if (g->graph[p]) (void)hipGraphDestroy(g->graph[p]);
  }
  delete g;
  return SIDON_OK;
}

// ------------------------------------------------------------ autotuner --
// The launch configurations the tuner tries, in the order of its trial table:
// waves per candidate (workgroup size), operand steps in flight, candidates a
// CU.  The built-in rule at every length is one of them.
static const int kTiles[SIDON_LS_TILES][3] = {
    {256, 2, 1}, {256, 4, 1}, {256, 2, 4}, {256, 4, 4}, {512, 2, 1},
    {512, 4, 1}, {512, 2, 2}, {512, 4, 2}, {1024, 2, 1}, {1024, 4, 1}};

extern "C" int sidon_ls_tile_config(int i, int* block, int* depth, int* per_cu) {
  if (i < 0 || i >= SIDON_LS_TILES) return SIDON_ERR_ARG;
  if (block) *block = kTiles[i][0];
  if (depth) *depth = kTiles[i][1];
  if (per_cu) *per_cu = kTiles[i][2];
  return SIDON_OK;
}

// Times the fused local search at length n in every configuration on the
// caller's scratch population (P x n, modified in place; restored from
// `reference` before each trial when given) and keeps the fastest as the rule
// for n.  The built-in configuration stays unless another beats it by more
// than 1% (two runs of one configuration differ by a few tenths of a percent).
extern "C" int sidon_tune_local_search(float* scratch, const float* reference, float* rows,
                                       int n, int P, int iters, int repeats, int* best_block,
                                       int* best_depth, int* best_per_cu, float* best_ms,
                                       float* trial_ms, hipStream_t stream) {
  if (!scratch || n < 1 || P < 1 || iters < 1 || repeats < 1) return SIDON_ERR_ARG;
  (void)sidon_set_ls_tile(n, 0, 0, 0);
  int dblock = 0, ddepth = 0, dper = 0;
  (void)sidon_get_ls_tile(n, &dblock, &ddepth, &dper);
  hipEvent_t e0, e1;
  SIDON_HIP_TRY(hipEventCreate(&e0));
  SIDON_HIP_TRY(hipEventCreate(&e1));
  float ms_of[SIDON_LS_TILES];
  int rc = SIDON_OK;
  const size_t bytes = (size_t)P * (size_t)n * sizeof(float);
  for (int c = 0; c < SIDON_LS_TILES && !rc; ++c) {
    ms_of[c] = -1.0f;
    if ((rc = sidon_set_ls_tile(n, kTiles[c][0], kTiles[c][1], kTiles[c][2]))) break;
    float tot = 0.0f;
    bool ok = true;
    for (int r = 0; r <= repeats && ok; ++r) {  // r = 0 warms up
      if (reference &&
          hipMemcpyAsync(scratch, reference, bytes, hipMemcpyDeviceToDevice, stream) != hipSuccess) {
        rc = SIDON_ERR_HIP;
        break;
      }
      (void)hipEventRecord(e0, stream);
      const int lrc = sidon_local_search(scratch, rows, rows ? rows + (size_t)P * (size_t)n : nullptr,
                                         nullptr, nullptr, n, P, iters,
                                         4.0f, 1e4f, 0.085f, 0.0015f, 0.9f, 0.999f, 0, stream);
      (void)hipEventRecord(e1, stream);
      if (lrc == SIDON_ERR_LDS || lrc == SIDON_ERR_UNSUPPORTED) { ok = false; break; }
      if (lrc) { rc = lrc; break; }
      if (hipEventSynchronize(e1) != hipSuccess) { rc = SIDON_ERR_HIP; break; }
      float ms = 0.0f;
      (void)hipEventElapsedTime(&ms, e0, e1);
      if (r > 0) tot += ms;
    }
    if (!rc && ok) ms_of[c] = tot / (float)repeats;
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  (void)sidon_set_ls_tile(n, 0, 0, 0);
  if (rc) return rc;
  bool any = false;
  for (int c = 0; c < SIDON_LS_TILES; ++c) any = any || ms_of[c] >= 0.0f;
  if (!any && sidon_lds_bytes(0, n, 1024) >= 0 &&
      sidon_lds_bytes(0, n, 1024) <= (long long)sidon_lds_limit()) {
    // above the on-chip form without rows: the launch has one configuration
    // (1024 threads, four steps in flight, the correlation output in the
    // caller's rows), which trials without rows cannot run
    if (best_block) *best_block = 1024;
    if (best_depth) *best_depth = 4;
    if (best_per_cu) *best_per_cu = 1;
    if (best_ms) *best_ms = -1.0f;
    if (trial_ms)
      for (int c = 0; c < SIDON_LS_TILES; ++c) trial_ms[c] = -1.0f;
    return SIDON_OK;
  }
  int dflt = -1, best = -1;
  for (int c = 0; c < SIDON_LS_TILES; ++c) {
    if (kTiles[c][0] == dblock && kTiles[c][1] == ddepth && kTiles[c][2] == dper) dflt = c;
    if (ms_of[c] >= 0.0f && (best < 0 || ms_of[c] < ms_of[best])) best = c;
  }
  if (best < 0) return SIDON_ERR_LDS;
  if (dflt >= 0 && ms_of[dflt] >= 0.0f && ms_of[best] > 0.99f * ms_of[dflt]) best = dflt;
  if (best != dflt)
    (void)sidon_set_ls_tile(n, kTiles[best][0], kTiles[best][1], kTiles[best][2]);
  if (best_block) *best_block = kTiles[best][0];
  if (best_depth) *best_depth = kTiles[best][1];
  if (best_per_cu) *best_per_cu = kTiles[best][2];
  if (best_ms) *best_ms = ms_of[best];
  if (trial_ms)
    for (int c = 0; c < SIDON_LS_TILES; ++c) trial_ms[c] = ms_of[c];
  return SIDON_OK;
}
