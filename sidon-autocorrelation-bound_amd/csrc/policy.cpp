// policy.cpp -- the island search policy of the reference's driver, as a
// host-only object behind the C ABI (no device calls; the driver moves the
// vectors).
//
// Reference: src/main.cu:461-472 (operator weights, tabu tables), :598-609
// (anneal tiers), :650-735 (best tracking, cell tables, island ranking, ring
// migration, offspring plan with tournaments, tabu override and stagnation
// restarts), :744-769 (best re-seed, credit assignment).  The random draws
// come from a std::mt19937_64 of the policy's own seed; islands are ranked by
// (fit, slot), so ties order the same way on every run.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <new>
#include <random>
#include <unordered_map>
#include <vector>

#include "../../include/sidon.h"

namespace {

constexpr int kOps = SIDON_POLICY_OPS;
constexpr int kClone = 20;
constexpr int kTabuTenure = 6;
constexpr int kDiversifying[7] = {5, 7, 9, 10, 8, 11, 12};

}  // namespace

struct sidon_policy {
  int P = 0, NI = 0, PI = 0, elite = 0;
  std::mt19937_64 rng;
  std::uniform_real_distribution<double> U01{0.0, 1.0};
  double opW[kOps], opUse[kOps], opWin[kOps];
  std::unordered_map<unsigned long long, int> tabuHits;
  std::unordered_map<unsigned long long, double> cellBest;
  double bestF = 1e30, lvlBest = 1e30;
  int sinceImp = 0;
  std::vector<double> fit, parentFit;
  std::vector<unsigned long long> sig;
  std::vector<int> p1, op;  // last plan, for the credit pass
  bool observed = false;
};

extern "C" int sidon_policy_create(int P, int islands, int elite, unsigned long long seed,
                                   sidon_policy** out) {
  if (!out || P < 1 || islands < 1 || P % islands != 0 || elite < 0 || elite > P / islands)
    return SIDON_ERR_ARG;
  sidon_policy* p = new (std::nothrow) sidon_policy;
  if (!p) return SIDON_ERR_ARG;
  p->P = P;
  p->NI = islands;
  p->PI = P / islands;
  p->elite = elite;
  p->rng.seed(seed);
  for (int o = 0; o < kOps; ++o) { p->opW[o] = 1.0; p->opUse[o] = 0.0; p->opWin[o] = 0.0; }
  p->opW[11] = 2.5;  // favour the structural prior early (main.cu:468)
  p->opW[12] = 1.8;
  p->opW[13] = 1.8;
  p->fit.assign(P, 1e30);
  p->parentFit.assign(P, 1e30);
  p->sig.assign(P, 0);
  p->p1.assign(P, 0);
  p->op.assign(P, kClone);
  *out = p;
  return SIDON_OK;
}

extern "C" void sidon_policy_destroy(sidon_policy* pol) { delete pol; }

extern "C" int sidon_policy_adopt(sidon_policy* pol, double f) {
  if (!pol || !(f > 0.0)) return SIDON_ERR_ARG;
  pol->bestF = std::min(pol->bestF, f);
  return SIDON_OK;
}

extern "C" int sidon_policy_new_level(sidon_policy* pol) {
  if (!pol) return SIDON_ERR_ARG;
  pol->lvlBest = 1e30;
  pol->sinceImp = 0;
  pol->observed = false;
  return SIDON_OK;
}

extern "C" int sidon_policy_observe(sidon_policy* pol, const double* fit,
                                    const unsigned long long* sig, int* best_slot,
                                    int* improved) {
  if (!pol || !fit || !sig) return SIDON_ERR_ARG;
  const int P = pol->P;
  int bi = 0;
  for (int p = 1; p < P; ++p)
    if (fit[p] < fit[bi]) bi = p;
  const bool imp = fit[bi] < pol->bestF - 1e-13;
  if (imp) pol->bestF = fit[bi];
  if (fit[bi] < pol->lvlBest - 1e-13) {
    pol->lvlBest = fit[bi];
    pol->sinceImp = 0;
  } else {
    ++pol->sinceImp;
  }
  for (int p = 0; p < P; ++p) {  // a cell revisited without improving gains a tabu hit
    auto it = pol->cellBest.find(sig[p]);
    if (it == pol->cellBest.end() || fit[p] < it->second - 1e-12) {
      pol->cellBest[sig[p]] = fit[p];
      pol->tabuHits[sig[p]] = 0;
    } else {
      pol->tabuHits[sig[p]] += 1;
    }
  }
  pol->fit.assign(fit, fit + P);
  pol->sig.assign(sig, sig + P);
  pol->observed = true;
  if (best_slot) *best_slot = bi;
  if (improved) *improved = imp ? 1 : 0;
  return SIDON_OK;
}

extern "C" int sidon_policy_plan(sidon_policy* pol, int gen, int* p1, int* p2, int* op,
                                 float* strength, int* mig_src, int* mig_dst, int* n_mig,
                                 int* reseed_slot) {
  if (!pol || !pol->observed || gen < 0 || !p1 || !p2 || !op || !strength) return SIDON_ERR_ARG;
  const int NI = pol->NI, PI = pol->PI;
  std::vector<std::vector<int>> isl(NI, std::vector<int>(PI));
  for (int i = 0; i < NI; ++i) {
    for (int r = 0; r < PI; ++r) isl[i][r] = i * PI + r;
    const std::vector<double>& f = pol->fit;
    std::sort(isl[i].begin(), isl[i].end(),
              [&](int a, int b) { return f[a] < f[b] || (f[a] == f[b] && a < b); });
  }
  int nm = 0;
  if (gen > 0 && gen % 12 == 0) {  // ring migration: best of island i -> worst of i+1
    for (int i = 0; i < NI && NI > 1; ++i) {
      const int src = isl[i][0], dst = isl[(i + 1) % NI][PI - 1];
      if (mig_src) mig_src[nm] = src;
      if (mig_dst) mig_dst[nm] = dst;
      ++nm;
      pol->fit[dst] = pol->fit[src];
      pol->sig[dst] = pol->sig[src];
    }
  }
  if (n_mig) *n_mig = nm;
  const double warm = std::min(1.0, gen / 40.0);
  const double stagn = std::min(1.0, pol->sinceImp / 25.0);
  std::discrete_distribution<int> pick(pol->opW, pol->opW + kOps);
  pol->parentFit = pol->fit;
  for (int i = 0; i < NI; ++i) {
    for (int r = 0; r < PI; ++r) {
      const int slot = i * PI + r;
      if (r < pol->elite) {  // elitism: clone untouched
        p1[slot] = p2[slot] = isl[i][r];
        op[slot] = kClone;
        strength[slot] = 0.0f;
        continue;
      }
      auto tourn = [&]() {
        const int a = isl[i][(int)(pol->U01(pol->rng) * PI)];
        const int b = isl[i][(int)(pol->U01(pol->rng) * PI)];
        return pol->fit[a] < pol->fit[b] ? a : b;
      };
      const int pa = tourn(), pb = tourn();
      int o = pick(pol->rng);
      // a support cell revisited too often without improving, and not
      // competitive with the incumbent, gets a structure-changing operator
      const auto hit = pol->tabuHits.find(pol->sig[pa]);
      const int hits = hit == pol->tabuHits.end() ? 0 : hit->second;
      const auto cb = pol->cellBest.find(pol->sig[pa]);
      const double cbest = cb == pol->cellBest.end() ? 1e30 : cb->second;
      if (hits > kTabuTenure && cbest > pol->bestF * (1.0 + 2e-4))
        o = kDiversifying[(int)(pol->U01(pol->rng) * 7)];
      if (stagn > 0.7 && pol->U01(pol->rng) < 0.10) o = 9;  // stagnation: restarts
      p1[slot] = pa;
      p2[slot] = pb;
      op[slot] = o;
      strength[slot] =
          (float)((0.03 + 0.35 * pol->U01(pol->rng)) * (0.5 + warm) * (1.0 + 1.5 * stagn));
      pol->opUse[o] += 1.0;
    }
  }
  pol->p1.assign(p1, p1 + pol->P);
  pol->op.assign(op, op + pol->P);
  if (reseed_slot) *reseed_slot = (gen % 10 == 9) ? isl[0][PI - 1] : -1;
  return SIDON_OK;
}

extern "C" int sidon_policy_credit(sidon_policy* pol, const double* child_fit) {
  if (!pol || !child_fit) return SIDON_ERR_ARG;
  for (int p = 0; p < pol->P; ++p) {
    const int o = pol->op[p];
    if (o >= 0 && o < kOps && child_fit[p] < pol->parentFit[pol->p1[p]]) pol->opWin[o] += 1.0;
  }
  for (int o = 0; o < kOps; ++o) {
    const double rate = pol->opUse[o] > 20 ? pol->opWin[o] / pol->opUse[o] : 0.25;
    pol->opW[o] = 0.85 * pol->opW[o] + 0.15 * (0.15 + 3.0 * rate);
    pol->opW[o] = std::max(0.05, std::min(4.0, pol->opW[o]));
    pol->opUse[o] *= 0.7;
    pol->opWin[o] *= 0.7;
  }
  return SIDON_OK;
}

extern "C" int sidon_policy_stats(const sidon_policy* pol, double* weights, double* best_f,
                                  double* level_best, int* cells, int* since_improved) {
  if (!pol) return SIDON_ERR_ARG;
  if (weights) std::copy(pol->opW, pol->opW + kOps, weights);
  if (best_f) *best_f = pol->bestF;
  if (level_best) *level_best = pol->lvlBest;
  if (cells) *cells = (int)pol->cellBest.size();
  if (since_improved) *since_improved = pol->sinceImp;
  return SIDON_OK;
}

extern "C" int sidon_policy_tiers(const int* cur_op, int P, int* tier, int* n_cold, int* n_med) {
  if (!cur_op || !tier || P < 1) return SIDON_ERR_ARG;
  const int cap = std::max(1, (int)(0.18 * P));
  int nc = 0, nm = 0;
  for (int p = 0; p < P; ++p) {
    const int o = cur_op[p];
    if ((o == 9 || o == 11) && nc < cap) {
      tier[p] = 0;
      ++nc;
    } else if (o == 9 || o == 11 || o == 10 || o == 5 || o == 7 || o == 8 || o == 12 ||
               o == 13) {
      tier[p] = 1;
      ++nm;
    } else {
      tier[p] = 2;
    }
  }
  if (n_cold) *n_cold = nc;
  if (n_med) *n_med = nm;
  return SIDON_OK;
}
