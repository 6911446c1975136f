// Autotuner of the exchange tiling: for the orbital sets of one call it times
// every combination of SYRK tile (64, 128), SYRK workgroups per launch (1, 2, 4
// or 8 rounds of two per CU) and half-transform AO tile (64, 128) with HIP
// events and keeps the fastest in the context.  Every point runs once untimed
// (so a switch of tiling pays its workspace growth and cold caches outside the
// clock) and is then timed kRuns times back to back; the median decides, so
// one slow run (a clock ramp, another queue's work) cannot pick the tiling.
// Shapes differ enough between molecules (n 100-600, o 10-120) that no single
// tiling wins everywhere.
#include <algorithm>

#include "mjk_internal.h"

using namespace mjk;

namespace {
constexpr int kRuns = 3;
}

extern "C" int mjk_autotune(mjk_context *c, int32_t nk, const int32_t *nocc,
                            const double *const *corb, double *klt, int64_t ldk,
                            int32_t precision, double *best_ms) {
  if (!c) return MJK_ERR_ARG;
  if (!nocc || !corb || !klt) return set_error(c, MJK_ERR_ARG, "autotune: null pointer");
  const bool was_profiling = c->profile;
  c->profile = false;
  hipEvent_t e0[kRuns] = {}, e1[kRuns] = {};
  int rc = MJK_OK;
  for (int r = 0; r < kRuns && rc == MJK_OK; ++r)
    if ((rc = hip_check(c, hipEventCreate(&e0[r]), "event")) == MJK_OK)
      rc = hip_check(c, hipEventCreate(&e1[r]), "event");
  const int tms[] = {64, 128};
  const int splits[] = {2 * c->ncu, 4 * c->ncu, 8 * c->ncu, 16 * c->ncu};
  const int tas[] = {64, 128};
  float best = 1e30f;
  int bt = 0, bs = 0, ba = 0;
  if (rc) goto done;
  // warm-up with the automatic choice (allocates the workspaces once)
  c->tile_syrk = c->split_target = c->ht_ta = 0;
  if ((rc = exchange(c, nk, nocc, corb, klt, ldk, precision))) goto done;
  for (int tm : tms) {
    if (c->n_pad % tm) continue;
    for (int sp : splits) {
      for (int ta : tas) {
        c->tile_syrk = tm;
        c->split_target = sp;
        c->ht_ta = ta;
        if ((rc = exchange(c, nk, nocc, corb, klt, ldk, precision))) goto done;  // untimed
        for (int r = 0; r < kRuns; ++r) {
          if ((rc = hip_check(c, hipEventRecord(e0[r], c->stream), "record"))) goto done;
          if ((rc = exchange(c, nk, nocc, corb, klt, ldk, precision))) goto done;
          if ((rc = hip_check(c, hipEventRecord(e1[r], c->stream), "record"))) goto done;
        }
        if ((rc = hip_check(c, hipEventSynchronize(e1[kRuns - 1]), "sync"))) goto done;
        float t[kRuns];
        for (int r = 0; r < kRuns; ++r)
          if ((rc = hip_check(c, hipEventElapsedTime(&t[r], e0[r], e1[r]), "elapsed"))) goto done;
        std::sort(t, t + kRuns);
        const float ms = t[kRuns / 2];
        if (ms < best) {
          best = ms;
          bt = tm;
          bs = sp;
          ba = ta;
        }
      }
    }
  }
  c->tile_syrk = bt;
  c->split_target = bs;
  c->ht_ta = ba;
  if (best_ms) *best_ms = best;
done:
  for (int r = 0; r < kRuns; ++r) {
    if (e0[r]) (void)hipEventDestroy(e0[r]);
    if (e1[r]) (void)hipEventDestroy(e1[r]);
  }
  c->profile = was_profiling;
  return rc;
}
