// C ABI of ragged groups: independent QG problems on grids of their own,
// advanced by shared launches over the members still running (see vardyn.h).
#include "vd_internal.h"

using namespace vd;

int vd_group_create(vd_plan* const* plans, int n, vd_group** out) {
  VD_REQUIRE(plans && out && n > 0, VD_ERR_INVALID_ARG, "null argument or empty group");
  *out = nullptr;
  for (int b = 0; b < n; ++b) {
    VD_REQUIRE(plans[b], VD_ERR_INVALID_ARG, "null plan in group");
    const Plan& p = plans[b]->p;
    const Plan& p0 = plans[0]->p;
    VD_REQUIRE(p.nb == 1, VD_ERR_INVALID_ARG, "group members must be single-problem plans");
    VD_REQUIRE(p.prec == p0.prec && p.scheme == p0.scheme && p.fold == p0.fold &&
                   p.device == p0.device && p.split_ns == p0.split_ns,
               VD_ERR_INVALID_ARG,
               "group members need the same precision, time scheme, sine-transform folding, "
               "split digits and device");
  }
  vd_group* grp = new vd_group();
  Group& g = grp->g;
  g.n = n;
  g.prec = plans[0]->p.prec;
  g.split_ns = plans[0]->p.split_ns;
  g.scheme = plans[0]->p.scheme;
  g.nstage = plans[0]->p.nstage;
  g.fold = plans[0]->p.fold;
  for (int b = 0; b < n; ++b) g.plans.push_back(&plans[b]->p);
  const int rc = g.prec == VD_F64 ? group_build<double>(g) : group_build<float>(g);
  if (rc != VD_OK) {
    group_release(g);
    delete grp;
    return rc;
  }
  *out = grp;
  return VD_OK;
}

int vd_group_destroy(vd_group* grp) {
  if (!grp) return VD_OK;
  group_release(grp->g);
  delete grp;
  return VD_OK;
}

int vd_group_offsets(const vd_group* grp, size_t* offsets) {
  VD_REQUIRE(grp && offsets, VD_ERR_INVALID_ARG, "null argument");
  for (int b = 0; b <= grp->g.n; ++b) offsets[b] = grp->g.foff[b];
  return VD_OK;
}

#define VD_GROUP_CHECK(grp, nact)                                                        \
  VD_REQUIRE(grp, VD_ERR_INVALID_ARG, "null group");                                     \
  VD_REQUIRE((nact) >= 0 && (nact) <= (grp)->g.n, VD_ERR_INVALID_ARG,                    \
             "nact must be in 0..number of members")

int vd_group_stage(vd_group* grp, int nact, const void* h, const void* q, const void* qbase,
                   const void* k0, const void* k1, double c0, double c1, double c2, void* qout,
                   void* kstore, void* stream) {
  VD_GROUP_CHECK(grp, nact);
  VD_REQUIRE(h && q && qbase && qout, VD_ERR_INVALID_ARG, "null argument");
  VD_REQUIRE(qout != h && qout != q, VD_ERR_INVALID_ARG,
             "qout must not alias h or q (the stencil reads them at neighbouring points)");
  const Group& g = grp->g;
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (g.prec == VD_F64)
    launch_stage_rg<double>(g, nact, (const double*)h, (const double*)q, (const double*)qbase,
                            (const double*)k0, (const double*)k1, c0, c1, c2, (double*)qout,
                            (double*)kstore, s);
  else
    launch_stage_rg<float>(g, nact, (const float*)h, (const float*)q, (const float*)qbase,
                           (const float*)k0, (const float*)k1, (float)c0, (float)c1, (float)c2,
                           (float*)qout, (float*)kstore, s);
  VD_HIP_TRY(hipGetLastError());
  return VD_OK;
}

int vd_group_inversion(vd_group* grp, int nact, const void* F, const void* Fsub, const void* hb,
                       void* out, int accumulate, void* stream) {
  VD_GROUP_CHECK(grp, nact);
  VD_REQUIRE(F && out, VD_ERR_INVALID_ARG, "null argument");
  const Group& g = grp->g;
  hipStream_t s = static_cast<hipStream_t>(stream);
  const int mode = accumulate ? INV_ACC : INV_FWD;
  if (g.prec == VD_F64)
    launch_inversion_rg<double>(g, nact, (const double*)F, (const double*)Fsub,
                                (const double*)hb, (double*)out, mode, s);
  else
    launch_inversion_rg<float>(g, nact, (const float*)F, (const float*)Fsub, (const float*)hb,
                               (float*)out, mode, s);
  VD_HIP_TRY(hipGetLastError());
  return VD_OK;
}

int vd_group_stage_adj(vd_group* grp, int nact, const void* h, const void* q, const void* s0,
                       const void* s1, const void* s2, double a0, double a1, double a2,
                       const void* acc0, const void* acc1, const void* acc2, void* lq_out,
                       void* lh_out, void* stream) {
  VD_GROUP_CHECK(grp, nact);
  VD_REQUIRE(h && q && s0 && lq_out && lh_out, VD_ERR_INVALID_ARG, "null argument");
  VD_REQUIRE(lq_out != s0 && lq_out != s1 && lq_out != s2, VD_ERR_INVALID_ARG,
             "lq_out must not alias s0..s2 (they are read at neighbouring points)");
  const Group& g = grp->g;
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (g.prec == VD_F64)
    launch_stage_adj_rg<double>(g, nact, (const double*)h, (const double*)q, (const double*)s0,
                                (const double*)s1, (const double*)s2, a0, a1, a2,
                                (const double*)acc0, (const double*)acc1, (const double*)acc2,
                                (double*)lq_out, (double*)lh_out, s);
  else
    launch_stage_adj_rg<float>(g, nact, (const float*)h, (const float*)q, (const float*)s0,
                               (const float*)s1, (const float*)s2, (float)a0, (float)a1,
                               (float)a2, (const float*)acc0, (const float*)acc1,
                               (const float*)acc2, (float*)lq_out, (float*)lh_out, s);
  VD_HIP_TRY(hipGetLastError());
  return VD_OK;
}

// Segment start and end of the members, as vd_qg_prep / vd_qg_finish /
// vd_qg_adj_init / vd_qg_adj_final compute them for each member's own plan.
int vd_group_prep(vd_group* grp, int nact, const void* h0, const void* hb, void* h, void* q,
                  void* hbm, void* qb, void* stream) {
  VD_GROUP_CHECK(grp, nact);
  VD_REQUIRE(h0 && hb && h && q, VD_ERR_INVALID_ARG, "null argument");
  for (const void* out : {(const void*)h, (const void*)q, (const void*)hbm, (const void*)qb})
    VD_REQUIRE(out != h0 && out != hb, VD_ERR_INVALID_ARG,
               "outputs must not alias h0 or hb (h2pv reads them at neighbouring points)");
  const Group& g = grp->g;
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (g.prec == VD_F64)
    launch_prep_rg<double>(g, nact, (const double*)h0, (const double*)hb, (double*)h, (double*)hbm,
                           (double*)q, (double*)qb, s);
  else
    launch_prep_rg<float>(g, nact, (const float*)h0, (const float*)hb, (float*)h, (float*)hbm,
                          (float*)q, (float*)qb, s);
  VD_HIP_TRY(hipGetLastError());
  return VD_OK;
}

int vd_group_finish(vd_group* grp, int nact, const void* h, void* h1, void* stream) {
  VD_GROUP_CHECK(grp, nact);
  VD_REQUIRE(h && h1, VD_ERR_INVALID_ARG, "null argument");
  const Group& g = grp->g;
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (g.prec == VD_F64)
    launch_finish_rg<double>(g, nact, (const double*)h, (double*)h1, s);
  else
    launch_finish_rg<float>(g, nact, (const float*)h, (float*)h1, s);
  VD_HIP_TRY(hipGetLastError());
  return VD_OK;
}

int vd_group_adj_init(vd_group* grp, int nact, const void* adh1, void* lh, void* lq, void* stream) {
  VD_GROUP_CHECK(grp, nact);
  VD_REQUIRE(adh1 && lh && lq, VD_ERR_INVALID_ARG, "null argument");
  const Group& g = grp->g;
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (g.prec == VD_F64)
    launch_adj_init_rg<double>(g, nact, (const double*)adh1, (double*)lh, (double*)lq, s);
  else
    launch_adj_init_rg<float>(g, nact, (const float*)adh1, (float*)lh, (float*)lq, s);
  VD_HIP_TRY(hipGetLastError());
  return VD_OK;
}

int vd_group_adj_final(vd_group* grp, int nact, const void* hcur, const void* lh, const void* lq,
                       void* adh0, void* stream) {
  VD_GROUP_CHECK(grp, nact);
  VD_REQUIRE(hcur && lh && lq && adh0, VD_ERR_INVALID_ARG, "null argument");
  VD_REQUIRE(adh0 != lq, VD_ERR_INVALID_ARG,
             "adh0 must not alias lq (h2pv^T reads it at neighbouring points)");
  const Group& g = grp->g;
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (g.prec == VD_F64)
    launch_adj_final_rg<double>(g, nact, (const double*)hcur, (const double*)lh,
                                (const double*)lq, (double*)adh0, s);
  else
    launch_adj_final_rg<float>(g, nact, (const float*)hcur, (const float*)lh, (const float*)lq,
                               (float*)adh0, s);
  VD_HIP_TRY(hipGetLastError());
  return VD_OK;
}

// Tile shapes and the fused small-member inversion of a group, picked by
// timing the alternatives on the group's own members (every member active).
// Decisions, each kept only where it applies:
//   split-product groups: per inversion stage, 64 x 32 or 32 x 32 split tiles
//     (chosen[0..3]: 0 or 1; -1 for other groups) -- the same bits either way;
//   the adjoint stage: 64 x 16 or 32 x 32 point tiles (chosen[4]: 0 or 1) --
//     the same bits either way;
//   equal-size small FP64 members: the inversion as one fused launch
//     (k_inv_small) or as the four GEMM stage launches (chosen[5]: 1 or 0; -1
//     where the fused form does not apply) -- results within rounding of
//     each other (the FP64 sums run in another order).
// Each candidate runs `reps` times per round, three rounds, fastest round
// kept.  The timing inputs are scratch fields of the group's layout filled
// with a finite constant; the call allocates and frees them, and synchronises
// the stream.
template <typename T>
static int group_autotune(Group& g, int reps, int* chosen, hipStream_t s) {
  const size_t nel = g.foff[g.n];
  T* buf = nullptr;
  VD_HIP_TRY(hipMalloc((void**)&buf, 6 * nel * sizeof(T)));
  int rc = VD_OK;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  auto done = [&](int r) {
    if (e0) hipEventDestroy(e0);
    if (e1) hipEventDestroy(e1);
    hipFree(buf);
    return r;
  };
  if (hipMemsetAsync(buf, 0x3F, 6 * nel * sizeof(T), s) != hipSuccess ||
      hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
    set_error("autotune setup failed");
    return done(VD_ERR_HIP);
  }
  T *f0 = buf, *f1 = buf + nel, *f2 = buf + 2 * nel, *f3 = buf + 3 * nel, *f4 = buf + 4 * nel;
  // fastest of three rounds of `reps` calls of each candidate; returns the index
  auto pick = [&](int ncand, auto&& run) -> int {
    float best = 1e30f;
    int bi = 0;
    for (int round = 0; round < 3; ++round)
      for (int c = 0; c < ncand; ++c) {
        run(c);  // warm
        hipEventRecord(e0, s);
        for (int r = 0; r < reps; ++r) run(c);
        hipEventRecord(e1, s);
        if (hipEventSynchronize(e1) != hipSuccess) rc = VD_ERR_HIP;
        float ms = 0;
        hipEventElapsedTime(&ms, e0, e1);
        if (ms < best) {
          best = ms;
          bi = c;
        }
      }
    return bi;
  };
  for (int st = 0; st < 4; ++st) chosen[st] = -1;
  if constexpr (sizeof(T) == 8) {
    if (g.split_ns > 0)
      for (int st = 0; st < 4; ++st) {
        g.split_shape[st] = pick(2, [&](int shape) {
          launch_split_stage_rg(g, st, shape, g.n, f0, nullptr, f1, f3, INV_FWD, s);
        });
        chosen[st] = g.split_shape[st];
      }
  }
  g.adj_narrow = pick(2, [&](int narrow) {
                   g.adj_narrow = narrow != 0;
                   launch_stage_adj_rg<T>(g, g.n, f0, f1, f2, nullptr, nullptr, T(1), T(0), T(0), nullptr,
                                          nullptr, nullptr, f3, f4, s);
                 }) != 0;
  chosen[4] = g.adj_narrow ? 1 : 0;
  chosen[5] = -1;
  if (g.small) {
    g.use_small = pick(2, [&](int fused) {
                    g.use_small = fused == 0;
                    launch_inversion_rg<T>(g, g.n, f0, nullptr, f1, f3, INV_FWD, s);
                  }) == 0;
    chosen[5] = g.use_small ? 1 : 0;
  }
  if (rc == VD_OK && hipGetLastError() != hipSuccess) rc = VD_ERR_HIP;
  if (rc != VD_OK) set_error("autotune: a timed launch failed");
  return done(rc);
}

int vd_group_autotune(vd_group* grp, int reps, int* chosen, void* stream) {
  VD_REQUIRE(grp && chosen, VD_ERR_INVALID_ARG, "null argument");
  VD_REQUIRE(reps > 0 && reps <= 1000, VD_ERR_INVALID_ARG, "reps must be in 1..1000");
  Group& g = grp->g;
  hipStream_t s = static_cast<hipStream_t>(stream);
  return g.prec == VD_F64 ? group_autotune<double>(g, reps, chosen, s)
                          : group_autotune<float>(g, reps, chosen, s);
}

int vd_group_set_tuning(vd_group* grp, const int* chosen) {
  VD_REQUIRE(grp && chosen, VD_ERR_INVALID_ARG, "null argument");
  Group& g = grp->g;
  for (int st = 0; st < 4; ++st)
    VD_REQUIRE(g.split_ns > 0 ? (chosen[st] == 0 || chosen[st] == 1) : chosen[st] == -1, VD_ERR_INVALID_ARG,
               "chosen[0..3] must be 0 or 1 in a split-product group and -1 otherwise");
  VD_REQUIRE(chosen[4] == 0 || chosen[4] == 1, VD_ERR_INVALID_ARG, "chosen[4] must be 0 or 1");
  VD_REQUIRE(g.small ? (chosen[5] == 0 || chosen[5] == 1) : chosen[5] == -1, VD_ERR_INVALID_ARG,
             "chosen[5] must be 0 or 1 where the fused small-member inversion applies and -1 otherwise");
  for (int st = 0; st < 4; ++st)
    if (g.split_ns > 0) g.split_shape[st] = chosen[st];
  g.adj_narrow = chosen[4] == 1;
  if (g.small) g.use_small = chosen[5] == 1;
  return VD_OK;
}
