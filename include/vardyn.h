/*
 * vardyn.h -- C ABI of libvardyn_hip.so, the MI355X (gfx950) implementation of
 * the QG 4D-Var hot path of datlas-ocean/VarDyn: the 1.5-layer QG model step,
 * its tangent-linear and adjoint, the L3 observation operator and the
 * reductions of the 4D-Var cost.
 *
 * Conventions
 *   - Every entry point returns an int status: VD_OK (0) or a negative
 *     VD_ERR_* code; vd_last_error() gives a thread-local message.
 *   - Fields are row-major (ny, nx) arrays in device memory, axis 0 = y,
 *     contiguous in x, of the plan precision (float for VD_F32, double for
 *     VD_F64).  A plan made for nbatch > 1 problems takes (nbatch, ny, nx)
 *     arrays instead and advances every member in the same launches.
 *     Observation values, weights, misfits and cost accumulators are
 *     double.
 *   - The caller owns every buffer passed in.  A plan owns its constants
 *     (sine matrices, 1/Lambda, masks) and a scratch set, allocated by
 *     vd_qg_plan_create; calls that share a plan must not overlap (the
 *     scratch is shared), calls on distinct plans may run concurrently.
 *   - `stream` is a hipStream_t (NULL = default stream).  Work is enqueued
 *     asynchronously; nothing in the step/obs calls allocates or
 *     synchronises, so a sequence of calls can be captured in a HIP graph.
 *
 * The reference is Python/JAX with no FFI; each entry point names the
 * reference function whose contract it takes over (file:line in
 * datlas-ocean/VarDyn).
 */
#ifndef VARDYN_H
#define VARDYN_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define VD_VERSION_MAJOR 0
#define VD_VERSION_MINOR 1

enum {
  VD_OK = 0,
  VD_ERR_INVALID_ARG = -1,
  VD_ERR_SHAPE = -2,
  VD_ERR_HIP = -3,
  VD_ERR_UNSUPPORTED = -4,
  VD_ERR_ALLOC = -5
};

enum { VD_F32 = 0, VD_F64 = 1 };
enum { VD_EULER = 0, VD_RK2 = 1, VD_RK3 = 2 };

/* Model parameters.  Mirrors the arguments of Qgm.__init__
 * (mapping/models/model_qg1l/jqgm.py:113) after its reductions:
 * dx, dy are the centre-pixel spacings (jqgm.py:126-128), f0 and c the
 * nanmeans of f and c (jqgm.py:140-151).  Host pointers; copied at create. */
typedef struct vd_qg_params {
  int ny, nx;            /* grid shape, >= 8 each */
  int precision;         /* VD_F32 or VD_F64 */
  int time_scheme;       /* VD_EULER, VD_RK2 or VD_RK3 (jqgm.py:398-477) */
  double dx, dy;         /* metres */
  double dt;             /* seconds */
  double g, f0, c;       /* gravity, Coriolis, gravity-wave speed */
  int has_kdiffus;       /* PV diffusion (jqgm.py:362-369) */
  double kdiffus;
  const int8_t* mask;    /* (ny, nx) mask or NULL for the land-free ring mask:
                            0 land, 1/2 boundary (q = -g f/c^2 hb, no
                            tendency), 3 active (rows/columns 2..n-3), as
                            jqgm.py:179-217 builds them; 4 interior point
                            without tendency (q from the Laplacian, rows and
                            columns 1..n-2) and 5 point with q = 0 and no
                            tendency: the rows next to and on the ring of a
                            QgmWithTiles tile inside the domain
                            (jqgm.py:1129-1230) */
  const double* beta;    /* (ny-2, nx) beta = (f[2:]-f[:-2])/(2 dy)
                            (jqgm.py:144) or NULL (f plane) */
  const double* fbathy;  /* (ny, nx) f * bathymetry_PV_term (jqgm.py:353-354)
                            or NULL */
  const double* mdt;     /* (ny, nx) mean dynamic topography (jqgm.py:532-555)
                            or NULL */
  int dst_dense;         /* 0: fold the sine transform by its symmetry when
                            ny-2 and nx-2 are even (half the GEMM work);
                            1: always the full (n-2)^2 sine matrices */
  int gemm_config;       /* GEMM tile configuration: 0 = chosen from the
                            problem shape; k in 1..11 forces configuration
                            k-1: single-LDS-buffer 32x32 (split-K), 64x64,
                            64x64 (split-K 2), 128x64, 128x64 (split-K 2);
                            double-buffered 32x32 (split-K 4), 64x64,
                            64x64 (split-K 2), 128x64, 64x32 (split-K 2);
                            single-LDS-buffer 128x128 */
  int nbatch;            /* independent problems on this grid (0 or 1 = one);
                            with nbatch > 1 every field argument of the step
                            and inversion calls is an (nbatch, ny, nx) array */
  const double* c_batch; /* (nbatch) gravity-wave speed per member, or NULL
                            for c everywhere */
  int dst_split;         /* 0: sine-matrix GEMMs in the plan precision on
                            FP64/FP32 MFMA; 5, 6 or 7 (FP64 plans): split
                            products on int8 MFMA, each FP64 operand written
                            as that many base-256 digits per line (Ozaki
                            scheme), relative GEMM error about 2^(5-8n);
                            interior sizes up to 1024 (folded) or 512; with
                            it, gemm_config k in 1..4 forces split tile
                            configuration k-1 (sine x data lines per
                            workgroup: 128x32, 64x64, 64x32, 32x32) */
} vd_qg_params;

typedef struct vd_plan vd_plan;

/* Library version. */
int vd_version(int* major, int* minor);

/* Message of the last failed call on this thread ("" if none). */
const char* vd_last_error(void);

/* Build a plan: Helmholtz eigenvalues, sine matrices, masks, scratch.
 * Replaces Qgm.__init__ (jqgm.py:113-236). */
int vd_qg_plan_create(const vd_qg_params* params, vd_plan** plan);
int vd_qg_plan_destroy(vd_plan* plan);

/* Bytes of trajectory tape for a segment of `nstep` steps (stage states of
 * every step plus the final state, for every batch member). */
int vd_qg_tape_bytes(const vd_plan* plan, int nstep, size_t* bytes);

/* Time every tile configuration of each of the four inversion GEMMs on the
 * device (`reps` launches each, on `stream`) and keep the fastest in the
 * plan; `chosen` (may be NULL) receives the four configuration indices.
 * Synchronises the stream.  No reference counterpart: the reference leaves
 * kernel choices to XLA. */
int vd_qg_autotune(vd_plan* plan, int reps, int* chosen, void* stream);

/* Bytes of plan-owned device memory (constants + scratch). */
int vd_qg_plan_bytes(const vd_plan* plan, size_t* bytes);

/* Diagnostics: device address and size of the plan's GEMM intermediates
 * (which = 0: the first/third GEMM output, 1: the scaled second GEMM
 * output), each (MY x MX) row-major with MY, MX the interior sizes rounded
 * up to 32.  Valid until the next call on the plan.  No reference
 * counterpart (the reference's inversion has no exposed intermediates). */
int vd_qg_debug_scratch(const vd_plan* plan, int which, void** ptr, int* rows, int* cols);

/* Forward propagation over `nstep` steps: h1 = M(h0; hb).  Replaces
 * Qgm.step (jqgm.py:517-557): q0 = h2pv(h0), q is prognostic, h re-diagnosed
 * by the Helmholtz inversion, NaN written on land in h1.  `tape` may be
 * NULL; otherwise it receives the stage states (vd_qg_tape_bytes).  `hb`
 * NULL means a zero boundary field (step_tgl and step_adj take the same). */
int vd_qg_step(vd_plan* plan, const void* h0, const void* hb, void* h1, int nstep,
               void* tape, void* stream);

/* Tangent-linear propagation dh1 = M'(h0; hb) dh0.  Replaces
 * Qgm.step_tgl (jqgm.py:559-563, jax.jvp of step).  The trajectory is
 * recomputed alongside; `h1` (may be NULL) receives the forward result. */
int vd_qg_step_tgl(vd_plan* plan, const void* dh0, const void* h0, const void* hb, void* dh1,
                   void* h1, int nstep, void* stream);

/* Adjoint propagation adh0 = M'(h0; hb)^T adh1.  Replaces Qgm.step_adj
 * (jqgm.py:565-569, jax.vjp of step).  If `tape_valid` is 0 the segment is
 * first recomputed into `tape` (required, vd_qg_tape_bytes); if 1, `tape`
 * must hold the tape written by vd_qg_step for the same (h0, hb, nstep). */
int vd_qg_step_adj(vd_plan* plan, const void* adh1, const void* h0, const void* hb, void* adh0,
                   int nstep, void* tape, int tape_valid, void* stream);

/* Elliptic inversion h = hb + S(S (q - qb)_int S ./ Lambda)S on the
 * interior, hb on the ring.  Replaces Qgm.pv2h (jqgm.py:295-319) and
 * inverse_elliptic_dst (jqgm.py:83-91).  qb may be NULL (zero). */
int vd_qg_pv2h(vd_plan* plan, const void* q, const void* hb, const void* qb, void* h,
               void* stream);

/* q = h2pv(h, hb) (jqgm.py:264-293). */
int vd_qg_h2pv(vd_plan* plan, const void* h, const void* hb, void* q, void* stream);

/* out = base + coef * rhs(h, q): the PV tendency of Qgm.rhs (jqgm.py:321-375,
 * with the upwind advection of Qgm.adv :377-396 and the beta, diffusion and
 * bathymetry terms the plan was created with) fused with one explicit stage
 * update -- Qgm.euler (jqgm.py:398-409) is base = q, coef = dt.  base NULL
 * means zero (out = coef * rhs).  rhs is zero outside the active points
 * (mask 3), as in the reference.  out may alias base, not h or q. */
int vd_qg_rhs(vd_plan* plan, const void* h, const void* q, const void* base, double coef,
              void* out, void* stream);

/* Transpose of vd_qg_rhs at (h, q): the vjp of coef * rhs(h2uv(h), q)
 * (Qgm.h2uv + Qgm.rhs, jqgm.py:238-262, 321-375; the reference gets it from
 * jax.vjp inside step_adj, jqgm.py:565-569) applied to the cotangent lam:
 *   lq = acc + coef * (d rhs / d q)^T lam,   lh = coef * (d rhs / d h)^T lam,
 * where h enters through the velocities.  acc NULL means zero and may alias
 * lq; lam is read at neighbouring points, so neither output may alias it or
 * h or q.  This is the adjoint RK stage kernel of vd_qg_step_adj. */
int vd_qg_rhs_adj(vd_plan* plan, const void* h, const void* q, const void* lam, double coef,
                  const void* acc, void* lq, void* lh, void* stream);

/* field[i] = value where the plan mask is 0 (land); the reference fills the
 * initial state with NaN there (mapping/src/mod.py:697-700). */
int vd_qg_fill_land(vd_plan* plan, void* field, double value, void* stream);

/* ---------------------------------------------------------------------------
 * The pieces of one Qgm.step call (jqgm.py:517-557), for drivers that advance
 * several problems step by step (ragged groups below).  vd_qg_step is
 * prep -> nstep x (stages + inversions) -> finish; vd_qg_step_adj is
 * adj_init -> reverse steps -> adj_final.
 * ------------------------------------------------------------------------- */
/* Segment start (jqgm.py:528-536): h = h0 (+ mdt), q = h2pv(h0, hb); hb NULL
 * means zero. */
int vd_qg_prep(vd_plan* plan, const void* h0, const void* hb, void* h, void* q, void* stream);
/* Segment end (jqgm.py:549-555): h1 = h (- mdt), NaN on land. */
int vd_qg_finish(vd_plan* plan, const void* h, void* h1, void* stream);
/* Start of the reverse sweep (adjoint of vd_qg_finish): lh = adh1 with zeros
 * on land, lq = 0. */
int vd_qg_adj_init(vd_plan* plan, const void* adh1, void* lh, void* lq, void* stream);
/* End of the reverse sweep (adjoint of vd_qg_prep at the segment start state
 * hcur): adh0 = lh + h2pv^T lq. */
int vd_qg_adj_final(vd_plan* plan, const void* hcur, const void* lh, const void* lq, void* adh0,
                    void* stream);

/* ---------------------------------------------------------------------------
 * Ragged groups: independent problems on grids of their own (own size, c,
 * mask, window), advanced by shared launches -- one launch per stage or
 * inversion covers every member that is still running.  No reference
 * counterpart: the reference runs such problem sets one process per
 * subdomain (mapping/run_assimilation.py:27-350).
 *
 * A flat field holds member b at elements offsets[b] .. offsets[b+1]-1 as an
 * (ny_b, nx_b) array.  The calls act on members 0 .. nact-1, so a driver
 * orders the members by window length (longest first) and passes the number
 * still running.  Per member the arithmetic is exactly that of its own plan
 * with GEMM configuration 0 (32 x 32 tiles), except the fused inversion of
 * equal-size small FP64 members (the same sums in another order).
 * ------------------------------------------------------------------------- */
typedef struct vd_group vd_group;

/* Group n plans (nbatch 1 each; same precision, time scheme and folding of
 * the sine transforms).  The plans must outlive the group. */
int vd_group_create(vd_plan* const* plans, int n, vd_group** group);
int vd_group_destroy(vd_group* group);
/* offsets[0..n]: element offset of each member's fields; offsets[n] = total. */
int vd_group_offsets(const vd_group* group, size_t* offsets);
/* One explicit RK stage (jqgm.py:398-477) for members 0..nact-1:
 * qout = qbase + c0 k0 + c1 k1 + c2 rhs(h, q); kstore (may be NULL)
 * receives rhs.  k0, k1 may be NULL. */
int vd_group_stage(vd_group* group, int nact, const void* h, const void* q, const void* qbase,
                   const void* k0, const void* k1, double c0, double c1, double c2, void* qout,
                   void* kstore, void* stream);
/* Helmholtz inversion (jqgm.py:295-319) for members 0..nact-1:
 * accumulate = 0: out = hb + P(F - Fsub) on the interior, hb on the ring
 * (Fsub, hb may be NULL = zero); accumulate = 1: out += P(F). */
int vd_group_inversion(vd_group* group, int nact, const void* F, const void* Fsub, const void* hb,
                       void* out, int accumulate, void* stream);
/* Adjoint of one RK stage for members 0..nact-1: with the cotangent
 * a0 s0 + a1 s1 + a2 s2 of rhs(h, q), lq_out = acc0 + acc1 + acc2 + d/dq and
 * lh_out = d/dh.  s0..s2 must not alias lq_out; acc* may. */
int vd_group_stage_adj(vd_group* group, int nact, const void* h, const void* q, const void* s0,
                       const void* s1, const void* s2, double a0, double a1, double a2,
                       const void* acc0, const void* acc1, const void* acc2, void* lq_out,
                       void* lh_out, void* stream);
/* Segment start and end of members 0..nact-1, each as vd_qg_prep /
 * vd_qg_finish / vd_qg_adj_init / vd_qg_adj_final compute it with the member's
 * own plan (jqgm.py:528-536, 549-555 and their adjoints), in one launch.
 * vd_group_prep: h = h0 (+ mdt), q = h2pv(h0, hb) and, where hbm / qb are not
 * NULL, hbm = hb (+ mdt), qb = h2pv(hb, hb); hb is required (a zero field for
 * none).  vd_group_adj_final: adh0 = lh + h2pv^T lq at the start state hcur. */
int vd_group_prep(vd_group* group, int nact, const void* h0, const void* hb, void* h, void* q,
                  void* hbm, void* qb, void* stream);
int vd_group_finish(vd_group* group, int nact, const void* h, void* h1, void* stream);
int vd_group_adj_init(vd_group* group, int nact, const void* adh1, void* lh, void* lq,
                      void* stream);
int vd_group_adj_final(vd_group* group, int nact, const void* hcur, const void* lh,
                       const void* lq, void* adh0, void* stream);
/* Tile shapes and inversion form of the group, picked by timing the
 * alternatives on its own members (no reference counterpart: the problem-set
 * engine's tuner).  chosen[0..5] receives the choices: [0..3] the split-product
 * tile shape of each inversion stage (0: 64 sine x 32 data lines, 1: 32 x 32;
 * -1 in groups without split products), [4] the adjoint-stage tile (0: 64 x 16
 * points, 1: 32 x 32), [5] the inversion of equal-size small FP64 members as
 * one fused launch (1) or as the four GEMM stages (0; -1 where the fused form
 * does not apply).  Choices [0..4] leave every result bit for bit unchanged;
 * [5] changes results within rounding.  Each candidate runs `reps` (1..1000)
 * times in each of three rounds; the call allocates and frees scratch fields
 * for the timing and synchronises the stream. */
int vd_group_autotune(vd_group* group, int reps, int* chosen, void* stream);
/* Apply choices in the form vd_group_autotune returns (for instance stored
 * from an earlier run); a value that does not apply to the group is an error. */
int vd_group_set_tuning(vd_group* group, const int* chosen);

/* ---------------------------------------------------------------------------
 * Balanced-motion wavelet basis (Basis_bm, mapping/src/basis.py:472-936):
 * SSH fields from wavelet coefficients (operg, basis.py:863-901) and the
 * transpose (operg_transpose, basis.py:903-936) for a batch of T basis times.
 * The host builds the tables (vardyn_amd/basis_bm.py): a rectilinear grid
 * (lon1 along x, lat1 along y, mask 1 = masked), per frequency DX (km), the
 * wave vectors kx, ky (nf x ntheta), the first coefficient, centre count and
 * time-centre count; per centre its frequency, index, lon, lat, cos(lat) and
 * point box (i0, i1, j0, j1 inclusive); per grid point the centres whose box
 * covers it (CSR).  Coefficients are laid out as in the reference:
 * frequency, time centre, space centre, direction, cos/sin.  The time factors
 * gt of a call are, per frequency, a T x nt_f row-major block, concatenated.
 * ------------------------------------------------------------------------- */
typedef struct vd_bm_basis vd_bm_basis;
typedef struct vd_bm_params {
  int ny, nx, nf, ntheta, nc;
  const double* lon1;
  const double* lat1;
  const uint8_t* mask;
  const double* dxf;
  const double* kx;
  const double* ky;
  const int32_t* cf;
  const int32_t* cP;
  const double* clon;
  const double* clat;
  const double* ccos;
  const int32_t* cbox;
  const int64_t* fbase;
  const int32_t* fnp;
  const int32_t* fnt;
  const int64_t* pptr;
  const int32_t* plist;
} vd_bm_params;

/* Upload the tables (host arrays) to the current device. */
int vd_bm_create(const vd_bm_params* params, vd_bm_basis** basis);
int vd_bm_destroy(vd_bm_basis* basis);
/* Number of coefficients and of time-factor columns (sum of nt_f). */
int vd_bm_info(const vd_bm_basis* basis, long long* nbasis, int* gt_total);
/* Device workspace of one call over T times. */
int vd_bm_workspace_size(const vd_bm_basis* basis, int T, size_t* bytes);
/* ssh[t] (T fields of ny x nx) = operg(time t, X); device pointers. */
int vd_bm_operg(vd_bm_basis* basis, int T, const double* gt, const double* X, double* ssh,
                void* work, void* stream);
/* adX += sum over the T times of operg_transpose(time t, adssh[t]). */
int vd_bm_operg_adj(vd_bm_basis* basis, int T, const double* gt, const double* adssh,
                    double* adX, void* work, void* stream);
/* Geostrophic velocities of the basis fields (compute_velocities,
 * basis.py:822-860; _ssh2uv / _ssh2uv_adj).  vd_bm_uv_tables uploads, once,
 * au = -g / f_on_u and dy_on_u on the u grid (ny x (nx+1)) and bv = g / f_on_v
 * and dx_on_v on the v grid ((ny+1) x nx), the reference's edge-padded
 * midpoint tables (basis.py:514-526), as host arrays.  vd_bm_ssh2uv: for T
 * fields, u = (au dP_y) / dy_on_u and v = (bv dP_x) / dx_on_v with P the SSH
 * padded by one edge row and column; vd_bm_ssh2uv_adj: adssh += the
 * transpose applied to (adu, adv), in the reference's order of operations.
 * Device pointers for the fields. */
int vd_bm_uv_tables(vd_bm_basis* basis, const double* au, const double* dy_on_u, const double* bv,
                    const double* dx_on_v);
int vd_bm_ssh2uv(vd_bm_basis* basis, int T, const double* ssh, double* u, double* v, void* stream);
int vd_bm_ssh2uv_adj(vd_bm_basis* basis, int T, const double* adu, const double* adv, double* adssh,
                     void* stream);

/* ---------------------------------------------------------------------------
 * 3-D wavelet basis (Basis_wavelet3d, mapping/src/basis.py:2640-2985): fields
 * from wavelet coefficients (operg, basis.py:2936-2960) and the transpose
 * (operg_transpose, basis.py:2962-2985) for a batch of T basis times, as
 * dense FP64 MFMA products over 16 x 16-point tiles.  The host builds the
 * tables (vardyn_amd/basis_wavelet3d.py): the grid's lon and lat (ny x nx,
 * any 2-D grid), per space frequency DX (km), the wave vectors kx, ky
 * (nf x ntheta), its first coefficient, first family and family count; per
 * family (a space centre with its 2 ntheta waves; families frequency-major)
 * its frequency, lon, lat and cos(lat); per tile (row-major over the
 * ceil(ny/16) x ceil(nx/16) tiles) a CSR of slots (tslot), each slot a family
 * whose support meets the tile (ascending) with the 256-bit membership of
 * the tile's points (point p = 16 r + c, bit p % 64 of word p / 64; the
 * reference's box test, masked points excluded).  Coefficients are laid out
 * as in the reference: space frequency, time block (time frequency, time
 * centre), space centre, direction, cos/sin.  The time factors gt of a call
 * are T x nf x ntb (ntb: time blocks per frequency), row-major.
 * ------------------------------------------------------------------------- */
typedef struct vd_w3_basis vd_w3_basis;
typedef struct vd_w3_params {
  int ny, nx, nf, ntheta, ntb, nfam;
  const double* lon;
  const double* lat;
  const double* dxs;
  const double* kx;
  const double* ky;
  const int32_t* fam_f;
  const double* fam_lon;
  const double* fam_lat;
  const double* fam_cos;
  const int64_t* fbase;
  const int32_t* ffam0;
  const int32_t* fnfam;
  const int64_t* tslot;
  const int32_t* sfam;
  const uint64_t* sbits;
} vd_w3_params;

/* Upload the tables and build the tile blocks of the space matrix on the
 * current device (synchronises once). */
int vd_w3_create(const vd_w3_params* params, vd_w3_basis** basis);
int vd_w3_destroy(vd_w3_basis* basis);
/* Number of coefficients, of tile entries (the stored block rows: the tile
 * blocks hold 256 doubles per entry) and of waves. */
int vd_w3_info(const vd_w3_basis* basis, long long* nbasis, long long* nentries, long long* nwave);
/* Device workspace of one call (any T: the calls run 16 times per pass). */
int vd_w3_workspace_size(const vd_w3_basis* basis, size_t* bytes);
/* out[t] (T fields of ny x nx) = operg(time t, X); device pointers. */
int vd_w3_operg(vd_w3_basis* basis, int T, const double* gt, const double* X, double* out, void* work,
                void* stream);
/* adX += sum over the T times of operg_transpose(time t, ad[t]). */
int vd_w3_operg_adj(vd_w3_basis* basis, int T, const double* gt, const double* ad, double* adX,
                    void* work, void* stream);

/* ---------------------------------------------------------------------------
 * Merge of subwindow analyses onto the target grid.  Replaces the numerical
 * core of compute_weights_map (mapping/src/run_assimilation.py:351-412, the
 * regular-grid branch of _build_interpolator :414-446), merge_output_date
 * (:537-607, without the netCDF load/save and with kernel None) and the blend
 * of merge_time_windows_outputs (:732-830).
 *
 * Subwindow s has ascending 1-D axes slat (sny[s] values) and slon (snx[s]),
 * concatenated over s.  Per target point and subwindow: scipy
 * RegularGridInterpolator semantics (linear, NaN outside), the raised-cosine
 * weight (ones with a single subwindow), its sum over the subwindows.
 * ------------------------------------------------------------------------- */
typedef struct vd_merge vd_merge;
typedef struct vd_merge_params {
  int ny, nx;              /* target grid */
  const double* lon;       /* ny * nx target longitudes (degrees) */
  const double* lat;       /* ny * nx target latitudes */
  const uint8_t* mask;     /* ny * nx, 1 = masked (NaN in the output), or null */
  int nsub;                /* subwindows */
  const int32_t* sny;      /* nsub */
  const int32_t* snx;      /* nsub */
  const double* slat;      /* sum of sny: latitude axes, strictly ascending each */
  const double* slon;      /* sum of snx: longitude axes, strictly ascending each */
  const int32_t* tri_slot; /* nsub or null: k >= 0 marks subwindow s as an irregular grid
                              (run_assimilation.py:436-446, LinearNDInterpolator) whose
                              triangle tables are block k of tri_idx / tri_w; -1 a
                              rectilinear grid (its axes in slat / slon; an irregular
                              subwindow's axes are not read) */
  const int32_t* tri_idx;  /* (irregular subwindows) x ny*nx x 3: the vertices (flat
                              indices into the subwindow's sny x snx points) of the
                              Delaunay triangle holding each target point, -1 outside
                              the convex hull */
  const double* tri_w;     /* the same shape: the barycentric weights (c0, c1, c2) */
} vd_merge_params;

/* Upload the grids (host arrays) and build the interpolation tables and
 * weights on the current device.  An irregular subwindow's value at a
 * target point is ((0 + c0 v0) + c1 v1) + c2 v2 over its triangle, in
 * LinearNDInterpolator's order. */
int vd_merge_create(const vd_merge_params* params, vd_merge** merge);
int vd_merge_destroy(vd_merge* merge);
/* Total subwindow points (sum of sny * snx). */
int vd_merge_info(const vd_merge* merge, long long* sub_points);
/* Copy the interpolated weights (nsub x ny x nx, NaN outside a subwindow)
 * and their sum (ny x nx) into device buffers; either may be null. */
int vd_merge_weights(const vd_merge* merge, double* weights, double* weights_sum, void* stream);
/* out[t] (T fields of ny x nx) = merge of the subwindows' fields of date t.
 * fields: subwindow s's T fields of sny[s] x snx[s] start at
 * T * (points of the subwindows before s).  Device pointers. */
int vd_merge_apply(vd_merge* merge, int T, const double* fields, double* out, void* stream);
/* out = w1 * a + w2 * b over n values (the blend of two time windows,
 * run_assimilation.py:797-818). */
int vd_merge_blend(long long n, double w1, double w2, const double* a, const double* b, double* out,
                   void* stream);

/* ---------------------------------------------------------------------------
 * L3 observation operator (Obsop_interp_l3_jax, mapping/src/obsop.py:375-678).
 * One table per observation time:
 *   row_ptr[n_obs+1], col[nnz] (int32 flat grid index), w[nnz] (double)
 *   y[n_obs], err[n_obs] (double);
 * and its transpose for the deterministic adjoint:
 *   tcol[n_cols] (int32 grid index), tptr[n_cols+1], trow[nnz], tw[nnz].
 * ------------------------------------------------------------------------- */
#define VD_REDUCE_SCRATCH 1024 /* doubles of scratch a reduction needs */

typedef struct vd_obs_table {
  int n_obs, nnz, n_cols;
  const int32_t* row_ptr;
  const int32_t* col;
  const double* w;
  const double* y;
  const double* err;
  const int32_t* tcol;
  const int32_t* tptr;
  const int32_t* trow;
  const double* tw;
  double* partial;       /* device scratch of VD_REDUCE_SCRATCH doubles */
} vd_obs_table;

/* misfit = where(isnan(Hx - y), 0, Hx - y) / err for x of the plan precision
 * (obsop.py:610-628); valid[o] = 1 where Hx - y is not NaN.  If `jo` is not
 * NULL, sum(misfit^2) * inv_sigma_r2 is ADDED to *jo (device double) with a
 * deterministic reduction (tools_4Dvar.py:287-292). */
int vd_obs_misfit(const vd_obs_table* tab, int precision, const void* x, double* misfit,
                  uint8_t* valid, double* jo, double inv_sigma_r2, void* stream);

/* adx += H^T (valid * cot / err), the vjp of the misfit (obsop.py:630-678),
 * gathered per grid point in a fixed order (bitwise reproducible). */
int vd_obs_adjoint(const vd_obs_table* tab, int precision, const double* cot,
                   const uint8_t* valid, void* adx, void* stream);

/* Operator construction (Obsop_interp_l3_jax._sparse_op, obsop.py:417-460,
 * which queries a scipy KDTree built over the grid's cartesian points,
 * obsop.py:424-425).  vd_obs_grid_create takes the n_grid x 3 cartesian km of
 * the grid points (geo2cart, grid.py:61-98; host memory) and dmax (km), and
 * bins the points into cubes of side >= dmax on the device.
 * vd_obs_grid_query returns, for each of n_obs observations (device xyz,
 * n_obs x 3 cartesian km), the npix (1..8) nearest grid points within dmax in
 * ascending distance (D = sqrt((dx^2 + dy^2) + dz^2), the KDTree's rounding;
 * exact ties of D^2 go to the lower grid index): col (device int32, n_obs x
 * npix; -1 past the neighbours found), w (device double, the normalised
 * Gaussian weights exp(-D^2 / (2 (dmax/2)^2)) of obsop.py:436-447; 0 past
 * them) and dist (device double or NULL; +inf past them).  Observations with
 * non-finite coordinates get no neighbours (the Python operator raises on
 * them, as the reference's KDTree query does).  The reference's Npix == 1 quirk
 * (an empty operator, obsop.py:449-460) is the caller's. */
typedef struct vd_obs_grid vd_obs_grid;
int vd_obs_grid_create(const double* xyz, int64_t n_grid, double dmax, vd_obs_grid** grid);
int vd_obs_grid_destroy(vd_obs_grid* grid);
int vd_obs_grid_query(const vd_obs_grid* grid, const double* xyz, int n_obs, int npix, int32_t* col,
                      double* w, double* dist, void* stream);

/* ---------------------------------------------------------------------------
 * Overlapping tiles of QgmWithTiles (jqgm.py:1033-1258).  Tile t = ty * ntx +
 * tx starts at (ys[ty], xs[tx]) (device int32 arrays of nty and ntx starts,
 * each in 0..n - ts; jqgm.py:1098-1113) and spans ts x ts points; a tile
 * stack `flat` holds nty * ntx tiles of ts * ts values.  Fields are (ny, nx)
 * of the given precision; w is the ts x ts window and wsum the (ny, nx) sum of
 * the windows over the tiles (jqgm.py:1073-1081).
 * ------------------------------------------------------------------------- */
/* flat_t = (full + add)[tile t] (add may be NULL): the tile slices of
 * QgmWithTiles.step_tile (jqgm.py:1177-1187, the MDT added at :1229-1231). */
int vd_tiles_gather(int precision, int ny, int nx, int ts, int nty, int ntx, const int* ys,
                    const int* xs, const void* full, const void* add, void* flat, void* stream);
/* full = sum_t w flat_t / wsum - sub (sub may be NULL), tiles summed in tile
 * order: the blend of QgmWithTiles.step (jqgm.py:1233-1243). */
int vd_tiles_blend(int precision, int ny, int nx, int ts, int nty, int ntx, const int* ys,
                   const int* xs, const void* w, const void* wsum, const void* flat, const void* sub,
                   void* full, void* stream);
/* Transposes for the adjoint: flat_t = w full[tile t] / wsum[tile t] ... */
int vd_tiles_blend_adj(int precision, int ny, int nx, int ts, int nty, int ntx, const int* ys,
                       const int* xs, const void* w, const void* wsum, const void* full, void* flat,
                       void* stream);
/* ... and full = sum_t flat_t placed at tile t (overwrites full). */
int vd_tiles_gather_adj(int precision, int ny, int nx, int ts, int nty, int ntx, const int* ys,
                        const int* xs, const void* flat, void* full, void* stream);

/* ---------------------------------------------------------------------------
 * Small device helpers for the 4D-Var driver (tools_4Dvar.py:252-400).
 * ------------------------------------------------------------------------- */
/* *out += alpha * sum_i x[i] * y[i] (double, deterministic); `scratch`
 * holds VD_REDUCE_SCRATCH doubles of device memory. */
int vd_dot_f64(int64_t n, const double* x, const double* y, double alpha, double* out,
               double* scratch, void* stream);

/* The same dot product, allocating nothing: the partial sums go to a device
 * array inside the library (VD_REDUCE_SCRATCH doubles, 8 KiB, one copy per
 * device, held until the process ends), so the caller passes no scratch.  No
 * counterpart in the reference: it is for callers that cannot allocate device
 * memory (code under HIP-graph capture, serving loops, harnesses that forbid
 * hipMalloc).  The first call on a device looks the array's address up
 * (hipGetSymbolAddress); make it outside graph capture.  n must not exceed
 * VD_DOT_NOALLOC_MAX_N (2^36 doubles, more than the device memory) and any
 * n up to it fits, since a call uses at most VD_REDUCE_SCRATCH partials.  All
 * calls on one device share the array: calls that overlap on one device (on
 * different streams) race. */
#define VD_DOT_NOALLOC_MAX_N ((int64_t)1 << 36)
int vd_dot_f64_noalloc(int64_t n, const double* x, const double* y, double alpha, double* out,
                       void* stream);

/* y[i] = a * x[i] + b * y[i] in the given precision (x, y same precision). */
int vd_axpby(int precision, int64_t n, double a, const void* x, double b, void* y, void* stream);

/* ---------------------------------------------------------------------------
 * Multilayer rotating shallow-water model on the Arakawa C grid (class SW,
 * mapping/models/model_qgsw/sw.py:67-937): the constructor (:117-323, with
 * the masks of masks.py:28-159), step (:752-904, RK3-SSP), step_tgl
 * (:906-923, the reference's jax.jvp of step) and step_adj (:925-937, its
 * jax.vjp with respect to the initial state).
 *
 * FP64.  Arrays are the reference's, x before y, y contiguous, one plane per
 * layer: u (nl, nx+1, ny), v (nl, nx, ny+1), h (nl, nx, ny) in physical
 * units (velocities in m/s, h the layer-thickness perturbation in m); f on
 * the (nx+1) x (ny+1) vorticity grid.  Supported: per-layer constant depths,
 * scalar dx / dy, any binary ocean mask, slip coefficient, linear bottom
 * drag on the last layer, the smooth thickness clamp, wind stress on the top
 * layer, Laplacian viscosity of the velocities and diffusion of the
 * thickness with zero-flux edges.  The sponge, open boundaries and the
 * barotropic filter have no parameter (the reference's defaults: none of
 * them).  External forcing
 * (Fu, Fv, Fh, direct momentum mode) is an argument of the step calls.
 * ------------------------------------------------------------------------- */
typedef struct vd_sw_model vd_sw_model;
typedef struct vd_sw_params {
  int nx, ny, nl;
  double dx, dy, dt;
  const double* H;        /* nl layer depths (m) */
  const double* g_prime;  /* nl reduced gravities (m/s^2), g_prime[0] the surface one */
  const double* f;        /* (nx+1) x (ny+1) Coriolis parameter on the vorticity grid */
  const uint8_t* mask;    /* nx x ny, 1 = ocean; null: all ocean */
  double slip_coef;       /* 1 free slip, 0 no slip */
  double bottom_drag_coef;
  double h_min, h_min_sharpness; /* smooth clamp of the total thickness (defaults 0.1 m, 10) */
  double visc_coef, diff_coef;   /* Laplacian viscosity of u, v and diffusion of h (m^2/s); <= 0: none
                                    (add_diffusion, add_h_diffusion, sw.py:511-565) */
  const double* wind_u;   /* top-layer wind-stress tendency on the u grid ((nx+1) x ny, scaled units:
                             tau_x / (rho_water H_wind) dx, zero on land and on the boundary rows),
                             or null for none (add_wind_forcing, sw.py:567-642) */
  const double* wind_v;   /* the same on the v grid (nx x (ny+1)), or null */
} vd_sw_params;

/* Build the masks and upload the grid constants to the current device. */
int vd_sw_create(const vd_sw_params* params, vd_sw_model** model);
int vd_sw_destroy(vd_sw_model* model);
/* Number of values of u, v and h (all layers). */
int vd_sw_info(const vd_sw_model* model, long long* nu, long long* nv, long long* nh);
/* (du, dv, dh) = time derivatives at the SCALED state (u dx, v dy, h dx dy),
 * compute_time_derivatives (sw.py:723-750); device pointers. */
int vd_sw_rhs(vd_sw_model* model, const double* u, const double* v, const double* h, double* du,
              double* dv, double* dh, void* stream);
/* (u1, v1, h1) = nstep RK3-SSP steps from (u0, v0, h0); device pointers.
 * (Fu, Fv, Fh): external forcing in physical units per second, per layer
 * (u, v and h shapes), added as x += dt F after every step (sw.py:752-904,
 * the Fu / Fv / Fh arguments of SW.step); a null field is zero. */
int vd_sw_step(vd_sw_model* model, const double* u0, const double* v0, const double* h0, const double* Fu,
               const double* Fv, const double* Fh, double* u1, double* v1, double* h1, int nstep, void* stream);
/* (du1, dv1, dh1) = tangent of step at (u0, v0, h0; F) applied to (du0, dv0, dh0; dF)
 * (jax.jvp of step, sw.py:906-923, with the forcing among the differentiated
 * inputs as in Model_qgsw.jstep_tgl, mod.py:3222-3253); null forcing fields
 * and null forcing tangents are zero. */
int vd_sw_step_tgl(vd_sw_model* model, const double* u0, const double* v0, const double* h0, const double* Fu,
                   const double* Fv, const double* Fh, const double* du0, const double* dv0, const double* dh0,
                   const double* dFu, const double* dFv, const double* dFh, double* du1, double* dv1,
                   double* dh1, int nstep, void* stream);
/* (au0, av0, ah0) = adjoint of step at (u0, v0, h0; F) applied to (au1, av1, ah1),
 * and (aFu, aFv, aFh) = the forcing cotangents (overwritten; null outputs are
 * not computed) -- jax.vjp of step (sw.py:925-937) as Model_qgsw.jstep_adj
 * uses it (mod.py:3255-3290).  The trajectory is recomputed in segments of
 * about sqrt(nstep) steps; the model keeps a device tape of (segments +
 * segment length) states, allocated on the first call that needs it. */
int vd_sw_step_adj(vd_sw_model* model, const double* u0, const double* v0, const double* h0, const double* Fu,
                   const double* Fv, const double* Fh, const double* au1, const double* av1, const double* ah1,
                   double* au0, double* av0, double* ah0, double* aFu, double* aFv, double* aFh, int nstep,
                   void* stream);
/* vd_sw_step that also keeps the scaled states (u dx, v dy, h dx dy, one
 * buffer [u | v | h] of nu + nv + nh values each) before steps 0, stride,
 * 2 stride, ... into `keep` (ceil(nstep / stride) states); stride 0 keeps
 * every step's start state and its two RK3 stage states (3 nstep states):
 * the trajectory a 4D-Var driver's forward sweep leaves for the reverse
 * sweep.  No reference counterpart (the reference's jax.vjp of a chunk
 * recomputes it). */
int vd_sw_step_keep(vd_sw_model* model, const double* u0, const double* v0, const double* h0, const double* Fu,
                    const double* Fv, const double* Fh, double* u1, double* v1, double* h1, int nstep, int stride,
                    double* keep, void* stream);
/* vd_sw_step_adj from states kept by vd_sw_step_keep with the same forcing,
 * nstep and stride: no forward pass; stride 0 recomputes nothing, stride 1
 * only each step's two stage states, a larger stride also each run of
 * `stride` steps once. */
int vd_sw_step_adj_kept(vd_sw_model* model, const double* Fu, const double* Fv, const double* Fh,
                        const double* keep, int stride, const double* au1, const double* av1, const double* ah1,
                        double* au0, double* av0, double* ah0, double* aFu, double* aFv, double* aFh, int nstep,
                        void* stream);
/* Precision conversion: dst[i] = (dst type) src[i]; NaN -> 0 if nan_to_zero. */
int vd_convert(int src_precision, const void* src, int dst_precision, void* dst, int64_t n,
               int nan_to_zero, void* stream);

#ifdef __cplusplus
}
#endif
#endif /* VARDYN_H */
