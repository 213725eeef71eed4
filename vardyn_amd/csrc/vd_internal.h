// Internal declarations shared by the HIP translation units of libvardyn_hip.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

#include "../../include/vardyn.h"

namespace vd {

void set_error(const std::string& msg);

#define VD_HIP_TRY(expr)                                                        \
  do {                                                                          \
    hipError_t e_ = (expr);                                                     \
    if (e_ != hipSuccess) {                                                     \
      ::vd::set_error(std::string(#expr " failed: ") + hipGetErrorString(e_));  \
      return VD_ERR_HIP;                                                        \
    }                                                                           \
  } while (0)

#define VD_REQUIRE(cond, code, msg)  \
  do {                               \
    if (!(cond)) {                   \
      ::vd::set_error(msg);          \
      return (code);                 \
    }                                \
  } while (0)

constexpr int kTile = 32;  // smallest GEMM output tile and the K chunk

// GEMM tile configurations: BM x BN output tile per workgroup, ks K groups of
// four waves, and the kernel: 0 = single LDS buffer (k_gemm), 1 = double
// buffer with transposed B (k_gemm2).  Configuration 0 picks ks = 4, 2 or 1
// from K.
struct TileCfg {
  int bm, bn, ks, kern;
};
constexpr int kNumTileCfg = 11;
inline constexpr TileCfg kTileCfgs[kNumTileCfg] = {
    {32, 32, 4, 0}, {64, 64, 1, 0}, {64, 64, 2, 0}, {128, 64, 1, 0}, {128, 64, 2, 0},
    {32, 32, 4, 1}, {64, 64, 1, 1}, {64, 64, 2, 1}, {128, 64, 1, 1}, {64, 32, 2, 1},
    {128, 128, 1, 0}};
inline bool tile_cfg_fits(int c, int M, int N, int K) {
  return M % kTileCfgs[c].bm == 0 && N % kTileCfgs[c].bn == 0 &&
         (c == 0 || (K / kTile) % kTileCfgs[c].ks == 0);
}

inline int round_up(int v, int m) { return (v + m - 1) / m * m; }

// Physical constants of one QG problem, passed by value to every stencil kernel.
template <typename T>
struct Phys {
  int ny, nx;
  T cu;       // -g/f/(4 dy): u = cu * (4-point difference of h)
  T cv;       //  g/f/(4 dx)
  T inv6dx, inv6dy;
  T kx, ky;   // Kdiffus/dx^2, Kdiffus/dy^2 (0 when absent)
  const T* gfc2_b;  // per batch member: g f / c^2
  size_t nf;        // ny * nx: stride between batch members of a field
  T gf_dx2, gf_dy2;  // g/f/dx^2, g/f/dy^2
  const uint8_t* mask;  // 0 land, 1 ring/coast, 2 next to ring/coast, 3 active
  const T* beta;        // (ny, nx) beta at T points, or nullptr (f plane)
  const T* fbathy;      // (ny, nx) f * bathymetry PV term, or nullptr
  int has_kdiff;
};

// Device-side state of one plan.  Buffers are typed by the plan precision.
struct Plan {
  int ny, nx, my, mx, MY, MX;
  int nb = 1;   // batch members (independent problems on the same grid)
  int prec;     // VD_F32 / VD_F64
  int scheme;   // VD_EULER / VD_RK2 / VD_RK3
  int nstage;
  double dx, dy, dt, g, f0, c, kdiffus;
  int has_beta, has_kdiff, has_bathy, has_mdt;
  int device;
  size_t es;    // element size
  size_t nf;    // ny*nx
  size_t ng;    // MY*MX

  // DST by symmetric folding (even interior sizes): two half-size GEMMs per
  // sine-matrix product instead of one full one.
  int fold = 0;
  int gemm_cfg[4] = {0, 0, 0, 0};  // tile configuration of the four inversion GEMMs
  int hy = 0, hx = 0, HY = 0, HX = 0;
  void* SyF = nullptr;  // [2][HY][HY]: rows 2r+p of S_y, columns 0..hy-1
  void* SxF = nullptr;  // [2][HX][HX]: SxF[p][l][c] = S_x[l][2c+p]

  // split-product GEMMs on int8 MFMA (0 = off): digits per operand line,
  // digit planes [par][split_ns][lines][K] and line exponents [par][lines] of
  // the sine operand of the left (S_y) and right (S_x) stages, lines
  // allocated to split_line_alloc(sine lines) (zero lines past the product)
  int split_ns = 0;
  void* S8y = nullptr;
  void* S8x = nullptr;
  int* eSy = nullptr;
  int* eSx = nullptr;

  // constants
  void* Sy = nullptr;
  void* Sx = nullptr;
  double* lam_cx = nullptr;  // [MX] 2 (cos(pi (j+1)/(nx-1)) - 1)/dx^2, zero past mx
  double* lam_cy = nullptr;  // [MY] the same in y
  double* gfc2_d = nullptr;  // [nb] g f / c^2 per member, FP64
  void* gfc2_b = nullptr;   // [nb]
  uint8_t* mask = nullptr;
  void* beta = nullptr;
  void* fbathy = nullptr;
  void* mdt = nullptr;

  // scratch (one call at a time per plan)
  void* gT[3] = {nullptr, nullptr, nullptr};
  void* fld[40] = {};
  int nfld = 0;
};

// A ragged group: independent problems on grids of their own (one plan each,
// nbatch 1), advanced by shared launches.  Member fields are concatenated in
// member order in every flat buffer; a launch covers the first `nact` members.
template <typename T>
struct RgMember {
  Phys<T> P;
  size_t foff;         // element offset of the member's fields
  const T* mdt;        // the member's MDT field, or NULL
};
struct Group {
  int n = 0, prec = 0, scheme = 0, nstage = 0, fold = 0;
  std::vector<Plan*> plans;   // not owned
  std::vector<size_t> foff;   // n + 1 prefix offsets
  void* mem = nullptr;        // RgMember<T>[n]
  int4* stage_tiles = nullptr;  // (member, tile row, tile column) of 64 x 16 point tiles
  std::vector<int> stage_prefix;
  // the same for the adjoint stage's tiles, both shapes: [0] 64 x 16, [1] 32 x 32;
  // adj_narrow picks [1] (by default when every member is at most 32 points
  // wide, or by vd_group_autotune)
  int4* adj_tiles[2] = {nullptr, nullptr};
  std::vector<int> adj_prefix[2];
  bool adj_narrow = false;
  bool small = false;           // equal-size small FP64 members: the fused inversion applies
  bool use_small = false;       // ... and is used (the default; vd_group_autotune may turn it off)
  void* small_mem = nullptr;    // RgSmall[n]
  void* gargs[4] = {nullptr, nullptr, nullptr, nullptr};  // RgGemm<T>[n] per inversion stage
  // (member, product, tile) of the 32 x 32 GEMM tiles per stage and K-split
  // class (4, 2, 1 K groups: configuration 0's choice for each member)
  int4* gtiles[4][3] = {};
  std::vector<int> gprefix[4][3];
  // split-product groups (split_ns > 0): per-stage member descriptors and
  // (member, product, tile) tables for both tile shapes, [0] 64 sine x 32 data
  // lines and [1] 32 x 32, and the shape each stage uses (split_shape)
  int split_ns = 0;
  void* sargs[2][4] = {};
  int4* stiles[2][4] = {};
  std::vector<int> sprefix[2][4];
  int split_shape[4] = {0, 0, 0, 0};
};

// Named scratch fields inside Plan::fld.
enum Field {
  F_HBM = 0, F_QB,
  F_HA, F_QA, F_HB, F_QB2, F_HS1, F_QS1, F_HS2, F_QS2, F_K0, F_K1,
  F_DH, F_DQ, F_DH2, F_DQ2, F_DHS1, F_DQS1, F_DHS2, F_DQS2, F_DK0, F_DK1,
  F_LH, F_LQ, F_LQ1, F_LQ2, F_LHS, F_LQX,
  F_HOUT,
  F_ZERO,  // never written: the zero boundary field
  F_COUNT
};

// Launchers implemented in vd_kernels.hip -----------------------------------
template <typename T> Phys<T> make_phys(const Plan& p);

// Helmholtz inversion: out = hb + S_y (S_y (F - Fsub) S_x ./ Lambda) S_x, with
// mode selecting the epilogue (see InvMode).
enum InvMode { INV_FWD = 0, INV_TL = 1, INV_ACC = 2 };
template <typename T>
void launch_inversion(const Plan& p, const T* F, const T* Fsub, const T* hb, T* out, int mode,
                      hipStream_t s);
// One of the four GEMMs of the inversion with tile configuration cfg.
template <typename T>
void launch_gemm_stage(const Plan& p, int stage, const T* F, const T* Fsub, const T* hb, T* out,
                       int mode, int cfg, hipStream_t s);
// (M, N, K, products) of GEMM stage `stage` of a plan.
void gemm_stage_shape(const Plan& p, int stage, int* M, int* N, int* K, int* nz);
// Split-product stage on int8 MFMA (vd_split.hip), FP64 plans with split_ns > 0,
// and its tile configurations.
void launch_split_stage(const Plan& p, int stage, const double* F, const double* Fsub,
                        const double* hb, double* out, int mode, int cfg, hipStream_t s);
int split_num_cfgs();
int split_max_k();
// Sine lines allocated per digit plane: a multiple of 64 (the zero lines past
// a product's own let ragged groups run every member on 64-line tiles).
inline int split_line_alloc(int lines) { return (lines + 63) / 64 * 64; }
int split_group_build(Group& g);
void launch_inversion_split_rg(const Group& g, int nact, const double* F, const double* Fsub,
                               const double* hb, double* out, int mode, hipStream_t s);
// One split-product stage of a ragged group with tile shape `shape` (0: 64 x 32, 1: 32 x 32).
void launch_split_stage_rg(const Group& g, int stage, int shape, int nact, const double* F,
                           const double* Fsub, const double* hb, double* out, int mode, hipStream_t s);
bool split_cfg_fits(const Plan& p, int stage, int c);
int split_default_cfg(const Plan& p, int stage);

// Prepare a segment: hcur = h0 + mdt, hbm = hb + mdt, q0 = h2pv(hcur, hbm), qb = h2pv(hbm, hbm).
template <typename T>
void launch_prep(const Plan& p, const T* h0, const T* hb, T* hcur, T* hbm, T* q0, T* qb,
                 hipStream_t s);
template <typename T>
void launch_h2pv(const Plan& p, const T* h, const T* hb, T* q, hipStream_t s);
template <typename T>
void launch_fill_land(const Plan& p, T* f, T value, hipStream_t s);
// h1 = NaN on land, else h - mdt.
template <typename T>
void launch_finish(const Plan& p, const T* h, T* h1, hipStream_t s);

// Forward stage: k = rhs(h2uv(h), q); out = qbase + sum_i c_i k_i where the
// k_i are (kprev0, kprev1, k) with coefficients (c0, c1, c2).  Optionally stores k.
template <typename T>
void launch_stage(const Plan& p, const T* h, const T* q, const T* qbase, const T* kprev0,
                  const T* kprev1, T c0, T c1, T c2, T* qout, T* kstore, hipStream_t s);

// Tangent-linear stage around (h, q).
template <typename T>
void launch_stage_tl(const Plan& p, const T* h, const T* q, const T* dh, const T* dq,
                     const T* dqbase, const T* dkprev0, const T* dkprev1, T c0, T c1, T c2,
                     T* dqout, T* dkstore, hipStream_t s);
template <typename T>
void launch_prep_tl(const Plan& p, const T* dh0, const T* hcur, T* dhcur, T* dq0, hipStream_t s);
template <typename T>
void launch_finish_tl(const Plan& p, const T* dh, T* dh1, hipStream_t s);

// Adjoint stage around (h, q): ak = mask * (a0 s0 + a1 s1 + a2 s2);
// lq_out = acc0 + acc1 + acc2 + adv^T ak (+ diffusion^T); lh_out = h2uv^T(...).
template <typename T>
void launch_stage_adj(const Plan& p, const T* h, const T* q, const T* s0, const T* s1,
                      const T* s2, T a0, T a1, T a2, const T* acc0, const T* acc1,
                      const T* acc2, T* lq_out, T* lh_out, hipStream_t s);
// Ragged groups (vd_kernels.hip): build the device tables of a group whose
// plans are set, free them, and the three per-step launches over the first
// nact members.
template <typename T> int group_build(Group& g);
void group_release(Group& g);
template <typename T>
void launch_stage_rg(const Group& g, int nact, const T* h, const T* q, const T* qbase,
                     const T* kprev0, const T* kprev1, T c0, T c1, T c2, T* qout, T* kstore,
                     hipStream_t s);
template <typename T>
void launch_inversion_rg(const Group& g, int nact, const T* F, const T* Fsub, const T* hb, T* out,
                         int mode, hipStream_t s);
template <typename T>
void launch_stage_adj_rg(const Group& g, int nact, const T* h, const T* q, const T* s0,
                         const T* s1, const T* s2, T a0, T a1, T a2, const T* acc0,
                         const T* acc1, const T* acc2, T* lq_out, T* lh_out, hipStream_t s);
// Segment start/end of members 0..nact-1 of a group (the plan launches'
// arithmetic per member; hbm and qb may be NULL).
template <typename T>
void launch_prep_rg(const Group& g, int nact, const T* h0, const T* hb, T* hcur, T* hbm, T* q0,
                    T* qb, hipStream_t s);
template <typename T>
void launch_finish_rg(const Group& g, int nact, const T* h, T* h1, hipStream_t s);
template <typename T>
void launch_adj_init_rg(const Group& g, int nact, const T* adh1, T* lh, T* lq, hipStream_t s);
template <typename T>
void launch_adj_final_rg(const Group& g, int nact, const T* hcur, const T* lh, const T* lq,
                         T* adh0, hipStream_t s);
// lh = adh1 with zeros on land; lq = 0.
template <typename T>
void launch_adj_init(const Plan& p, const T* adh1, T* lh, T* lq, hipStream_t s);
// adh0 = lh + h2pv^T(lq) at the NaN pattern of hcur.
template <typename T>
void launch_adj_final(const Plan& p, const T* hcur, const T* lh, const T* lq, T* adh0,
                      hipStream_t s);

}  // namespace vd

// Opaque handles of the C ABI.
struct vd_plan {
  vd::Plan p;
};
struct vd_group {
  vd::Group g;
};
