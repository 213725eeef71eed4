// C ABI of libvardyn_hip: plan construction and the step / tangent-linear /
// adjoint drivers of the QG model (mapping/models/model_qg1l/jqgm.py,
// datlas-ocean/VarDyn).
#include "vd_internal.h"

#include <utility>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

namespace vd {

static thread_local std::string g_last_error;
void set_error(const std::string& msg) { g_last_error = msg; }

// Ring mask of jqgm.py:179-193 (no land).
static void ring_mask(int ny, int nx, std::vector<uint8_t>& m) {
  m.assign((size_t)ny * nx, 3);
  auto at = [&](int i, int j) -> uint8_t& { return m[(size_t)i * nx + j]; };
  for (int j = 0; j < nx; ++j) { at(0, j) = 1; at(ny - 1, j) = 1; }
  for (int i = 0; i < ny; ++i) { at(i, 0) = 1; at(i, nx - 1) = 1; }
  for (int j = 1; j < nx - 1; ++j) { at(1, j) = 2; at(ny - 2, j) = 2; at(ny - 3, j) = 2; }
  for (int i = 1; i < ny - 1; ++i) { at(i, 1) = 2; at(i, nx - 2) = 2; at(i, nx - 3) = 2; }
}

template <typename T>
static int upload(const std::vector<double>& h, void** dptr) {
  std::vector<T> t(h.size());
  for (size_t i = 0; i < h.size(); ++i) t[i] = static_cast<T>(h[i]);
  VD_HIP_TRY(hipMalloc(dptr, t.size() * sizeof(T)));
  VD_HIP_TRY(hipMemcpy(*dptr, t.data(), t.size() * sizeof(T), hipMemcpyHostToDevice));
  return VD_OK;
}

// Digit planes of a sine operand for the split-product GEMMs (vd_split.hip):
// op(par, line, k) for par < npar, line < lines, k < K is written as ns
// balanced base-256 digits per element, the bytes of (Y + C) ^ C with
// Y = rint(x 2^(8 ns - e)), e = ilogb(max |line|) + 3 and C = 0x80...80 --
// the split the kernel applies to its data operand.  Planes [par][ns]
// (digit 1 first) of lines x K bytes each, stored in MFMA fragment order:
// element (line, k) at ((line/32 * K/32 + k/32) * 64 + lane) * 16 + k%16 with
// lane = line%32 + 32 (k%32 / 16), the 16 bytes lane `lane` of a wave loads
// for its 32-line block and K step (lines and K are multiples of 32).
// Exponents [par][lines].  The planes hold split_line_alloc(lines) lines,
// the ones past `lines` zero.
template <typename Op>
static int upload_split(int npar, int lines, int K, int ns, Op op, void** d8, int** de) {
  const int LA = split_line_alloc(lines);
  std::vector<int8_t> pl((size_t)npar * ns * LA * K, 0);
  std::vector<int> ex((size_t)npar * LA, 0);
  const unsigned long long C = 0x8080808080808080ull >> (8 * (8 - ns));
  for (int par = 0; par < npar; ++par)
    for (int ln = 0; ln < lines; ++ln) {
      double m = 0.0;
      for (int k = 0; k < K; ++k) m = std::max(m, std::fabs(op(par, ln, k)));
      if (m < 0x1p-900) continue;  // zero line: zero digits
      const int e = std::ilogb(m) + 3;
      ex[(size_t)par * LA + ln] = e;
      for (int k = 0; k < K; ++k) {
        const long long Y = std::llrint(std::ldexp(op(par, ln, k), 8 * ns - e));
        const unsigned long long Z = ((unsigned long long)Y + C) ^ C;
        const size_t frag = ((size_t)(ln / 32) * (K / 32) + k / 32) * 64 + (ln % 32) + 32 * ((k % 32) / 16);
        for (int d = 0; d < ns; ++d)
          pl[((size_t)par * ns + d) * LA * K + frag * 16 + k % 16] = (int8_t)(uint8_t)(Z >> (8 * (ns - 1 - d)));
      }
    }
  VD_HIP_TRY(hipMalloc(d8, pl.size()));
  VD_HIP_TRY(hipMemcpy(*d8, pl.data(), pl.size(), hipMemcpyHostToDevice));
  VD_HIP_TRY(hipMalloc((void**)de, ex.size() * sizeof(int)));
  VD_HIP_TRY(hipMemcpy(*de, ex.data(), ex.size() * sizeof(int), hipMemcpyHostToDevice));
  return VD_OK;
}

static void release_plan_buffers(Plan& p) {
  hipFree(p.SyF); hipFree(p.SxF); hipFree(p.gfc2_b);
  hipFree(p.S8y); hipFree(p.S8x); hipFree(p.eSy); hipFree(p.eSx);
  hipFree(p.Sy); hipFree(p.Sx); hipFree(p.lam_cx); hipFree(p.lam_cy); hipFree(p.gfc2_d); hipFree(p.mask);
  hipFree(p.beta); hipFree(p.fbathy); hipFree(p.mdt);
  for (auto& g : p.gT) hipFree(g);
  for (auto& f : p.fld) hipFree(f);
}

template <typename T>
static int build_constants(Plan& p, const vd_qg_params& prm) {
  // Orthonormal DST-I matrices S_m[j][k] = sqrt(2/(m+1)) sin(pi (j+1)(k+1)/(m+1)),
  // zero-padded to a multiple of the GEMM tile (jqgm.py:75-81).
  auto sine = [](int m, int M) {
    std::vector<double> S((size_t)M * M, 0.0);
    const double a = std::sqrt(2.0 / (m + 1));
    for (int j = 0; j < m; ++j)
      for (int k = 0; k < m; ++k) {
        // reduce (j+1)(k+1) mod 2(m+1) to keep the sine argument small
        const long r = ((long)(j + 1) * (k + 1)) % (2L * (m + 1));
        S[(size_t)j * M + k] = a * std::sin(M_PI * (double)r / (m + 1));
      }
    return S;
  };
  int rc;
  const std::vector<double> SyP = sine(p.my, p.MY), SxP = sine(p.mx, p.MX);
  if ((rc = upload<T>(SyP, &p.Sy)) != VD_OK) return rc;
  if ((rc = upload<T>(SxP, &p.Sx)) != VD_OK) return rc;
  std::vector<double> fy, fx;
  if (p.fold) {
    // SyF[p][r][k] = S_y[2r+p][k], SxF[p][l][c] = S_x[l][2c+p] for r, k, l, c < m/2
    const std::vector<double> Sy = sine(p.my, p.my), Sx = sine(p.mx, p.mx);
    fy.assign(2 * (size_t)p.HY * p.HY, 0.0);
    fx.assign(2 * (size_t)p.HX * p.HX, 0.0);
    for (int par = 0; par < 2; ++par) {
      for (int r = 0; r < p.hy; ++r)
        for (int k = 0; k < p.hy; ++k)
          fy[((size_t)par * p.HY + r) * p.HY + k] = Sy[(size_t)(2 * r + par) * p.my + k];
      for (int l = 0; l < p.hx; ++l)
        for (int c = 0; c < p.hx; ++c)
          fx[((size_t)par * p.HX + l) * p.HX + c] = Sx[(size_t)l * p.mx + 2 * c + par];
    }
    if ((rc = upload<T>(fy, &p.SyF)) != VD_OK) return rc;
    if ((rc = upload<T>(fx, &p.SxF)) != VD_OK) return rc;
  }
  if (p.split_ns > 0) {
    // the sine operand of the split GEMMs: rows of S_y (left stages) and
    // columns of S_x (right stages), K contiguous
    const int ns = p.split_ns;
    if (p.fold) {
      const int HY = p.HY, HX = p.HX;
      rc = upload_split(2, HY, HY, ns, [&](int par, int r, int k) { return fy[((size_t)par * HY + r) * HY + k]; },
                        &p.S8y, &p.eSy);
      if (rc != VD_OK) return rc;
      rc = upload_split(2, HX, HX, ns, [&](int par, int c, int l) { return fx[((size_t)par * HX + l) * HX + c]; },
                        &p.S8x, &p.eSx);
    } else {
      const int MY = p.MY, MX = p.MX;
      rc = upload_split(1, MY, MY, ns, [&](int, int j, int k) { return SyP[(size_t)j * MY + k]; }, &p.S8y, &p.eSy);
      if (rc != VD_OK) return rc;
      rc = upload_split(1, MX, MX, ns, [&](int, int j, int l) { return SxP[(size_t)l * MX + j]; }, &p.S8x, &p.eSx);
    }
    if (rc != VD_OK) return rc;
  }
  // Lambda = g/f [2(cos(pi(j+1)/(nx-1))-1)/dx^2 + 2(cos(pi(i+1)/(ny-1))-1)/dy^2] - g f/c^2
  // (jqgm.py:165-172) is separable: the x and y terms as two tables and g f/c^2 per batch
  // member (c may differ); the GEMM epilogue forms 1/Lambda from them.
  std::vector<double> cx(p.MX, 0.0), cy(p.MY, 0.0), gfc2(p.nb);
  for (int j = 0; j < p.mx; ++j) cx[j] = 2 * (std::cos(M_PI / (p.nx - 1) * (j + 1)) - 1) / (p.dx * p.dx);
  for (int i = 0; i < p.my; ++i) cy[i] = 2 * (std::cos(M_PI / (p.ny - 1) * (i + 1)) - 1) / (p.dy * p.dy);
  for (int b = 0; b < p.nb; ++b) {
    const double cb = prm.c_batch ? prm.c_batch[b] : p.c;
    VD_REQUIRE(cb > 0, VD_ERR_INVALID_ARG, "c_batch values must be positive");
    gfc2[b] = p.g * p.f0 / (cb * cb);
  }
  {
    void *dcx = nullptr, *dcy = nullptr, *dg = nullptr;
    if ((rc = upload<double>(cx, &dcx)) != VD_OK) return rc;
    p.lam_cx = static_cast<double*>(dcx);
    if ((rc = upload<double>(cy, &dcy)) != VD_OK) return rc;
    p.lam_cy = static_cast<double*>(dcy);
    if ((rc = upload<double>(gfc2, &dg)) != VD_OK) return rc;
    p.gfc2_d = static_cast<double*>(dg);
  }
  if ((rc = upload<T>(gfc2, &p.gfc2_b)) != VD_OK) return rc;

  std::vector<uint8_t> mask;
  if (prm.mask) {
    mask.resize(p.nf);
    for (size_t i = 0; i < p.nf; ++i) {
      VD_REQUIRE(prm.mask[i] >= 0 && prm.mask[i] <= 5, VD_ERR_INVALID_ARG, "mask values must be 0..5");
      mask[i] = (uint8_t)prm.mask[i];
    }
  } else {
    ring_mask(p.ny, p.nx, mask);
  }
  // The stencils read up to two cells around an active (mask 3) point and
  // one around a Laplacian-only (mask 4) point: active points lie in rows and
  // columns 2..n-3 (the reference's full-domain ring keeps them in 2..n-4;
  // QgmWithTiles tiles use 2..n-3), mask-4 points in 1..n-2.
  for (int i = 0; i < p.ny; ++i)
    for (int j = 0; j < p.nx; ++j) {
      const uint8_t m = mask[(size_t)i * p.nx + j];
      if (m == 3)
        VD_REQUIRE(i >= 2 && i <= p.ny - 3 && j >= 2 && j <= p.nx - 3, VD_ERR_INVALID_ARG,
                   "mask value 3 (active) is only allowed in rows/columns 2..n-3");
      if (m == 4)
        VD_REQUIRE(i >= 1 && i <= p.ny - 2 && j >= 1 && j <= p.nx - 2, VD_ERR_INVALID_ARG,
                   "mask value 4 is only allowed in rows/columns 1..n-2");
    }
  VD_HIP_TRY(hipMalloc((void**)&p.mask, p.nf));
  VD_HIP_TRY(hipMemcpy(p.mask, mask.data(), p.nf, hipMemcpyHostToDevice));

  if (p.has_beta) {
    std::vector<double> b(p.nf, 0.0);
    for (int i = 1; i < p.ny - 1; ++i)
      for (int j = 0; j < p.nx; ++j) b[(size_t)i * p.nx + j] = prm.beta[(size_t)(i - 1) * p.nx + j];
    if ((rc = upload<T>(b, &p.beta)) != VD_OK) return rc;
  }
  if (p.has_bathy) {
    std::vector<double> b(prm.fbathy, prm.fbathy + p.nf);
    if ((rc = upload<T>(b, &p.fbathy)) != VD_OK) return rc;
  }
  if (p.has_mdt) {
    std::vector<double> b(prm.mdt, prm.mdt + p.nf);
    if ((rc = upload<T>(b, &p.mdt)) != VD_OK) return rc;
  }
  for (auto& g : p.gT) {
    VD_HIP_TRY(hipMalloc(&g, p.nb * p.ng * sizeof(T)));
    VD_HIP_TRY(hipMemset(g, 0, p.nb * p.ng * sizeof(T)));
  }
  for (int f = 0; f < F_COUNT; ++f) {
    VD_HIP_TRY(hipMalloc(&p.fld[f], p.nb * p.nf * sizeof(T)));
    VD_HIP_TRY(hipMemset(p.fld[f], 0, p.nb * p.nf * sizeof(T)));
  }
  p.nfld = F_COUNT;
  VD_HIP_TRY(hipDeviceSynchronize());
  return VD_OK;
}

template <typename T>
static inline T* fld(const Plan& p, int id) { return static_cast<T*>(p.fld[id]); }

// Tape slot (step n, stage st): two fields (h, q).
template <typename T>
struct Tape {
  T* base;
  size_t nf;
  int S;
  T* h(int n, int st) const { return base + ((size_t)(n * S + st) * 2) * nf; }
  T* q(int n, int st) const { return base + ((size_t)(n * S + st) * 2 + 1) * nf; }
  // nf here is the size of one batched field: nb * ny * nx
};

// ---------------------------------------------------------------------------
// Forward: Qgm.step (jqgm.py:517-557) with one_step/euler/rk2/rk3 (:398-503).
// ---------------------------------------------------------------------------
template <typename T>
static int qg_forward(Plan& p, const T* h0, const T* hb, T* h1, int nstep, T* tape_ptr,
                      hipStream_t s) {
  const T dt = T(p.dt);
  Tape<T> tp{tape_ptr, p.nb * p.nf, p.nstage};
  T* hbm = fld<T>(p, F_HBM);
  T* qb = fld<T>(p, F_QB);
  T* hc = tape_ptr ? tp.h(0, 0) : fld<T>(p, F_HA);
  T* qc = tape_ptr ? tp.q(0, 0) : fld<T>(p, F_QA);
  // A NULL boundary field means zero: then (without an MDT) h_b and q_b vanish
  // and the inversion skips them.
  const bool zero_b = hb == nullptr && !p.has_mdt;
  if (!hb) hb = fld<T>(p, F_ZERO);
  const T* qbi = zero_b ? nullptr : qb;
  const T* hbi = zero_b ? nullptr : hbm;
  launch_prep<T>(p, h0, hb, hc, hbm, qc, qb, s);
  for (int n = 0; n < nstep; ++n) {
    T *hn, *qn, *hs1, *qs1, *hs2, *qs2;
    if (tape_ptr) {
      hn = tp.h(n + 1, 0); qn = tp.q(n + 1, 0);
      hs1 = p.nstage > 1 ? tp.h(n, 1) : nullptr; qs1 = p.nstage > 1 ? tp.q(n, 1) : nullptr;
      hs2 = p.nstage > 2 ? tp.h(n, 2) : nullptr; qs2 = p.nstage > 2 ? tp.q(n, 2) : nullptr;
    } else {
      const bool even = (n % 2) == 0;
      hn = fld<T>(p, even ? F_HB : F_HA); qn = fld<T>(p, even ? F_QB2 : F_QA);
      hs1 = fld<T>(p, F_HS1); qs1 = fld<T>(p, F_QS1);
      hs2 = fld<T>(p, F_HS2); qs2 = fld<T>(p, F_QS2);
    }
    if (p.scheme == VD_EULER) {
      launch_stage<T>(p, hc, qc, qc, nullptr, nullptr, 0, 0, dt, qn, nullptr, s);
    } else if (p.scheme == VD_RK2) {
      launch_stage<T>(p, hc, qc, qc, nullptr, nullptr, 0, 0, T(0.5) * dt, qs1, nullptr, s);
      launch_inversion<T>(p, qs1, qbi, hbi, hs1, INV_FWD, s);
      launch_stage<T>(p, hs1, qs1, qc, nullptr, nullptr, 0, 0, dt, qn, nullptr, s);
    } else {
      T* k0 = fld<T>(p, F_K0);
      T* k1 = fld<T>(p, F_K1);
      launch_stage<T>(p, hc, qc, qc, nullptr, nullptr, 0, 0, T(0.5) * dt, qs1, k0, s);
      launch_inversion<T>(p, qs1, qbi, hbi, hs1, INV_FWD, s);
      launch_stage<T>(p, hs1, qs1, qc, k0, nullptr, -dt, 0, T(2) * dt, qs2, k1, s);
      launch_inversion<T>(p, qs2, qbi, hbi, hs2, INV_FWD, s);
      launch_stage<T>(p, hs2, qs2, qc, k0, k1, dt / T(6), T(4) * dt / T(6), dt / T(6), qn,
                      nullptr, s);
    }
    launch_inversion<T>(p, qn, qbi, hbi, hn, INV_FWD, s);
    hc = hn;
    qc = qn;
  }
  launch_finish<T>(p, hc, h1, s);
  VD_HIP_TRY(hipGetLastError());
  return VD_OK;
}

// ---------------------------------------------------------------------------
// Tangent linear (jvp of step, jqgm.py:559-563): primal and tangent advance
// together, stage by stage.
// ---------------------------------------------------------------------------
template <typename T>
static int qg_tangent(Plan& p, const T* dh0, const T* h0, const T* hb, T* dh1, T* h1, int nstep,
                      hipStream_t s) {
  const T dt = T(p.dt);
  T* hbm = fld<T>(p, F_HBM);
  T* qb = fld<T>(p, F_QB);
  T* hc = fld<T>(p, F_HA);
  T* qc = fld<T>(p, F_QA);
  T* dhc = fld<T>(p, F_DH);
  T* dqc = fld<T>(p, F_DQ);
  // A NULL boundary field means zero: then (without an MDT) h_b and q_b vanish
  // and the inversion skips them.
  const bool zero_b = hb == nullptr && !p.has_mdt;
  if (!hb) hb = fld<T>(p, F_ZERO);
  const T* qbi = zero_b ? nullptr : qb;
  const T* hbi = zero_b ? nullptr : hbm;
  launch_prep<T>(p, h0, hb, hc, hbm, qc, qb, s);
  launch_prep_tl<T>(p, dh0, hc, dhc, dqc, s);
  T *hs1 = fld<T>(p, F_HS1), *qs1 = fld<T>(p, F_QS1), *hs2 = fld<T>(p, F_HS2),
    *qs2 = fld<T>(p, F_QS2);
  T *dhs1 = fld<T>(p, F_DHS1), *dqs1 = fld<T>(p, F_DQS1), *dhs2 = fld<T>(p, F_DHS2),
    *dqs2 = fld<T>(p, F_DQS2);
  for (int n = 0; n < nstep; ++n) {
    const bool even = (n % 2) == 0;
    T* hn = fld<T>(p, even ? F_HB : F_HA);
    T* qn = fld<T>(p, even ? F_QB2 : F_QA);
    T* dhn = fld<T>(p, even ? F_DH2 : F_DH);
    T* dqn = fld<T>(p, even ? F_DQ2 : F_DQ);
    if (p.scheme == VD_EULER) {
      launch_stage<T>(p, hc, qc, qc, nullptr, nullptr, 0, 0, dt, qn, nullptr, s);
      launch_stage_tl<T>(p, hc, qc, dhc, dqc, dqc, nullptr, nullptr, 0, 0, dt, dqn, nullptr, s);
    } else if (p.scheme == VD_RK2) {
      const T h2 = T(0.5) * dt;
      launch_stage<T>(p, hc, qc, qc, nullptr, nullptr, 0, 0, h2, qs1, nullptr, s);
      launch_stage_tl<T>(p, hc, qc, dhc, dqc, dqc, nullptr, nullptr, 0, 0, h2, dqs1, nullptr, s);
      launch_inversion<T>(p, qs1, qbi, hbi, hs1, INV_FWD, s);
      launch_inversion<T>(p, dqs1, nullptr, nullptr, dhs1, INV_TL, s);
      launch_stage<T>(p, hs1, qs1, qc, nullptr, nullptr, 0, 0, dt, qn, nullptr, s);
      launch_stage_tl<T>(p, hs1, qs1, dhs1, dqs1, dqc, nullptr, nullptr, 0, 0, dt, dqn, nullptr,
                         s);
    } else {
      T *k0 = fld<T>(p, F_K0), *k1 = fld<T>(p, F_K1);
      T *dk0 = fld<T>(p, F_DK0), *dk1 = fld<T>(p, F_DK1);
      const T h2 = T(0.5) * dt;
      launch_stage<T>(p, hc, qc, qc, nullptr, nullptr, 0, 0, h2, qs1, k0, s);
      launch_stage_tl<T>(p, hc, qc, dhc, dqc, dqc, nullptr, nullptr, 0, 0, h2, dqs1, dk0, s);
      launch_inversion<T>(p, qs1, qbi, hbi, hs1, INV_FWD, s);
      launch_inversion<T>(p, dqs1, nullptr, nullptr, dhs1, INV_TL, s);
      launch_stage<T>(p, hs1, qs1, qc, k0, nullptr, -dt, 0, T(2) * dt, qs2, k1, s);
      launch_stage_tl<T>(p, hs1, qs1, dhs1, dqs1, dqc, dk0, nullptr, -dt, 0, T(2) * dt, dqs2, dk1,
                         s);
      launch_inversion<T>(p, qs2, qbi, hbi, hs2, INV_FWD, s);
      launch_inversion<T>(p, dqs2, nullptr, nullptr, dhs2, INV_TL, s);
      const T c1 = dt / T(6), c2 = T(4) * dt / T(6);
      launch_stage<T>(p, hs2, qs2, qc, k0, k1, c1, c2, c1, qn, nullptr, s);
      launch_stage_tl<T>(p, hs2, qs2, dhs2, dqs2, dqc, dk0, dk1, c1, c2, c1, dqn, nullptr, s);
    }
    launch_inversion<T>(p, qn, qbi, hbi, hn, INV_FWD, s);
    launch_inversion<T>(p, dqn, nullptr, nullptr, dhn, INV_TL, s);
    hc = hn; qc = qn; dhc = dhn; dqc = dqn;
  }
  launch_finish_tl<T>(p, dhc, dh1, s);
  if (h1) launch_finish<T>(p, hc, h1, s);
  VD_HIP_TRY(hipGetLastError());
  return VD_OK;
}

// ---------------------------------------------------------------------------
// Adjoint (vjp of step, jqgm.py:565-569): reverse sweep over the tape.
// ---------------------------------------------------------------------------
template <typename T>
static int qg_adjoint(Plan& p, const T* adh1, const T* h0, const T* hb, T* adh0, int nstep,
                      T* tape_ptr, int tape_valid, hipStream_t s) {
  int rc;
  if (!tape_valid) {
    if ((rc = qg_forward<T>(p, h0, hb, fld<T>(p, F_HOUT), nstep, tape_ptr, s)) != VD_OK) return rc;
  }
  const T dt = T(p.dt);
  Tape<T> tp{tape_ptr, p.nb * p.nf, p.nstage};
  T* lh = fld<T>(p, F_LH);
  T* lq = fld<T>(p, F_LQ);
  T* lq1 = fld<T>(p, F_LQ1);
  T* lq2 = fld<T>(p, F_LQ2);
  T* lhs = fld<T>(p, F_LHS);
  // The adjoint stage reads its cotangents (s0..s2) at neighbouring points, so
  // it must not write into one of them: the Euler stage and the last RK3
  // stage, whose result replaces lq, write lqx and the two are swapped.
  T* lqx = fld<T>(p, F_LQX);
  launch_adj_init<T>(p, adh1, lh, lq, s);
  for (int n = nstep - 1; n >= 0; --n) {
    // h_{n+1} = pv2h(q_{n+1}): lq += P lh
    launch_inversion<T>(p, lh, nullptr, nullptr, lq, INV_ACC, s);
    if (p.scheme == VD_EULER) {
      launch_stage_adj<T>(p, tp.h(n, 0), tp.q(n, 0), lq, nullptr, nullptr, dt, 0, 0, lq, nullptr,
                          nullptr, lqx, lh, s);
      std::swap(lq, lqx);
    } else if (p.scheme == VD_RK2) {
      launch_stage_adj<T>(p, tp.h(n, 1), tp.q(n, 1), lq, nullptr, nullptr, dt, 0, 0, nullptr,
                          nullptr, nullptr, lq1, lhs, s);
      launch_inversion<T>(p, lhs, nullptr, nullptr, lq1, INV_ACC, s);
      launch_stage_adj<T>(p, tp.h(n, 0), tp.q(n, 0), lq1, nullptr, nullptr, T(0.5) * dt, 0, 0, lq,
                          lq1, nullptr, lq, lh, s);
    } else {
      launch_stage_adj<T>(p, tp.h(n, 2), tp.q(n, 2), lq, nullptr, nullptr, dt / T(6), 0, 0,
                          nullptr, nullptr, nullptr, lq2, lhs, s);
      launch_inversion<T>(p, lhs, nullptr, nullptr, lq2, INV_ACC, s);
      launch_stage_adj<T>(p, tp.h(n, 1), tp.q(n, 1), lq, lq2, nullptr, T(4) * dt / T(6),
                          T(2) * dt, 0, nullptr, nullptr, nullptr, lq1, lhs, s);
      launch_inversion<T>(p, lhs, nullptr, nullptr, lq1, INV_ACC, s);
      launch_stage_adj<T>(p, tp.h(n, 0), tp.q(n, 0), lq, lq2, lq1, dt / T(6), -dt, T(0.5) * dt,
                          lq, lq2, lq1, lqx, lh, s);
      std::swap(lq, lqx);
    }
  }
  launch_adj_final<T>(p, tp.h(0, 0), lh, lq, adh0, s);
  VD_HIP_TRY(hipGetLastError());
  return VD_OK;
}

void gemm_stage_shape(const Plan& p, int stage, int* M, int* N, int* K, int* nz) {
  if (p.fold) {
    const bool left = stage == 0 || stage == 2;  // S_y from the left
    *M = left ? p.HY : p.MY;
    *N = left ? p.MX : p.HX;
    *K = left ? p.HY : p.HX;
    *nz = 2 * p.nb;
  } else {
    *M = p.MY;
    *N = p.MX;
    *K = (stage == 0 || stage == 2) ? p.MY : p.MX;
    *nz = p.nb;
  }
}

// Largest tile that still gives at least one workgroup per CU (256 CUs).
static int default_tile_cfg(const Plan& p, int stage) {
  int M, N, K, nz;
  gemm_stage_shape(p, stage, &M, &N, &K, &nz);
  const int cand[3] = {3, 1, 0};  // 128x64, 64x64, 32x32
  for (int c : cand) {
    if (!tile_cfg_fits(c, M, N, K)) continue;
    if (c == 0 || (long)nz * (M / kTileCfgs[c].bm) * (N / kTileCfgs[c].bn) >= 256) return c;
  }
  return 0;
}

template <typename T>
static int autotune(Plan& p, int reps, hipStream_t s) {
  hipEvent_t e0, e1;
  VD_HIP_TRY(hipEventCreate(&e0));
  VD_HIP_TRY(hipEventCreate(&e1));
  T* F = fld<T>(p, F_HA);
  T* out = fld<T>(p, F_HB);
  for (int st = 0; st < 4; ++st) {
    int M, N, K, nz;
    gemm_stage_shape(p, st, &M, &N, &K, &nz);
    float best = 1e30f;
    int best_cfg = p.split_ns > 0 ? split_default_cfg(p, st) : 0;
    const int ncfg = p.split_ns > 0 ? split_num_cfgs() : kNumTileCfg;
    // three timing rounds over the configurations, interleaved, keeping each
    // configuration's fastest round (single rounds of a few launches vary by
    // several percent from run to run)
    std::vector<float> tbest(ncfg, 1e30f);
    for (int round = 0; round < 3; ++round)
      for (int c = 0; c < ncfg; ++c) {
        if (p.split_ns > 0 ? !split_cfg_fits(p, st, c) : !tile_cfg_fits(c, M, N, K)) continue;
        launch_gemm_stage<T>(p, st, F, nullptr, nullptr, out, INV_FWD, c, s);  // warm
        VD_HIP_TRY(hipEventRecord(e0, s));
        for (int r = 0; r < reps; ++r) launch_gemm_stage<T>(p, st, F, nullptr, nullptr, out, INV_FWD, c, s);
        VD_HIP_TRY(hipEventRecord(e1, s));
        VD_HIP_TRY(hipEventSynchronize(e1));
        float ms = 0;
        VD_HIP_TRY(hipEventElapsedTime(&ms, e0, e1));
        tbest[c] = std::min(tbest[c], ms);
      }
    for (int c = 0; c < ncfg; ++c)
      if (tbest[c] < best) {
        best = tbest[c];
        best_cfg = c;
      }
    p.gemm_cfg[st] = best_cfg;
  }
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  VD_HIP_TRY(hipGetLastError());
  return VD_OK;
}

}  // namespace vd

using namespace vd;


extern "C" {

int vd_version(int* major, int* minor) {
  if (major) *major = VD_VERSION_MAJOR;
  if (minor) *minor = VD_VERSION_MINOR;
  return VD_OK;
}

const char* vd_last_error(void) { return g_last_error.c_str(); }

int vd_qg_plan_create(const vd_qg_params* prm, vd_plan** out) {
  VD_REQUIRE(prm && out, VD_ERR_INVALID_ARG, "null argument");
  VD_REQUIRE(prm->ny >= 8 && prm->nx >= 8, VD_ERR_SHAPE, "grid must be at least 8x8");
  VD_REQUIRE(prm->precision == VD_F32 || prm->precision == VD_F64, VD_ERR_INVALID_ARG,
             "precision must be VD_F32 or VD_F64");
  VD_REQUIRE(prm->time_scheme >= VD_EULER && prm->time_scheme <= VD_RK3, VD_ERR_INVALID_ARG,
             "time_scheme must be VD_EULER, VD_RK2 or VD_RK3");
  VD_REQUIRE(prm->dx > 0 && prm->dy > 0 && prm->f0 != 0 && prm->c > 0 && prm->g > 0,
             VD_ERR_INVALID_ARG, "dx, dy, c, g must be positive and f0 non-zero");
  VD_REQUIRE(prm->nbatch >= 0 && prm->nbatch <= 4096, VD_ERR_INVALID_ARG,
             "nbatch must be in 0..4096");
  VD_REQUIRE(prm->dst_split == 0 || (prm->dst_split >= 5 && prm->dst_split <= 7), VD_ERR_INVALID_ARG,
             "dst_split must be 0, 5, 6 or 7");
  VD_REQUIRE(prm->dst_split == 0 || prm->precision == VD_F64, VD_ERR_INVALID_ARG,
             "dst_split needs an FP64 plan");
  VD_REQUIRE(prm->gemm_config >= 0 && prm->gemm_config <= (prm->dst_split ? split_num_cfgs() : kNumTileCfg),
             VD_ERR_INVALID_ARG,
             prm->dst_split ? "gemm_config must be in 0..4 with dst_split" : "gemm_config must be in 0..11");
  vd_plan* h = new vd_plan();
  Plan& p = h->p;
  p.ny = prm->ny; p.nx = prm->nx;
  p.my = p.ny - 2; p.mx = p.nx - 2;
  p.MY = round_up(p.my, kTile); p.MX = round_up(p.mx, kTile);
  p.fold = (!prm->dst_dense && p.my % 2 == 0 && p.mx % 2 == 0) ? 1 : 0;
  p.hy = p.my / 2; p.hx = p.mx / 2;
  const int forced_cfg = prm->gemm_config - 1;
  p.nb = prm->nbatch <= 0 ? 1 : prm->nbatch;
  p.HY = round_up(p.hy, kTile); p.HX = round_up(p.hx, kTile);
  p.split_ns = prm->dst_split;
  {
    const int ky = p.fold ? p.HY : p.MY, kx = p.fold ? p.HX : p.MX;
    if (p.split_ns > 0 && (ky > split_max_k() || kx > split_max_k())) {
      set_error("dst_split supports sine products with K up to " + std::to_string(split_max_k()) +
                " (interior sizes up to 1024 folded, 512 dense)");
      delete h;
      return VD_ERR_UNSUPPORTED;
    }
  }
  p.prec = prm->precision;
  p.scheme = prm->time_scheme;
  p.nstage = p.scheme == VD_EULER ? 1 : (p.scheme == VD_RK2 ? 2 : 3);
  p.dx = prm->dx; p.dy = prm->dy; p.dt = prm->dt;
  p.g = prm->g; p.f0 = prm->f0; p.c = prm->c;
  p.has_kdiff = prm->has_kdiffus; p.kdiffus = prm->kdiffus;
  p.has_beta = prm->beta != nullptr;
  p.has_bathy = prm->fbathy != nullptr;
  p.has_mdt = prm->mdt != nullptr;
  p.es = p.prec == VD_F64 ? 8 : 4;
  p.nf = (size_t)p.ny * p.nx;
  p.ng = (size_t)p.MY * p.MX;
  if (hipGetDevice(&p.device) != hipSuccess) {
    set_error("hipGetDevice failed");
    delete h;
    return VD_ERR_HIP;
  }
  for (int st = 0; st < 4; ++st)
    p.gemm_cfg[st] = forced_cfg >= 0 ? forced_cfg
                                     : (p.split_ns > 0 ? split_default_cfg(p, st) : default_tile_cfg(p, st));
  int rc = p.prec == VD_F64 ? build_constants<double>(p, *prm) : build_constants<float>(p, *prm);
  if (rc != VD_OK) {
    const std::string msg = g_last_error;
    release_plan_buffers(p);
    delete h;
    set_error(msg);
    return rc;
  }
  *out = h;
  return VD_OK;
}

int vd_qg_plan_destroy(vd_plan* h) {
  if (!h) return VD_OK;
  release_plan_buffers(h->p);
  delete h;
  return VD_OK;
}

int vd_qg_tape_bytes(const vd_plan* h, int nstep, size_t* bytes) {
  VD_REQUIRE(h && bytes, VD_ERR_INVALID_ARG, "null argument");
  VD_REQUIRE(nstep >= 0, VD_ERR_INVALID_ARG, "nstep must be >= 0");
  const Plan& p = h->p;
  *bytes = ((size_t)nstep * p.nstage + 1) * 2 * p.nb * p.nf * p.es;
  return VD_OK;
}

int vd_qg_plan_bytes(const vd_plan* h, size_t* bytes) {
  VD_REQUIRE(h && bytes, VD_ERR_INVALID_ARG, "null argument");
  const Plan& p = h->p;
  size_t b = ((size_t)p.MY * p.MY + (size_t)p.MX * p.MX + p.nb * p.ng * 3) * p.es + p.nf +
             ((size_t)p.MY + p.MX + p.nb) * sizeof(double) +
             (size_t)F_COUNT * p.nb * p.nf * p.es;
  if (p.split_ns > 0) {
    const int npar = p.fold ? 2 : 1;
    const size_t ly = p.fold ? p.HY : p.MY, lx = p.fold ? p.HX : p.MX;
    const size_t ay = split_line_alloc((int)ly), ax = split_line_alloc((int)lx);
    b += (size_t)npar * (p.split_ns * (ay * ly + ax * lx) + (ay + ax) * sizeof(int));
  }
  if (p.has_beta) b += p.nf * p.es;
  if (p.has_bathy) b += p.nf * p.es;
  if (p.has_mdt) b += p.nf * p.es;
  *bytes = b;
  return VD_OK;
}

int vd_qg_autotune(vd_plan* h, int reps, int* chosen, void* stream) {
  VD_REQUIRE(h, VD_ERR_INVALID_ARG, "null argument");
  VD_REQUIRE(reps >= 1 && reps <= 1000, VD_ERR_INVALID_ARG, "reps must be in 1..1000");
  Plan& p = h->p;
  hipStream_t s = static_cast<hipStream_t>(stream);
  const int rc = p.prec == VD_F64 ? autotune<double>(p, reps, s) : autotune<float>(p, reps, s);
  if (rc != VD_OK) return rc;
  if (chosen)
    for (int st = 0; st < 4; ++st) chosen[st] = p.gemm_cfg[st];
  return VD_OK;
}

int vd_qg_debug_scratch(const vd_plan* h, int which, void** ptr, int* rows, int* cols) {
  VD_REQUIRE(h && ptr && rows && cols, VD_ERR_INVALID_ARG, "null argument");
  VD_REQUIRE(which == 0 || which == 1, VD_ERR_INVALID_ARG, "which must be 0 or 1");
  *ptr = h->p.gT[which];
  *rows = h->p.MY;
  *cols = h->p.MX;
  return VD_OK;
}

int vd_qg_step(vd_plan* h, const void* h0, const void* hb, void* h1, int nstep, void* tape,
               void* stream) {
  VD_REQUIRE(h && h0 && h1, VD_ERR_INVALID_ARG, "null argument");
  VD_REQUIRE(nstep >= 0, VD_ERR_INVALID_ARG, "nstep must be >= 0");
  Plan& p = h->p;
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (p.prec == VD_F64)
    return qg_forward<double>(p, (const double*)h0, (const double*)hb, (double*)h1, nstep,
                              (double*)tape, s);
  return qg_forward<float>(p, (const float*)h0, (const float*)hb, (float*)h1, nstep, (float*)tape,
                           s);
}

int vd_qg_step_tgl(vd_plan* h, const void* dh0, const void* h0, const void* hb, void* dh1,
                   void* h1, int nstep, void* stream) {
  VD_REQUIRE(h && dh0 && h0 && dh1, VD_ERR_INVALID_ARG, "null argument");
  VD_REQUIRE(nstep >= 0, VD_ERR_INVALID_ARG, "nstep must be >= 0");
  Plan& p = h->p;
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (p.prec == VD_F64)
    return qg_tangent<double>(p, (const double*)dh0, (const double*)h0, (const double*)hb,
                              (double*)dh1, (double*)h1, nstep, s);
  return qg_tangent<float>(p, (const float*)dh0, (const float*)h0, (const float*)hb, (float*)dh1,
                           (float*)h1, nstep, s);
}

int vd_qg_step_adj(vd_plan* h, const void* adh1, const void* h0, const void* hb, void* adh0,
                   int nstep, void* tape, int tape_valid, void* stream) {
  VD_REQUIRE(h && adh1 && h0 && adh0 && tape, VD_ERR_INVALID_ARG, "null argument");
  VD_REQUIRE(nstep >= 0, VD_ERR_INVALID_ARG, "nstep must be >= 0");
  Plan& p = h->p;
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (p.prec == VD_F64)
    return qg_adjoint<double>(p, (const double*)adh1, (const double*)h0, (const double*)hb,
                              (double*)adh0, nstep, (double*)tape, tape_valid, s);
  return qg_adjoint<float>(p, (const float*)adh1, (const float*)h0, (const float*)hb,
                           (float*)adh0, nstep, (float*)tape, tape_valid, s);
}

int vd_qg_pv2h(vd_plan* h, const void* q, const void* hb, const void* qb, void* hout,
               void* stream) {
  VD_REQUIRE(h && q && hout, VD_ERR_INVALID_ARG, "null argument");
  Plan& p = h->p;
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (p.prec == VD_F64)
    launch_inversion<double>(p, (const double*)q, (const double*)qb, (const double*)hb,
                             (double*)hout, INV_FWD, s);
  else
    launch_inversion<float>(p, (const float*)q, (const float*)qb, (const float*)hb, (float*)hout,
                            INV_FWD, s);
  VD_HIP_TRY(hipGetLastError());
  return VD_OK;
}

int vd_qg_h2pv(vd_plan* h, const void* hh, const void* hb, void* q, void* stream) {
  VD_REQUIRE(h && hh && hb && q, VD_ERR_INVALID_ARG, "null argument");
  Plan& p = h->p;
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (p.prec == VD_F64)
    launch_h2pv<double>(p, (const double*)hh, (const double*)hb, (double*)q, s);
  else
    launch_h2pv<float>(p, (const float*)hh, (const float*)hb, (float*)q, s);
  VD_HIP_TRY(hipGetLastError());
  return VD_OK;
}

int vd_qg_rhs(vd_plan* h, const void* hh, const void* q, const void* base, double coef, void* out,
              void* stream) {
  VD_REQUIRE(h && hh && q && out, VD_ERR_INVALID_ARG, "null argument");
  VD_REQUIRE(out != hh && out != q, VD_ERR_INVALID_ARG,
             "out must not alias h or q (the stencil reads them at neighbouring points)");
  Plan& p = h->p;
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (p.prec == VD_F64) {
    const double* b = base ? (const double*)base : fld<double>(p, F_ZERO);
    launch_stage<double>(p, (const double*)hh, (const double*)q, b, nullptr, nullptr, 0.0, 0.0,
                         coef, (double*)out, nullptr, s);
  } else {
    const float* b = base ? (const float*)base : fld<float>(p, F_ZERO);
    launch_stage<float>(p, (const float*)hh, (const float*)q, b, nullptr, nullptr, 0.f, 0.f,
                        (float)coef, (float*)out, nullptr, s);
  }
  VD_HIP_TRY(hipGetLastError());
  return VD_OK;
}

int vd_qg_rhs_adj(vd_plan* h, const void* hh, const void* q, const void* lam, double coef,
                  const void* acc, void* lq, void* lh, void* stream) {
  VD_REQUIRE(h && hh && q && lam && lq && lh, VD_ERR_INVALID_ARG, "null argument");
  VD_REQUIRE(lq != lam && lh != lam && lq != hh && lq != q && lh != hh && lh != q && lq != lh &&
                 acc != lh,
             VD_ERR_INVALID_ARG,
             "lq and lh must not alias lam, h, q or each other (the stencil reads lam, h and q at "
             "neighbouring points)");
  Plan& p = h->p;
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (p.prec == VD_F64) {
    launch_stage_adj<double>(p, (const double*)hh, (const double*)q, (const double*)lam, nullptr,
                             nullptr, coef, 0.0, 0.0, (const double*)acc, nullptr, nullptr,
                             (double*)lq, (double*)lh, s);
  } else {
    launch_stage_adj<float>(p, (const float*)hh, (const float*)q, (const float*)lam, nullptr,
                            nullptr, (float)coef, 0.f, 0.f, (const float*)acc, nullptr, nullptr,
                            (float*)lq, (float*)lh, s);
  }
  VD_HIP_TRY(hipGetLastError());
  return VD_OK;
}

int vd_qg_prep(vd_plan* h, const void* h0, const void* hb, void* hh, void* q, void* stream) {
  VD_REQUIRE(h && h0 && hh && q, VD_ERR_INVALID_ARG, "null argument");
  VD_REQUIRE(hh != h0 && q != h0, VD_ERR_INVALID_ARG,
             "h and q must not alias h0 (q = h2pv(h0) reads h0 at neighbouring points)");
  Plan& p = h->p;
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (p.prec == VD_F64)
    launch_prep<double>(p, (const double*)h0, hb ? (const double*)hb : fld<double>(p, F_ZERO),
                        (double*)hh, fld<double>(p, F_HBM), (double*)q, fld<double>(p, F_QB), s);
  else
    launch_prep<float>(p, (const float*)h0, hb ? (const float*)hb : fld<float>(p, F_ZERO),
                       (float*)hh, fld<float>(p, F_HBM), (float*)q, fld<float>(p, F_QB), s);
  VD_HIP_TRY(hipGetLastError());
  return VD_OK;
}

int vd_qg_finish(vd_plan* h, const void* hh, void* h1, void* stream) {
  VD_REQUIRE(h && hh && h1, VD_ERR_INVALID_ARG, "null argument");
  Plan& p = h->p;
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (p.prec == VD_F64)
    launch_finish<double>(p, (const double*)hh, (double*)h1, s);
  else
    launch_finish<float>(p, (const float*)hh, (float*)h1, s);
  VD_HIP_TRY(hipGetLastError());
  return VD_OK;
}

int vd_qg_adj_init(vd_plan* h, const void* adh1, void* lh, void* lq, void* stream) {
  VD_REQUIRE(h && adh1 && lh && lq, VD_ERR_INVALID_ARG, "null argument");
  Plan& p = h->p;
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (p.prec == VD_F64)
    launch_adj_init<double>(p, (const double*)adh1, (double*)lh, (double*)lq, s);
  else
    launch_adj_init<float>(p, (const float*)adh1, (float*)lh, (float*)lq, s);
  VD_HIP_TRY(hipGetLastError());
  return VD_OK;
}

int vd_qg_adj_final(vd_plan* h, const void* hcur, const void* lh, const void* lq, void* adh0,
                    void* stream) {
  VD_REQUIRE(h && hcur && lh && lq && adh0, VD_ERR_INVALID_ARG, "null argument");
  VD_REQUIRE(adh0 != lq, VD_ERR_INVALID_ARG,
             "adh0 must not alias lq (it is read at neighbouring points)");
  Plan& p = h->p;
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (p.prec == VD_F64)
    launch_adj_final<double>(p, (const double*)hcur, (const double*)lh, (const double*)lq,
                             (double*)adh0, s);
  else
    launch_adj_final<float>(p, (const float*)hcur, (const float*)lh, (const float*)lq,
                            (float*)adh0, s);
  VD_HIP_TRY(hipGetLastError());
  return VD_OK;
}

int vd_qg_fill_land(vd_plan* h, void* field, double value, void* stream) {
  VD_REQUIRE(h && field, VD_ERR_INVALID_ARG, "null argument");
  Plan& p = h->p;
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (p.prec == VD_F64)
    launch_fill_land<double>(p, (double*)field, value, s);
  else
    launch_fill_land<float>(p, (float*)field, (float)value, s);
  VD_HIP_TRY(hipGetLastError());
  return VD_OK;
}

}  // extern "C"
