// Multilayer rotating shallow-water model on the Arakawa C grid
// (mapping/models/model_qgsw/sw.py:67-937): forward RK3-SSP steps, their
// tangent-linear model and a hand-written adjoint.
//
// One RHS evaluation (compute_time_derivatives, sw.py:723-750) is two
// kernels:
//   k_sw_diag  per (i, j, layer): the vorticity on the w grid, the clamped
//              total thickness h_tot and kinetic energy + pressure (the
//              pressure couples the layers: a thread reads the column's
//              thicknesses)
//   k_sw_tend  per (i, j, layer): the thickness tendency (the divergence of
//              the two upwind-reconstructed fluxes of h_tot), the u and v
//              tendencies (vortex force from the reconstructed vorticity,
//              Coriolis, gradient of kinetic energy + pressure, wind, bottom
//              drag, viscosity; thickness diffusion)
//              and the RK3-SSP stage update written to the next state buffer
// Both are templates over the scalar type: double for the model, a dual
// number (value, tangent) for the tangent-linear model, whose value part is
// the model's arithmetic.
//
// The adjoint of one RHS evaluation, F'(x)^T mu, is three gather kernels, so
// every sum has a fixed order (no atomics) and the result is deterministic:
//   k_sw_adj_face     per face and velocity point: the flux cotangent times
//                     the partial derivatives of the flux with respect to its
//                     six stencil values (a reverse sweep through the
//                     reconstruction) and its velocity
//   k_sw_adj_gather1  per w point and layer: the vorticity cotangent; per h
//                     cell and layer: the h_tot, kinetic-energy and pressure
//                     cotangents (the pressure's layer sums in reverse)
//   k_sw_adj_gather2  per u / v point and h cell: the state cotangents
// and the RK3 stages run in reverse with the stage states recomputed from the
// step's start state.  Steps are reversed segment by segment: the start
// states of segments of about sqrt(n) steps are kept from a first forward
// sweep, each segment's step states are recomputed before it is reversed.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "vd_internal.h"
#include "vd_sw_core.h"

namespace vd {
namespace sw {

typedef Dn<1> D1;

struct St {
  Fld u, v, h;
};
struct StOut {
  FldOut u, v, h;
};

// ------------------------------------------------------------ forward pieces
// h flux along x through the face at u point (iu, j), 1 <= iu <= nx-1:
// between cells iu-1 and iu, stencil cells iu-3 .. iu+2, velocity U = u/dx^2.
template <typename S>
__device__ __forceinline__ S h_flux_x(const SwGrid& g, Fld ht, Fld u, int l, int iu, int j) {
  const int a = iu - 1, n = g.nx;
  S s[6];
#pragma unroll
  for (int k = 0; k < 6; ++k) {
    const int c = a - 2 + k;
    s[k] = (c >= 0 && c < n) ? Ld<S>::at(ht.v, ht.t, ih_(g, l, c, j)) : Cst<S>::make(0.0);
  }
  const S vel = Ld<S>::at(u.v, u.t, iu_(g, l, iu, j)) / (g.dx * g.dx);
  return vel * face_recon(s, val(vel) > 0, g.code_hx[iu * g.ny + j], a >= 1 && a + 2 <= n - 1,
                          a >= 2 && a + 3 <= n - 1);
}
// h flux along y through the face at v point (i, jv), 1 <= jv <= ny-1.
template <typename S>
__device__ __forceinline__ S h_flux_y(const SwGrid& g, Fld ht, Fld v, int l, int i, int jv) {
  const int a = jv - 1, n = g.ny;
  S s[6];
#pragma unroll
  for (int k = 0; k < 6; ++k) {
    const int c = a - 2 + k;
    s[k] = (c >= 0 && c < n) ? Ld<S>::at(ht.v, ht.t, ih_(g, l, i, c)) : Cst<S>::make(0.0);
  }
  const S vel = Ld<S>::at(v.v, v.t, iv_(g, l, i, jv)) / (g.dy * g.dy);
  return vel * face_recon(s, val(vel) > 0, g.code_hy[i * (g.ny + 1) + jv], a >= 1 && a + 2 <= n - 1,
                          a >= 2 && a + 3 <= n - 1);
}
// V interpolated to u point (iu, j) and U to v point (i, jv) (interp_TP).
template <typename S>
__device__ __forceinline__ S vm_at(const SwGrid& g, Fld v, int l, int iu, int j) {
  auto V = [&](int a, int b) { return Ld<S>::at(v.v, v.t, iv_(g, l, a, b)) / (g.dy * g.dy); };
  return 0.25 * (((V(iu, j + 1) + V(iu, j)) + V(iu - 1, j + 1)) + V(iu - 1, j));
}
template <typename S>
__device__ __forceinline__ S um_at(const SwGrid& g, Fld u, int l, int i, int jv) {
  auto U = [&](int a, int b) { return Ld<S>::at(u.v, u.t, iu_(g, l, a, b)) / (g.dx * g.dx); };
  return 0.25 * (((U(i + 1, jv) + U(i + 1, jv - 1)) + U(i, jv)) + U(i, jv - 1));
}
// Vorticity stencils: along y at u point (iu, j) (w points j-2 .. j+3 of
// row iu), along x at v point (i, jv) (w points i-2 .. i+3 of column jv).
template <typename S>
__device__ __forceinline__ void w_sten_y(const SwGrid& g, Fld om, int l, int iu, int j, S* s) {
#pragma unroll
  for (int k = 0; k < 6; ++k) {
    const int c = j - 2 + k;
    s[k] = (c >= 0 && c <= g.ny) ? Ld<S>::at(om.v, om.t, iw_(g, l, iu, c)) : Cst<S>::make(0.0);
  }
}
template <typename S>
__device__ __forceinline__ void w_sten_x(const SwGrid& g, Fld om, int l, int i, int jv, S* s) {
#pragma unroll
  for (int k = 0; k < 6; ++k) {
    const int c = i - 2 + k;
    s[k] = (c >= 0 && c <= g.nx) ? Ld<S>::at(om.v, om.t, iw_(g, l, c, jv)) : Cst<S>::make(0.0);
  }
}

// Diagnostics of layer blockIdx.z at (i, j): vorticity on the w grid, h_tot
// and kinetic energy + pressure on the h grid.  The pressure of layer l
// needs the thicknesses of every layer at the column; a thread reads them
// all (the same sums, in the same order, as one column walk).
template <typename S>
__global__ __launch_bounds__(256) void k_sw_diag(SwGrid g, St x, FldOut om, FldOut ht, FldOut kep) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x, i = blockIdx.y * blockDim.y + threadIdx.y;
  const int l = blockIdx.z;
  if (i > g.nx || j > g.ny) return;
  Ld<S>::put(om.v, om.t, iw_(g, l, i, j), omega_at<S>(g, x.u, x.v, l, i, j));
  if (i >= g.nx || j >= g.ny) return;
  // eta = reverse cumsum of h / area over the layers, p = cumsum of g' eta
  // (sw.py:686-687, reverse_cumsum :46-54)
  S tot = Cst<S>::make(0.0);
  for (int k = 0; k < g.nl; ++k) {
    const S xk = Ld<S>::at(x.h.v, x.h.t, ih_(g, k, i, j)) / g.area;
    tot = k == 0 ? xk : tot + xk;
  }
  S cum = Cst<S>::make(0.0), p = Cst<S>::make(0.0), xl = Cst<S>::make(0.0);
  for (int k = 0; k <= l; ++k) {
    const S xk = Ld<S>::at(x.h.v, x.h.t, ih_(g, k, i, j)) / g.area;
    cum = k == 0 ? xk : cum + xk;
    const S eta = (xk + tot) - cum;
    p = k == 0 ? g.gp[0] * eta : p + g.gp[k] * eta;
    xl = xk;
  }
  (void)xl;
  const double mh = g.mask_h[i * g.ny + j];
  const double idx2 = g.dx * g.dx, idy2 = g.dy * g.dy;
  // comp_ke (finite_diff.py:26-30): 0.25 (u U + u U + v V + v V)
  const S ua = Ld<S>::at(x.u.v, x.u.t, iu_(g, l, i + 1, j)), ub = Ld<S>::at(x.u.v, x.u.t, iu_(g, l, i, j));
  const S va = Ld<S>::at(x.v.v, x.v.t, iv_(g, l, i, j + 1)), vb = Ld<S>::at(x.v.v, x.v.t, iv_(g, l, i, j));
  const S ke = 0.25 * (((ua * (ua / idx2) + ub * (ub / idx2)) + va * (va / idy2)) + vb * (vb / idy2));
  Ld<S>::put(kep.v, kep.t, ih_(g, l, i, j), ke * mh + p);
  const S hh = Ld<S>::at(x.h.v, x.h.t, ih_(g, l, i, j));
  Ld<S>::put(ht.v, ht.t, ih_(g, l, i, j), smooth_clamp(g.href[l] + hh, g.h_min_area, g.sharp));
}

// Five-point Laplacians of the edge-padded physical fields (add_diffusion,
// sw.py:511-546; add_h_diffusion, :548-565): u / dx at an interior u point,
// padded along y (the boundary u rows are its x neighbours); v / dy at an
// interior v point, padded along x; h / area at any cell, padded both ways.
template <typename S>
__device__ __forceinline__ S lap5(const S& c, const S& xm, const S& xp, const S& ym, const S& yp, const SwGrid& g) {
  return ((xp - 2.0 * c) + xm) / (g.dx * g.dx) + ((yp - 2.0 * c) + ym) / (g.dy * g.dy);
}
template <typename S>
__device__ __forceinline__ S lap_u(const SwGrid& g, Fld u, int l, int i, int j) {
  auto q = [&](int a, int b) { return Ld<S>::at(u.v, u.t, iu_(g, l, a, min(max(b, 0), g.ny - 1))) / g.dx; };
  return lap5<S>(q(i, j), q(i - 1, j), q(i + 1, j), q(i, j - 1), q(i, j + 1), g);
}
template <typename S>
__device__ __forceinline__ S lap_v(const SwGrid& g, Fld v, int l, int i, int j) {
  auto q = [&](int a, int b) { return Ld<S>::at(v.v, v.t, iv_(g, l, min(max(a, 0), g.nx - 1), b)) / g.dy; };
  return lap5<S>(q(i, j), q(i - 1, j), q(i + 1, j), q(i, j - 1), q(i, j + 1), g);
}
template <typename S>
__device__ __forceinline__ S lap_h(const SwGrid& g, Fld h, int l, int i, int j) {
  auto q = [&](int a, int b) {
    return Ld<S>::at(h.v, h.t, ih_(g, l, min(max(a, 0), g.nx - 1), min(max(b, 0), g.ny - 1))) / g.area;
  };
  return lap5<S>(q(i, j), q(i - 1, j), q(i + 1, j), q(i, j - 1), q(i, j + 1), g);
}

// RK3-SSP stage STAGE at state x (sw.py:834-857): k = F(x), then
//   0: x' = x + dt k            (k stored as k0)
//   1: x' = x + dt/4 (k - 3 k0)  (k stored as k1)
//   2: x' = x + dt/12 (8 k - k1 - k0)
// The thickness fluxes of the block's faces are computed once each into LDS
// (the x faces i0 .. i0+4 of its 64 columns, the y faces j0 .. j0+64 of its 4
// rows) and differenced from there: 10 flux evaluations per thread group of
// 64 instead of 16 when each cell evaluated its four faces.
constexpr int kSwBX = 64, kSwBY = 4;
template <typename S, int STAGE>
__global__ __launch_bounds__(256) void k_sw_tend(SwGrid g, St x, Fld om, Fld ht, Fld kep, St k0, St k1,
                                                 StOut kst, StOut xn) {
  const int tx = threadIdx.x, ty = threadIdx.y;
  const int j0 = blockIdx.x * kSwBX, i0 = blockIdx.y * kSwBY;
  const int j = j0 + tx, i = i0 + ty;
  const int l = blockIdx.z;
  const S zero = Cst<S>::make(0.0);
  __shared__ S fxs[kSwBY + 1][kSwBX];  // x face (u point) i0 + r at column j0 + c
  __shared__ S fys[kSwBY][kSwBX + 1];  // y face (v point) j0 + c at row i0 + r
  auto fx = [&](int iu, int jj) { return (iu >= 1 && iu <= g.nx - 1 && jj < g.ny) ? h_flux_x<S>(g, ht, x.u, l, iu, jj) : zero; };
  auto fy = [&](int ii, int jv) { return (ii < g.nx && jv >= 1 && jv <= g.ny - 1) ? h_flux_y<S>(g, ht, x.v, l, ii, jv) : zero; };
  fxs[ty][tx] = fx(i, j);
  fys[ty][tx] = fy(i, j);
  if (ty == 0) fxs[kSwBY][tx] = fx(i0 + kSwBY, j);
  if (ty == 1 && tx < kSwBY) fys[tx][kSwBX] = fy(i0 + tx, j0 + kSwBX);
  __syncthreads();
  if (i > g.nx || j > g.ny) return;
  const double dt = g.dt;
  auto update = [&](const Fld& xf, const Fld& k0f, const Fld& k1f, const FldOut& kf, const FldOut& out, int idx,
                    const S& k) {
    const S xv = Ld<S>::at(xf.v, xf.t, idx);
    S r;
    if (STAGE == 0) {
      Ld<S>::put(kf.v, kf.t, idx, k);
      r = xv + dt * k;
    } else if (STAGE == 1) {
      Ld<S>::put(kf.v, kf.t, idx, k);
      r = xv + (dt / 4.0) * (k - 3.0 * Ld<S>::at(k0f.v, k0f.t, idx));
    } else {
      r = xv + (dt / 12.0) * ((8.0 * k - Ld<S>::at(k1f.v, k1f.t, idx)) - Ld<S>::at(k0f.v, k0f.t, idx));
    }
    Ld<S>::put(out.v, out.t, idx, r);
  };
  // thickness (advection_h, sw.py:465-479; div_nofluxbc, finite_diff.py:42-52)
  if (i < g.nx && j < g.ny) {
    S divx = zero, divy = zero;
    if (i >= 1 && i <= g.nx - 2) divx = fxs[ty + 1][tx] - fxs[ty][tx];
    if (j >= 1 && j <= g.ny - 2) divy = fys[ty][tx + 1] - fys[ty][tx];
    S k = -(divy + divx) * g.mask_h[i * g.ny + j];
    if (g.diff > 0) k = k + ((g.diff * lap_h<S>(g, x.h, l, i, j)) * g.area) * g.mask_h[i * g.ny + j];
    update(x.h, k0.h, k1.h, kst.h, xn.h, ih_(g, l, i, j), k);
  }
  // u momentum (advection_momentum, sw.py:481-510)
  if (j < g.ny) {
    S k = zero;
    if (i >= 1 && i <= g.nx - 1) {
      const S vm = vm_at<S>(g, x.v, l, i, j);
      S s[6];
      w_sten_y<S>(g, om, l, i, j, s);
      const S fw = vm * face_recon(s, val(vm) > 0, g.code_wy[i * g.ny + j], j >= 1 && j + 2 <= g.ny,
                                   j >= 2 && j + 3 <= g.ny);
      k = fw + g.fstar_u[i * g.ny + j] * vm;
      k = k - ((Ld<S>::at(kep.v, kep.t, ih_(g, l, i, j)) - Ld<S>::at(kep.v, kep.t, ih_(g, l, i - 1, j))) + 0.0);
      // wind forcing of the top layer, then bottom drag (sw.py:504-505)
      if (l == 0 && g.wind_u) k = k + Cst<S>::make(g.wind_u[i * g.ny + j]);
      if (l == g.nl - 1) k = k - g.drag * Ld<S>::at(x.u.v, x.u.t, iu_(g, l, i, j));
      if (g.visc > 0) k = k + (g.visc * lap_u<S>(g, x.u, l, i, j)) * g.dx;
      k = k * g.mask_u[i * g.ny + j];
    }
    update(x.u, k0.u, k1.u, kst.u, xn.u, iu_(g, l, i, j), k);
  }
  // v momentum
  if (i < g.nx) {
    S k = zero;
    if (j >= 1 && j <= g.ny - 1) {
      const S um = um_at<S>(g, x.u, l, i, j);
      S s[6];
      w_sten_x<S>(g, om, l, i, j, s);
      const S fw = um * face_recon(s, val(um) > 0, g.code_wx[i * (g.ny + 1) + j], i >= 1 && i + 2 <= g.nx,
                                   i >= 2 && i + 3 <= g.nx);
      k = -(fw + g.fstar_v[i * (g.ny + 1) + j] * um);
      k = k - ((Ld<S>::at(kep.v, kep.t, ih_(g, l, i, j)) - Ld<S>::at(kep.v, kep.t, ih_(g, l, i, j - 1))) + 0.0);
      if (l == 0 && g.wind_v) k = k + Cst<S>::make(g.wind_v[i * (g.ny + 1) + j]);
      if (l == g.nl - 1) k = k - g.drag * Ld<S>::at(x.v.v, x.v.t, iv_(g, l, i, j));
      if (g.visc > 0) k = k + (g.visc * lap_v<S>(g, x.v, l, i, j)) * g.dy;
      k = k * g.mask_v[i * (g.ny + 1) + j];
    }
    update(x.v, k0.v, k1.v, kst.v, xn.v, iv_(g, l, i, j), k);
  }
}

// set_input_uvh / get_physical_uvh (sw.py:423-449) and their adjoints on
// one field of a state (the caller passes the field's metric factor d and
// its mask, `plane` values per layer): mode 0 out = (in mask) d,
// 1 out = in / d, 2 out = (in d) mask.  (The factor and mask are arguments:
// choosing them from a field index inside the kernel was compiled wrongly
// by this ROCm's compiler, the third branch taking the second one's factor.)
__global__ void k_sw_scale_field(const double* in, double* out, const double* mask, int plane, double d, int mode,
                                 int n) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  const double m = mask[e % plane];
  const double x = in ? in[e] : 0.0;  // no input: a zero field
  out[e] = mode == 0 ? (x * m) * d : (mode == 1 ? x / d : (x * d) * m);
}

// ------------------------------------------------------------ adjoint
// Workspace planes (one (nx+1) x (ny+1) plane per layer each): the face
// partial products G (24 planes: h flux along x and y, vorticity flux along
// y on u points and along x on v points, 6 stencil values each), the
// velocity cotangents of the h fluxes (aUf, aVf), of the interpolated
// velocities (aVm on u points, aUm on v points), the vorticity cotangent,
// the kinetic-energy cotangent and the thickness cotangent of gather1.
enum { kGx = 0, kGy = 6, kGwy = 12, kGwx = 18, kAUf = 24, kAVf, kAVm, kAUm, kAom, kAKE, kAhp, kAdjPlanes };

struct AdjWs {
  double* p;  // kAdjPlanes x nl planes
  int nw;     // (nx+1) (ny+1)
  __device__ __forceinline__ double* pl(int q, int l, int nl) const { return p + ((size_t)q * nl + l) * nw; }
};

// Partials of vel * recon(s) with respect to s[0..5] (times vel) and vel,
// by the reverse sweep of the reconstruction.
__device__ __forceinline__ void face_partials(const double* s, double vel, uint8_t code, bool ok4, bool ok6,
                                              double* ds, double* dvel) {
  double g[6];
  *dvel = face_recon_grad(s, vel > 0, code, ok4, ok6, g);
#pragma unroll
  for (int k = 0; k < 6; ++k) ds[k] = vel * g[k];
}

__device__ __forceinline__ double mu_h(const SwGrid& g, const double* lh, int l, int i, int j) {
  return (i >= 0 && i < g.nx && j >= 0 && j < g.ny) ? lh[ih_(g, l, i, j)] * g.mask_h[i * g.ny + j] : 0.0;
}
__device__ __forceinline__ double mu_u(const SwGrid& g, const double* lu, int l, int i, int j) {
  return (i >= 1 && i <= g.nx - 1 && j >= 0 && j < g.ny) ? lu[iu_(g, l, i, j)] * g.mask_u[i * g.ny + j] : 0.0;
}
__device__ __forceinline__ double mu_v(const SwGrid& g, const double* lv, int l, int i, int j) {
  return (i >= 0 && i < g.nx && j >= 1 && j <= g.ny - 1) ? lv[iv_(g, l, i, j)] * g.mask_v[i * (g.ny + 1) + j]
                                                         : 0.0;
}

// mu = (lu, lv, lh) is the cotangent of F(x) = (dt_u, dt_v, dt_h); x the
// state, om / ht the diagnostics of x.
__global__ __launch_bounds__(256) void k_sw_adj_face(SwGrid g, const double* xu, const double* xv,
                                                     const double* om, const double* ht, const double* lu,
                                                     const double* lv, const double* lh, AdjWs ws) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x, i = blockIdx.y * blockDim.y + threadIdx.y;
  const int l = blockIdx.z, nl = g.nl;
  if (i > g.nx || j > g.ny) return;
  const int w = i * (g.ny + 1) + j;
  double ds[6], dvel;
  // h flux along x at u point (i, j)
  {
    double G[6] = {0, 0, 0, 0, 0, 0}, aU = 0.0;
    if (i >= 1 && i <= g.nx - 1 && j < g.ny) {
      const double muF = (i <= g.nx - 2 ? mu_h(g, lh, l, i, j) : 0.0) - (i >= 2 ? mu_h(g, lh, l, i - 1, j) : 0.0);
      const int a = i - 1, n = g.nx;
      double s[6];
      for (int k = 0; k < 6; ++k) {
        const int c = a - 2 + k;
        s[k] = (c >= 0 && c < n) ? ht[ih_(g, l, c, j)] : 0.0;
      }
      const double vel = xu[iu_(g, l, i, j)] / (g.dx * g.dx);
      face_partials(s, vel, g.code_hx[i * g.ny + j], a >= 1 && a + 2 <= n - 1, a >= 2 && a + 3 <= n - 1, ds, &dvel);
      for (int k = 0; k < 6; ++k) G[k] = muF * ds[k];
      aU = muF * dvel;
    }
    for (int k = 0; k < 6; ++k) ws.pl(kGx + k, l, nl)[w] = G[k];
    ws.pl(kAUf, l, nl)[w] = aU;
  }
  // h flux along y at v point (i, j)
  {
    double G[6] = {0, 0, 0, 0, 0, 0}, aV = 0.0;
    if (i < g.nx && j >= 1 && j <= g.ny - 1) {
      const double muF = (j <= g.ny - 2 ? mu_h(g, lh, l, i, j) : 0.0) - (j >= 2 ? mu_h(g, lh, l, i, j - 1) : 0.0);
      const int a = j - 1, n = g.ny;
      double s[6];
      for (int k = 0; k < 6; ++k) {
        const int c = a - 2 + k;
        s[k] = (c >= 0 && c < n) ? ht[ih_(g, l, i, c)] : 0.0;
      }
      const double vel = xv[iv_(g, l, i, j)] / (g.dy * g.dy);
      face_partials(s, vel, g.code_hy[i * (g.ny + 1) + j], a >= 1 && a + 2 <= n - 1, a >= 2 && a + 3 <= n - 1, ds,
                    &dvel);
      for (int k = 0; k < 6; ++k) G[k] = muF * ds[k];
      aV = muF * dvel;
    }
    for (int k = 0; k < 6; ++k) ws.pl(kGy + k, l, nl)[w] = G[k];
    ws.pl(kAVf, l, nl)[w] = aV;
  }
  const Fld omf{om, nullptr}, uf{xu, nullptr}, vf{xv, nullptr};
  // u momentum at (i, j): vortex force along y, Coriolis
  {
    double G[6] = {0, 0, 0, 0, 0, 0}, aVm = 0.0;
    if (i >= 1 && i <= g.nx - 1 && j < g.ny) {
      const double mu = mu_u(g, lu, l, i, j);
      const double vm = vm_at<double>(g, vf, l, i, j);
      double s[6];
      w_sten_y<double>(g, omf, l, i, j, s);
      face_partials(s, vm, g.code_wy[i * g.ny + j], j >= 1 && j + 2 <= g.ny, j >= 2 && j + 3 <= g.ny, ds, &dvel);
      for (int k = 0; k < 6; ++k) G[k] = mu * ds[k];
      aVm = mu * (dvel + g.fstar_u[i * g.ny + j]);
    }
    for (int k = 0; k < 6; ++k) ws.pl(kGwy + k, l, nl)[w] = G[k];
    ws.pl(kAVm, l, nl)[w] = aVm;
  }
  // v momentum at (i, j)
  {
    double G[6] = {0, 0, 0, 0, 0, 0}, aUm = 0.0;
    if (i < g.nx && j >= 1 && j <= g.ny - 1) {
      const double mu = mu_v(g, lv, l, i, j);
      const double um = um_at<double>(g, uf, l, i, j);
      double s[6];
      w_sten_x<double>(g, omf, l, i, j, s);
      face_partials(s, um, g.code_wx[i * (g.ny + 1) + j], i >= 1 && i + 2 <= g.nx, i >= 2 && i + 3 <= g.nx, ds,
                    &dvel);
      for (int k = 0; k < 6; ++k) G[k] = -mu * ds[k];
      aUm = -mu * (dvel + g.fstar_v[i * (g.ny + 1) + j]);
    }
    for (int k = 0; k < 6; ++k) ws.pl(kGwx + k, l, nl)[w] = G[k];
    ws.pl(kAUm, l, nl)[w] = aUm;
  }
}

__global__ __launch_bounds__(256) void k_sw_adj_gather1(SwGrid g, const double* xh, const double* lu,
                                                        const double* lv, AdjWs ws) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x, i = blockIdx.y * blockDim.y + threadIdx.y;
  const int l = blockIdx.z, nl = g.nl, W = g.ny + 1;
  if (i > g.nx || j > g.ny) return;
  const int w = i * W + j;
  // vorticity cotangent at w point (i, j): the u points (i, j+2-k) of the flux
  // along y and the v points (i+2-k, j) of the flux along x
  {
    double a = 0.0;
    for (int k = 0; k < 6; ++k) {
      const int jj = j + 2 - k;
      if (jj >= 0 && jj <= g.ny) a += ws.pl(kGwy + k, l, nl)[i * W + jj];
    }
    for (int k = 0; k < 6; ++k) {
      const int ii = i + 2 - k;
      if (ii >= 0 && ii <= g.nx) a += ws.pl(kGwx + k, l, nl)[ii * W + j];
    }
    ws.pl(kAom, l, nl)[w] = a;
  }
  if (i >= g.nx || j >= g.ny) return;
  // h cell (i, j) of layer l: h_tot through the faces (i+3-k, j) and
  // (i, j+3-k), kinetic energy + pressure through the u points i, i+1 and
  // v points j, j+1 of every layer (the pressure of layer k depends on the
  // thickness of layer l for all k)
  double a = 0.0;
  for (int k = 0; k < 6; ++k) {
    const int ii = i + 3 - k;
    if (ii >= 0 && ii <= g.nx) a += ws.pl(kGx + k, l, nl)[ii * W + j];
  }
  for (int k = 0; k < 6; ++k) {
    const int jj = j + 3 - k;
    if (jj >= 0 && jj <= g.ny) a += ws.pl(kGy + k, l, nl)[i * W + jj];
  }
  const double mh = g.mask_h[i * g.ny + j];
  const double hh = xh[ih_(g, l, i, j)];
  const double ahp = a * smooth_clamp_deriv(g.href[l] + hh, g.h_min_area, g.sharp);
  auto akep = [&](int k) {
    return -mu_u(g, lu, k, i, j) + mu_u(g, lu, k, i + 1, j) - mu_v(g, lv, k, i, j) + mu_v(g, lv, k, i, j + 1);
  };
  ws.pl(kAKE, l, nl)[w] = akep(l) * mh;
  // p_k = sum_{m <= k} g'_m eta_m, eta_m = sum_{n >= m} h_n / area:
  // d/dh_l = (1/area) sum_{m <= l} g'_m sum_{k >= m} akep_k
  double ap[8];
  for (int k = 0; k < nl; ++k) ap[k] = akep(k);
  double rs = 0.0, cum = 0.0;
  double aeta[8];
  for (int k = nl - 1; k >= 0; --k) {
    rs += ap[k];
    aeta[k] = g.gp[k] * rs;
  }
  for (int m = 0; m <= l; ++m) cum += aeta[m];
  ws.pl(kAhp, l, nl)[w] = ahp + cum / g.area;
}

// Transposed edge-padded Laplacian at point (i, j): the cotangents m0 of the
// point and of its four neighbours (zero where the forward Laplacian is not
// evaluated), the padding along a direction folding the outside neighbour of
// the first / last point back onto it (fold_* = 1 there).
__device__ __forceinline__ double lap5_adj(double m0, double xm, double xp, double ym, double yp, int fold_x,
                                           int fold_y, const SwGrid& g) {
  return ((xm + xp) + (fold_x - 2.0) * m0) / (g.dx * g.dx) + ((ym + yp) + (fold_y - 2.0) * m0) / (g.dy * g.dy);
}

// The stage combination that follows an adjoint RHS evaluation in the
// reversed RK3 step, y_k = a_k y_k + b_k L (a_k = 0: y_k = b_k L) over the
// whole state, applied by k_sw_adj_gather2 to each value of L it finishes
// (the arrays y_k are not read by the evaluation itself).
struct StageComb {
  double* y[2];
  double a[2], b[2];
  __device__ __forceinline__ void apply(size_t e, double L) const {
#pragma unroll
    for (int k = 0; k < 2; ++k)
      if (y[k]) y[k][e] = a[k] == 0.0 ? b[k] * L : a[k] * y[k][e] + b[k] * L;
  }
};

// lam_out (u, v, h) += F'(x)^T mu, given gather1's planes; then the stage
// combination cb of the new values.
__global__ __launch_bounds__(256) void k_sw_adj_gather2(SwGrid g, const double* xu, const double* xv,
                                                        const double* lu, const double* lv, const double* lh,
                                                        AdjWs ws, double* ou, double* ov, double* oh,
                                                        StageComb cb) {
  const size_t nu = (size_t)g.nl * (g.nx + 1) * g.ny, nv = (size_t)g.nl * g.nx * (g.ny + 1);
  const int j = blockIdx.x * blockDim.x + threadIdx.x, i = blockIdx.y * blockDim.y + threadIdx.y;
  const int l = blockIdx.z, nl = g.nl, W = g.ny + 1;
  if (i > g.nx || j > g.ny) return;
  auto P = [&](int q, int a, int b) { return ws.pl(q, l, nl)[a * W + b]; };
  const double idx2 = 1.0 / (g.dx * g.dx), idy2 = 1.0 / (g.dy * g.dy);
  if (j < g.ny) {  // u point (i, j)
    double aU = P(kAUf, i, j) + 0.25 * ((i >= 1 ? P(kAUm, i - 1, j) + P(kAUm, i - 1, j + 1) : 0.0) +
                                        (i < g.nx ? P(kAUm, i, j) + P(kAUm, i, j + 1) : 0.0));
    const double u = xu[iu_(g, l, i, j)];
    const double ake = (i >= 1 ? P(kAKE, i - 1, j) : 0.0) + (i <= g.nx - 1 ? P(kAKE, i, j) : 0.0);
    double a = aU * idx2 + 0.5 * u * idx2 * ake;
    a += -g.wAC[i * W + j] * P(kAom, i, j) + g.wAC[i * W + j + 1] * P(kAom, i, j + 1);
    if (l == nl - 1) a += -g.drag * mu_u(g, lu, l, i, j);
    if (g.visc > 0)  // (nu lap(u / dx)) dx
      a += g.visc * lap5_adj(mu_u(g, lu, l, i, j), mu_u(g, lu, l, i - 1, j), mu_u(g, lu, l, i + 1, j),
                             mu_u(g, lu, l, i, j - 1), mu_u(g, lu, l, i, j + 1), 0,
                             (j == 0) + (j == g.ny - 1), g);
    const double L = ou[iu_(g, l, i, j)] + a;
    ou[iu_(g, l, i, j)] = L;
    cb.apply(iu_(g, l, i, j), L);
  }
  if (i < g.nx) {  // v point (i, j)
    double aV = P(kAVf, i, j) + 0.25 * ((j >= 1 ? P(kAVm, i, j - 1) + P(kAVm, i + 1, j - 1) : 0.0) +
                                        (j < g.ny ? P(kAVm, i, j) + P(kAVm, i + 1, j) : 0.0));
    const double v = xv[iv_(g, l, i, j)];
    const double ake = (j >= 1 ? P(kAKE, i, j - 1) : 0.0) + (j <= g.ny - 1 ? P(kAKE, i, j) : 0.0);
    double a = aV * idy2 + 0.5 * v * idy2 * ake;
    a += g.wAB[i * W + j] * P(kAom, i, j) - g.wAB[(i + 1) * W + j] * P(kAom, i + 1, j);
    if (l == nl - 1) a += -g.drag * mu_v(g, lv, l, i, j);
    if (g.visc > 0)  // (nu lap(v / dy)) dy
      a += g.visc * lap5_adj(mu_v(g, lv, l, i, j), mu_v(g, lv, l, i - 1, j), mu_v(g, lv, l, i + 1, j),
                             mu_v(g, lv, l, i, j - 1), mu_v(g, lv, l, i, j + 1), (i == 0) + (i == g.nx - 1), 0,
                             g);
    const double L = ov[iv_(g, l, i, j)] + a;
    ov[iv_(g, l, i, j)] = L;
    cb.apply(nu + iv_(g, l, i, j), L);
  }
  if (i < g.nx && j < g.ny) {  // h cell (i, j); ((kappa lap(h / area)) area) mask_h
    double a = P(kAhp, i, j);
    if (g.diff > 0)
      a += g.diff * lap5_adj(mu_h(g, lh, l, i, j), mu_h(g, lh, l, i - 1, j), mu_h(g, lh, l, i + 1, j),
                             mu_h(g, lh, l, i, j - 1), mu_h(g, lh, l, i, j + 1), (i == 0) + (i == g.nx - 1),
                             (j == 0) + (j == g.ny - 1), g);
    const double L = oh[ih_(g, l, i, j)] + a;
    oh[ih_(g, l, i, j)] = L;
    cb.apply(nu + nv + ih_(g, l, i, j), L);
  }
}

// y_k = a_k y_k + b_k x (a_k = 0: y_k = b_k x) for up to three outputs.
__global__ void k_sw_comb(int n, const double* x, double* y0, double a0, double b0, double* y1, double a1, double b1,
                          double* y2, double a2, double b2, double* y3, double a3, double b3) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  const double xv = x[e];
  if (y0) y0[e] = a0 == 0.0 ? b0 * xv : a0 * y0[e] + b0 * xv;
  if (y1) y1[e] = a1 == 0.0 ? b1 * xv : a1 * y1[e] + b1 * xv;
  if (y2) y2[e] = a2 == 0.0 ? b2 * xv : a2 * y2[e] + b2 * xv;
  if (y3) y3[e] = a3 == 0.0 ? b3 * xv : a3 * y3[e] + b3 * xv;
}

// ------------------------------------------------------------ host: masks
// avg_pool2d (tools.py:9-67) of an n0 x n1 array: zero padding (p0, p1),
// window k0 x k1, stride 1, sum divided by div.
struct Arr {
  int n0, n1;
  std::vector<double> a;
  double at(int i, int j) const { return (i >= 0 && i < n0 && j >= 0 && j < n1) ? a[(size_t)i * n1 + j] : 0.0; }
};
static Arr pool(const Arr& x, int k0, int k1, int p0, int p1, double div) {
  Arr o;
  o.n0 = x.n0 + 2 * p0 - k0 + 1;
  o.n1 = x.n1 + 2 * p1 - k1 + 1;
  o.a.assign((size_t)o.n0 * o.n1, 0.0);
  for (int i = 0; i < o.n0; ++i)
    for (int j = 0; j < o.n1; ++j) {
      double sum = 0.0;
      for (int di = 0; di < k0; ++di)
        for (int dj = 0; dj < k1; ++dj) sum += x.at(i + di - p0, j + dj - p1);
      o.a[(size_t)i * o.n1 + j] = sum / div;
    }
  return o;
}
static Arr cmp(const Arr& x, bool (*f)(double)) {
  Arr o = x;
  for (auto& e : o.a) e = f(e) ? 1.0 : 0.0;
  return o;
}
static Arr land(const Arr& x, const Arr& y) {
  Arr o = x;
  for (size_t k = 0; k < o.a.size(); ++k) o.a[k] = (x.a[k] != 0.0 && y.a[k] != 0.0) ? 1.0 : 0.0;
  return o;
}
static Arr lnot(const Arr& x) {
  Arr o = x;
  for (auto& e : o.a) e = e != 0.0 ? 0.0 : 1.0;
  return o;
}

struct Masks {
  Arr h, u, v, w, wAB, wAC;
  std::vector<uint8_t> hx, wy, hy, wx;
};

// Stencil class of a face from its masks: which of the 2-, 4- and
// 6-point reconstructions the flux adds.
static std::vector<uint8_t> codes(const Arr& m2, const Arr& m4, const Arr& m6) {
  std::vector<uint8_t> c(m2.a.size());
  for (size_t k = 0; k < c.size(); ++k)
    c[k] = (m2.a[k] != 0.0 ? kSten2 : 0) | (m4.a[k] != 0.0 ? kSten4 : 0) | (m6.a[k] != 0.0 ? kSten6 : 0);
  return c;
}

// masks.py:64-159
static Masks build_masks(int nx, int ny, const uint8_t* mask, double slip) {
  Masks M;
  M.h.n0 = nx;
  M.h.n1 = ny;
  M.h.a.assign((size_t)nx * ny, 1.0);
  if (mask)
    for (size_t k = 0; k < M.h.a.size(); ++k) M.h.a[k] = mask[k] ? 1.0 : 0.0;
  auto gt34 = [](double x) { return x > 3.0 / 4.0; };
  auto gt18 = [](double x) { return x > 1.0 / 8.0; };
  auto lt78 = [](double x) { return x < 7.0 / 8.0; };
  auto lt1112 = [](double x) { return x < 11.0 / 12.0; };
  auto eq1 = [](double x) { return x == 1.0; };
  M.u = cmp(pool(M.h, 2, 1, 1, 0, 2), gt34);
  M.v = cmp(pool(M.h, 1, 2, 0, 1, 2), gt34);
  M.w = cmp(pool(M.h, 2, 2, 1, 1, 4), gt18);
  const Arr one_v = cmp(pool(M.v, 2, 1, 1, 0, 1), eq1);
  const Arr one_u = cmp(pool(M.u, 1, 2, 0, 1, 1), eq1);
  const Arr vb = land(M.w, land(one_v, lnot(one_u)));
  const Arr hb = land(M.w, land(one_u, lnot(one_v)));
  const Arr co = land(M.w, land(one_u, one_v));
  const Arr valid = land(M.w, land(lnot(vb), land(lnot(hb), lnot(co))));
  // omega = valid curl + (1 - slip) cornerout curl + alpha (vertical dx_v
  // - horizontal dy_u)  (sw.py:653-671), as wAB dx_v - wAC dy_u
  const double alpha = 2.0 * (1.0 - slip);
  M.wAB = M.w;
  M.wAC = M.w;
  for (size_t k = 0; k < M.w.a.size(); ++k) {
    const double A = valid.a[k] + co.a[k] * (1.0 - slip);
    M.wAB.a[k] = A + vb.a[k] * alpha;
    M.wAC.a[k] = A + hb.a[k] * alpha;
  }
  // h along x on the u grid
  {
    const Arr eq2 = land(land(cmp(pool(M.h, 2, 1, 1, 0, 2), gt34), cmp(pool(M.h, 4, 1, 2, 0, 4), lt78)), M.u);
    const Arr gt4 = land(lnot(eq2), M.u);
    const Arr eq4 = land(cmp(pool(M.h, 6, 1, 3, 0, 6), lt1112), gt4);
    M.hx = codes(eq2, eq4, land(lnot(eq4), gt4));
  }
  // vorticity along y on the u grid: the flux blends the 6-point
  // reconstruction with the ">= 4 points" mask (sw.py:285-293)
  {
    const Arr eq2 = land(land(cmp(pool(M.w, 1, 2, 0, 0, 2), gt34), cmp(pool(M.w, 1, 4, 0, 1, 4), lt78)), M.u);
    const Arr gt4 = land(lnot(eq2), M.u);
    const Arr eq4 = land(cmp(pool(M.w, 1, 6, 0, 2, 6), lt1112), gt4);
    M.wy = codes(eq2, eq4, gt4);
  }
  // h along y on the v grid
  {
    const Arr eq2 = land(land(cmp(pool(M.h, 1, 2, 0, 1, 2), gt34), cmp(pool(M.h, 1, 4, 0, 2, 4), lt78)), M.v);
    const Arr gt4 = land(lnot(eq2), M.v);
    const Arr eq4 = land(cmp(pool(M.h, 1, 6, 0, 3, 6), lt1112), gt4);
    M.hy = codes(eq2, eq4, land(lnot(eq4), gt4));
  }
  // vorticity along x on the v grid
  {
    const Arr eq2 = land(land(cmp(pool(M.w, 2, 1, 0, 0, 2), gt34), cmp(pool(M.w, 4, 1, 1, 0, 4), lt78)), M.v);
    const Arr gt4 = land(lnot(eq2), M.v);
    const Arr eq4 = land(cmp(pool(M.w, 6, 1, 2, 0, 6), lt1112), gt4);
    M.wx = codes(eq2, eq4, land(lnot(eq4), gt4));
  }
  return M;
}

}  // namespace sw
}  // namespace vd

// ------------------------------------------------------------ C ABI
using namespace vd::sw;

struct vd_sw_model {
  int dev = 0;
  SwGrid g{};
  size_t nu = 0, nv = 0, nh = 0, ns = 0, nwl = 0;
  std::vector<void*> allocs;
  double *SA = nullptr, *SB = nullptr, *SC = nullptr, *TA = nullptr, *TB = nullptr, *TC = nullptr;
  double *K0 = nullptr, *K1 = nullptr, *TK0 = nullptr, *TK1 = nullptr;
  double *om = nullptr, *ht = nullptr, *kep = nullptr, *tom = nullptr, *tht = nullptr, *tkep = nullptr;
  double* ws = nullptr;
  double *L = nullptr, *LK0 = nullptr, *LK1 = nullptr, *LK2 = nullptr;
  double *F = nullptr, *dF = nullptr, *aF = nullptr;  // scaled forcing, its tangent and cotangent
  double* tape = nullptr;    // segment start states of vd_sw_step_adj
  size_t tape_states = 0;
  double* segbuf = nullptr;  // the step states of one segment in the reverse sweep
  size_t seg_states = 0;
  ~vd_sw_model() {
    for (void* p : allocs) (void)hipFree(p);
    if (tape) (void)hipFree(tape);
    if (segbuf) (void)hipFree(segbuf);
  }
};

namespace {

template <typename T>
int sw_alloc(vd_sw_model* m, size_t n, T** p) {
  void* q = nullptr;
  VD_HIP_TRY(hipMalloc(&q, std::max<size_t>(n, 1) * sizeof(T)));
  m->allocs.push_back(q);
  *p = static_cast<T*>(q);
  return VD_OK;
}
template <typename T>
int sw_upload(vd_sw_model* m, const std::vector<T>& h, const T** p) {
  T* q = nullptr;
  int rc = sw_alloc(m, h.size(), &q);
  if (rc) return rc;
  VD_HIP_TRY(hipMemcpy(q, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice));
  *p = q;
  return VD_OK;
}

St st_of(const vd_sw_model* m, const double* base, const double* tan) {
  return St{{base, tan}, {base + m->nu, tan ? tan + m->nu : nullptr},
            {base + m->nu + m->nv, tan ? tan + m->nu + m->nv : nullptr}};
}
StOut sto_of(const vd_sw_model* m, double* base, double* tan) {
  return StOut{{base, tan}, {base + m->nu, tan ? tan + m->nu : nullptr},
               {base + m->nu + m->nv, tan ? tan + m->nu + m->nv : nullptr}};
}

dim3 blk() { return dim3(kSwBX, kSwBY, 1); }
dim3 grd(const SwGrid& g, int nz) { return dim3((g.ny + 1 + kSwBX - 1) / kSwBX, (g.nx + 1 + kSwBY - 1) / kSwBY, nz); }

// One RK3-SSP step x <- step(x) on buffer x (value, tangent when TAN),
// through the stage buffers B, C.
template <bool TAN>
int sw_rk3(vd_sw_model* m, double* xv, double* xt, hipStream_t s, double* x1 = nullptr, double* x2 = nullptr,
           double* x3 = nullptr) {
  typedef typename std::conditional<TAN, D1, double>::type S;
  const SwGrid& g = m->g;
  // value buffers of the stage states: x1, x2 (default SB, SC) and the
  // step's result x3 (default: over x0); x0 is left intact when x3 is given
  double* bv[4] = {xv, x1 ? x1 : m->SB, x2 ? x2 : m->SC, x3 ? x3 : xv};
  double* bt[3] = {xt, TAN ? m->TB : nullptr, TAN ? m->TC : nullptr};
  const St k0 = st_of(m, m->K0, TAN ? m->TK0 : nullptr), k1 = st_of(m, m->K1, TAN ? m->TK1 : nullptr);
  const StOut k0o = sto_of(m, m->K0, TAN ? m->TK0 : nullptr), k1o = sto_of(m, m->K1, TAN ? m->TK1 : nullptr);
  const FldOut omo{m->om, TAN ? m->tom : nullptr}, hto{m->ht, TAN ? m->tht : nullptr},
      kepo{m->kep, TAN ? m->tkep : nullptr};
  const Fld omi{m->om, TAN ? m->tom : nullptr}, hti{m->ht, TAN ? m->tht : nullptr},
      kepi{m->kep, TAN ? m->tkep : nullptr};
  for (int stg = 0; stg < 3; ++stg) {
    const St x = st_of(m, bv[stg], bt[stg]);
    const StOut xn = sto_of(m, bv[stg + 1], bt[(stg + 1) % 3]);
    hipLaunchKernelGGL(k_sw_diag<S>, grd(g, g.nl), blk(), 0, s, g, x, omo, hto, kepo);
    if (stg == 0) hipLaunchKernelGGL((k_sw_tend<S, 0>), grd(g, g.nl), blk(), 0, s, g, x, omi, hti, kepi, k0, k1, k0o, xn);
    else if (stg == 1)
      hipLaunchKernelGGL((k_sw_tend<S, 1>), grd(g, g.nl), blk(), 0, s, g, x, omi, hti, kepi, k0, k1, k1o, xn);
    else
      hipLaunchKernelGGL((k_sw_tend<S, 2>), grd(g, g.nl), blk(), 0, s, g, x, omi, hti, kepi, k0, k1, k1o, xn);
  }
  VD_HIP_TRY(hipGetLastError());
  return VD_OK;
}

void scale3(const vd_sw_model* m, const double* iu, const double* iv, const double* ih, double* ou, double* ov,
            double* oh, int mode, hipStream_t s) {
  const SwGrid& g = m->g;
  const double* ins[3] = {iu, iv, ih};
  double* outs[3] = {ou, ov, oh};
  const double* masks[3] = {g.mask_u, g.mask_v, g.mask_h};
  const int planes[3] = {(g.nx + 1) * g.ny, g.nx * (g.ny + 1), g.nx * g.ny};
  const double fac[3] = {g.dx, g.dy, g.area};
  const size_t cnt[3] = {m->nu, m->nv, m->nh};
  for (int f = 0; f < 3; ++f) {
    const int c = (int)cnt[f];
    if (!outs[f]) continue;
    hipLaunchKernelGGL(k_sw_scale_field, dim3((c + 255) / 256), dim3(256), 0, s, ins[f], outs[f], masks[f], planes[f],
                       fac[f], mode, c);
  }
}

// F'(x)^T mu added to L, x a scaled state, mu (three fields) at mu_base.
int sw_rhs_adj(vd_sw_model* m, const double* x, const double* mu, hipStream_t s,
               StageComb cb = StageComb{{nullptr, nullptr}, {0, 0}, {0, 0}}) {
  const SwGrid& g = m->g;
  const St xs = st_of(m, x, nullptr);
  hipLaunchKernelGGL(k_sw_diag<double>, grd(g, g.nl), blk(), 0, s, g, xs, FldOut{m->om, nullptr},
                     FldOut{m->ht, nullptr}, FldOut{m->kep, nullptr});
  AdjWs ws{m->ws, (int)m->nwl};
  const double *mu_u = mu, *mu_v = mu + m->nu, *mu_h = mu + m->nu + m->nv;
  hipLaunchKernelGGL(k_sw_adj_face, grd(g, g.nl), blk(), 0, s, g, x, x + m->nu, m->om, m->ht, mu_u, mu_v, mu_h, ws);
  hipLaunchKernelGGL(k_sw_adj_gather1, grd(g, g.nl), blk(), 0, s, g, x + m->nu + m->nv, mu_u, mu_v, ws);
  hipLaunchKernelGGL(k_sw_adj_gather2, grd(g, g.nl), blk(), 0, s, g, x, x + m->nu, mu_u, mu_v, mu_h, ws, m->L,
                     m->L + m->nu, m->L + m->nu + m->nv, cb);
  VD_HIP_TRY(hipGetLastError());
  return VD_OK;
}

void comb(const vd_sw_model* m, const double* x, double* y0, double a0, double b0, double* y1, double a1, double b1,
          double* y2, double a2, double b2, hipStream_t s, double* y3 = nullptr, double a3 = 0, double b3 = 0) {
  const int n = (int)m->ns;
  hipLaunchKernelGGL(k_sw_comb, dim3((n + 255) / 256), dim3(256), 0, s, n, x, y0, a0, b0, y1, a1, b1, y2, a2, b2,
                     y3, a3, b3);
}

// Scaled forcing (set_input_uvh of Fu, Fv, Fh; a missing field is zero)
// written to dst; false when all three are missing.
bool load_forcing(const vd_sw_model* m, const double* Fu, const double* Fv, const double* Fh, double* dst,
                  hipStream_t s) {
  if (!Fu && !Fv && !Fh) return false;
  scale3(m, Fu, Fv, Fh, dst, dst + m->nu, dst + m->nu + m->nv, 0, s);
  return true;
}

// x <- x + dt F: the external forcing after each RK3 step (sw.py:875-877).
void add_forcing(const vd_sw_model* m, double* x, const double* F, hipStream_t s) {
  comb(m, F, x, 1.0, m->g.dt, nullptr, 0, 0, nullptr, 0, 0, s);
}

// L <- (one RK3 step at x0)^T L
// L <- (one RK3 step at x0)^T L with the stage states x1, x2 given.
// With aF, the forcing cotangent aF += dt L of the step's result (x3 + dt F)
// is collected in the same launch as the first stage combination.
int sw_rk3_adj_states(vd_sw_model* m, const double* x0, const double* x1, const double* x2, hipStream_t s,
                      double* aF = nullptr) {
  const double dt = m->g.dt;
  int rc;
  // x3 = x2 + dt/12 (8 k2 - k1 - k0)
  comb(m, m->L, m->LK2, 0.0, (dt / 12.0) * 8.0, m->LK1, 0.0, -(dt / 12.0), m->LK0, 0.0, -(dt / 12.0), s, aF, 1.0,
       dt);
  // x2 = x1 + dt/4 (k1 - 3 k0): its combination fused into the evaluation's last kernel
  if ((rc = sw_rhs_adj(m, x2, m->LK2, s,
                       StageComb{{m->LK1, m->LK0}, {1.0, 1.0}, {dt / 4.0, -3.0 * (dt / 4.0)}})))
    return rc;
  // x1 = x0 + dt k0
  if ((rc = sw_rhs_adj(m, x1, m->LK1, s, StageComb{{m->LK0, nullptr}, {1.0, 0.0}, {dt, 0.0}}))) return rc;
  return sw_rhs_adj(m, x0, m->LK0, s);
}

// L <- (one RK3 step at x0)^T L, the stage states recomputed from x0.
int sw_rk3_adj(vd_sw_model* m, const double* x0, hipStream_t s) {
  // stage states x1 (SB), x2 (SC) from x0: the forward stages 0 and 1
  {
    const SwGrid& g = m->g;
    const St k0 = st_of(m, m->K0, nullptr), k1 = st_of(m, m->K1, nullptr);
    const StOut k0o = sto_of(m, m->K0, nullptr), k1o = sto_of(m, m->K1, nullptr);
    const FldOut omo{m->om, nullptr}, hto{m->ht, nullptr}, kepo{m->kep, nullptr};
    const Fld omi{m->om, nullptr}, hti{m->ht, nullptr}, kepi{m->kep, nullptr};
    St x = st_of(m, x0, nullptr);
    hipLaunchKernelGGL(k_sw_diag<double>, grd(g, g.nl), blk(), 0, s, g, x, omo, hto, kepo);
    hipLaunchKernelGGL((k_sw_tend<double, 0>), grd(g, g.nl), blk(), 0, s, g, x, omi, hti, kepi, k0, k1, k0o,
                       sto_of(m, m->SB, nullptr));
    x = st_of(m, m->SB, nullptr);
    hipLaunchKernelGGL(k_sw_diag<double>, grd(g, g.nl), blk(), 0, s, g, x, omo, hto, kepo);
    hipLaunchKernelGGL((k_sw_tend<double, 1>), grd(g, g.nl), blk(), 0, s, g, x, omi, hti, kepi, k0, k1, k1o,
                       sto_of(m, m->SC, nullptr));
  }
  return sw_rk3_adj_states(m, x0, m->SB, m->SC, s);
}

}  // namespace

extern "C" {

int vd_sw_create(const vd_sw_params* P, vd_sw_model** out) {
  VD_REQUIRE(P && out, VD_ERR_INVALID_ARG, "null argument");
  *out = nullptr;
  VD_REQUIRE(P->nx >= 6 && P->ny >= 6, VD_ERR_INVALID_ARG, "sw: the grid needs at least 6 x 6 cells");
  VD_REQUIRE(P->nl >= 1 && P->nl <= 8, VD_ERR_INVALID_ARG, "sw: 1 to 8 layers");
  VD_REQUIRE((long long)P->nl * (P->nx + 1) * (P->ny + 1) < (1LL << 30), VD_ERR_INVALID_ARG, "sw: grid too large");
  VD_REQUIRE(P->dx > 0 && P->dy > 0 && P->dt > 0, VD_ERR_INVALID_ARG, "sw: dx, dy and dt must be positive");
  VD_REQUIRE(P->H && P->g_prime && P->f, VD_ERR_INVALID_ARG, "sw: H, g_prime and f are required");
  VD_REQUIRE(P->slip_coef >= 0 && P->slip_coef <= 1, VD_ERR_INVALID_ARG, "sw: slip coefficient must be in [0, 1]");
  VD_REQUIRE(P->h_min_sharpness > 0, VD_ERR_INVALID_ARG, "sw: h_min_sharpness must be positive");
  vd_sw_model* m = new vd_sw_model();
  (void)hipGetDevice(&m->dev);
  SwGrid& g = m->g;
  g.nx = P->nx;
  g.ny = P->ny;
  g.nl = P->nl;
  g.dx = P->dx;
  g.dy = P->dy;
  g.area = P->dx * P->dy;
  g.dt = P->dt;
  g.h_min_area = P->h_min * g.area;
  g.sharp = P->h_min_sharpness;
  g.drag = P->bottom_drag_coef;
  // only a positive coefficient diffuses (sw.py:517, 555)
  g.visc = P->visc_coef > 0 ? P->visc_coef : 0.0;
  g.diff = P->diff_coef > 0 ? P->diff_coef : 0.0;
  for (int l = 0; l < 8; ++l) {
    g.href[l] = l < P->nl ? P->H[l] * g.area : 0.0;
    g.gp[l] = l < P->nl ? P->g_prime[l] : 0.0;
  }
  const int nx = P->nx, ny = P->ny, nl = P->nl;
  const Masks M = build_masks(nx, ny, P->mask, P->slip_coef);
  std::vector<double> fu((size_t)(nx + 1) * ny), fv((size_t)nx * (ny + 1));
  for (int i = 0; i <= nx; ++i)
    for (int j = 0; j < ny; ++j)
      fu[(size_t)i * ny + j] = 0.5 * (P->f[(size_t)i * (ny + 1) + j + 1] + P->f[(size_t)i * (ny + 1) + j]) * g.area;
  for (int i = 0; i < nx; ++i)
    for (int j = 0; j <= ny; ++j)
      fv[(size_t)i * (ny + 1) + j] =
          0.5 * (P->f[(size_t)(i + 1) * (ny + 1) + j] + P->f[(size_t)i * (ny + 1) + j]) * g.area;
  int rc = VD_OK;
  m->nu = (size_t)nl * (nx + 1) * ny;
  m->nv = (size_t)nl * nx * (ny + 1);
  m->nh = (size_t)nl * nx * ny;
  m->ns = m->nu + m->nv + m->nh;
  m->nwl = (size_t)(nx + 1) * (ny + 1);
  const size_t nw = nl * m->nwl;
  if ((rc = sw_upload(m, M.u.a, &g.mask_u)) || (rc = sw_upload(m, M.v.a, &g.mask_v)) ||
      (rc = sw_upload(m, M.h.a, &g.mask_h)) || (rc = sw_upload(m, M.wAB.a, &g.wAB)) ||
      (rc = sw_upload(m, M.wAC.a, &g.wAC)) || (rc = sw_upload(m, fu, &g.fstar_u)) ||
      (rc = sw_upload(m, fv, &g.fstar_v)) || (rc = sw_upload(m, M.hx, &g.code_hx)) ||
      (rc = sw_upload(m, M.hy, &g.code_hy)) || (rc = sw_upload(m, M.wy, &g.code_wy)) ||
      (rc = sw_upload(m, M.wx, &g.code_wx)) || (rc = sw_alloc(m, m->ns, &m->SA)) ||
      (rc = sw_alloc(m, m->ns, &m->SB)) || (rc = sw_alloc(m, m->ns, &m->SC)) || (rc = sw_alloc(m, m->ns, &m->TA)) ||
      (rc = sw_alloc(m, m->ns, &m->TB)) || (rc = sw_alloc(m, m->ns, &m->TC)) || (rc = sw_alloc(m, m->ns, &m->K0)) ||
      (rc = sw_alloc(m, m->ns, &m->K1)) || (rc = sw_alloc(m, m->ns, &m->TK0)) ||
      (rc = sw_alloc(m, m->ns, &m->TK1)) || (rc = sw_alloc(m, nw, &m->om)) || (rc = sw_alloc(m, m->nh, &m->ht)) ||
      (rc = sw_alloc(m, m->nh, &m->kep)) || (rc = sw_alloc(m, nw, &m->tom)) || (rc = sw_alloc(m, m->nh, &m->tht)) ||
      (rc = sw_alloc(m, m->nh, &m->tkep)) || (rc = sw_alloc(m, (size_t)kAdjPlanes * nw, &m->ws)) ||
      (rc = sw_alloc(m, m->ns, &m->L)) || (rc = sw_alloc(m, m->ns, &m->LK0)) || (rc = sw_alloc(m, m->ns, &m->LK1)) ||
      (rc = sw_alloc(m, m->ns, &m->LK2)) || (rc = sw_alloc(m, m->ns, &m->F)) || (rc = sw_alloc(m, m->ns, &m->dF)) ||
      (rc = sw_alloc(m, m->ns, &m->aF))) {
    delete m;
    return rc;
  }
  // the wind tendency is a constant of the plan (no state dependence: nothing
  // in the tangent or the adjoint)
  if (P->wind_u || P->wind_v) {
    std::vector<double> wu((size_t)(nx + 1) * ny, 0.0), wv((size_t)nx * (ny + 1), 0.0);
    if (P->wind_u) std::copy(P->wind_u, P->wind_u + wu.size(), wu.begin());
    if (P->wind_v) std::copy(P->wind_v, P->wind_v + wv.size(), wv.begin());
    if ((rc = sw_upload(m, wu, &g.wind_u)) || (rc = sw_upload(m, wv, &g.wind_v))) {
      delete m;
      return rc;
    }
  }
  *out = m;
  return VD_OK;
}

int vd_sw_destroy(vd_sw_model* m) {
  delete m;
  return VD_OK;
}

int vd_sw_info(const vd_sw_model* m, long long* nu, long long* nv, long long* nh) {
  VD_REQUIRE(m, VD_ERR_INVALID_ARG, "null model");
  if (nu) *nu = (long long)m->nu;
  if (nv) *nv = (long long)m->nv;
  if (nh) *nh = (long long)m->nh;
  return VD_OK;
}

int vd_sw_rhs(vd_sw_model* m, const double* u, const double* v, const double* h, double* du, double* dv,
              double* dh, void* stream) {
  VD_REQUIRE(m && u && v && h && du && dv && dh, VD_ERR_INVALID_ARG, "null argument");
  hipStream_t s = static_cast<hipStream_t>(stream);
  const SwGrid& g = m->g;
  VD_HIP_TRY(hipMemcpyAsync(m->SB, u, m->nu * sizeof(double), hipMemcpyDeviceToDevice, s));
  VD_HIP_TRY(hipMemcpyAsync(m->SB + m->nu, v, m->nv * sizeof(double), hipMemcpyDeviceToDevice, s));
  VD_HIP_TRY(hipMemcpyAsync(m->SB + m->nu + m->nv, h, m->nh * sizeof(double), hipMemcpyDeviceToDevice, s));
  const St x = st_of(m, m->SB, nullptr);
  hipLaunchKernelGGL(k_sw_diag<double>, grd(g, g.nl), blk(), 0, s, g, x, FldOut{m->om, nullptr},
                     FldOut{m->ht, nullptr}, FldOut{m->kep, nullptr});
  // stage 0 with the tendency stored in K0 (the updated state goes to SC)
  hipLaunchKernelGGL((k_sw_tend<double, 0>), grd(g, g.nl), blk(), 0, s, g, x, Fld{m->om, nullptr},
                     Fld{m->ht, nullptr}, Fld{m->kep, nullptr}, st_of(m, m->K0, nullptr), st_of(m, m->K1, nullptr),
                     sto_of(m, m->K0, nullptr), sto_of(m, m->SC, nullptr));
  VD_HIP_TRY(hipMemcpyAsync(du, m->K0, m->nu * sizeof(double), hipMemcpyDeviceToDevice, s));
  VD_HIP_TRY(hipMemcpyAsync(dv, m->K0 + m->nu, m->nv * sizeof(double), hipMemcpyDeviceToDevice, s));
  VD_HIP_TRY(hipMemcpyAsync(dh, m->K0 + m->nu + m->nv, m->nh * sizeof(double), hipMemcpyDeviceToDevice, s));
  VD_HIP_TRY(hipGetLastError());
  return VD_OK;
}

int vd_sw_step(vd_sw_model* m, const double* u0, const double* v0, const double* h0, const double* Fu,
               const double* Fv, const double* Fh, double* u1, double* v1, double* h1, int nstep, void* stream) {
  VD_REQUIRE(m && u0 && v0 && h0 && u1 && v1 && h1, VD_ERR_INVALID_ARG, "null argument");
  VD_REQUIRE(nstep >= 0, VD_ERR_INVALID_ARG, "sw: nstep must be >= 0");
  hipStream_t s = static_cast<hipStream_t>(stream);
  const bool forced = load_forcing(m, Fu, Fv, Fh, m->F, s);
  scale3(m, u0, v0, h0, m->SA, m->SA + m->nu, m->SA + m->nu + m->nv, 0, s);
  for (int k = 0; k < nstep; ++k) {
    int rc = sw_rk3<false>(m, m->SA, nullptr, s);
    if (rc) return rc;
    if (forced) add_forcing(m, m->SA, m->F, s);
  }
  scale3(m, m->SA, m->SA + m->nu, m->SA + m->nu + m->nv, u1, v1, h1, 1, s);
  VD_HIP_TRY(hipGetLastError());
  return VD_OK;
}

int vd_sw_step_tgl(vd_sw_model* m, const double* u0, const double* v0, const double* h0, const double* Fu,
                   const double* Fv, const double* Fh, const double* du0, const double* dv0, const double* dh0,
                   const double* dFu, const double* dFv, const double* dFh, double* du1, double* dv1, double* dh1,
                   int nstep, void* stream) {
  VD_REQUIRE(m && u0 && v0 && h0 && du0 && dv0 && dh0 && du1 && dv1 && dh1, VD_ERR_INVALID_ARG, "null argument");
  VD_REQUIRE(nstep >= 0, VD_ERR_INVALID_ARG, "sw: nstep must be >= 0");
  hipStream_t s = static_cast<hipStream_t>(stream);
  const bool forced = load_forcing(m, Fu, Fv, Fh, m->F, s);
  const bool dforced = load_forcing(m, dFu, dFv, dFh, m->dF, s);
  scale3(m, u0, v0, h0, m->SA, m->SA + m->nu, m->SA + m->nu + m->nv, 0, s);
  scale3(m, du0, dv0, dh0, m->TA, m->TA + m->nu, m->TA + m->nu + m->nv, 0, s);
  for (int k = 0; k < nstep; ++k) {
    int rc = sw_rk3<true>(m, m->SA, m->TA, s);
    if (rc) return rc;
    if (forced) add_forcing(m, m->SA, m->F, s);
    if (dforced) add_forcing(m, m->TA, m->dF, s);
  }
  scale3(m, m->TA, m->TA + m->nu, m->TA + m->nu + m->nv, du1, dv1, dh1, 1, s);
  VD_HIP_TRY(hipGetLastError());
  return VD_OK;
}

// Reverse sweep over nstep steps from kept scaled states: state k*seg at
// ck + k*ns (k < ceil(nstep/seg)); L holds the cotangent of the final scaled
// state and is overwritten by the initial one; aF (when wanted) collects the
// scaled forcing cotangent.  Inside a sub-segment the step states are
// recomputed into the model's segment buffer (none for seg = 1).
static int sw_reverse(vd_sw_model* m, bool forced, const double* ck, int seg, int nstep, bool want_aF,
                      hipStream_t s) {
  const size_t ns = m->ns;
  if (seg == 0) {
    // every step's start state and its two RK3 stage states were kept
    for (int k = nstep - 1; k >= 0; --k) {
      const double* xk = ck + (size_t)3 * k * ns;
      int rc = sw_rk3_adj_states(m, xk, xk + ns, xk + 2 * ns, s, want_aF ? m->aF : nullptr);
      if (rc) return rc;
    }
    return VD_OK;
  }
  const int nseg = nstep > 0 ? (nstep + seg - 1) / seg : 0;
  double* st = nullptr;
  if (seg > 1) {
    if (m->seg_states < (size_t)seg) {
      if (m->segbuf) VD_HIP_TRY(hipFree(m->segbuf));
      m->segbuf = nullptr;
      m->seg_states = 0;
      VD_HIP_TRY(hipMalloc(&m->segbuf, (size_t)seg * ns * sizeof(double)));
      m->seg_states = seg;
    }
    st = m->segbuf;
  }
  for (int q = nseg - 1; q >= 0; --q) {
    const int s0 = q * seg, s1 = std::min(nstep, s0 + seg);
    if (seg > 1) {
      VD_HIP_TRY(hipMemcpyAsync(m->SA, ck + (size_t)q * ns, ns * sizeof(double), hipMemcpyDeviceToDevice, s));
      for (int k = s0; k < s1; ++k) {
        VD_HIP_TRY(
            hipMemcpyAsync(st + (size_t)(k - s0) * ns, m->SA, ns * sizeof(double), hipMemcpyDeviceToDevice, s));
        if (k + 1 < s1) {
          int rc = sw_rk3<false>(m, m->SA, nullptr, s);
          if (rc) return rc;
          if (forced) add_forcing(m, m->SA, m->F, s);
        }
      }
    }
    for (int k = s1 - 1; k >= s0; --k) {
      // x_{k+1} = RK3(x_k) + dt F: F collects dt L_{k+1} before L goes back through the step
      if (want_aF) add_forcing(m, m->aF, m->L, s);
      const double* xk = seg > 1 ? st + (size_t)(k - s0) * ns : ck + (size_t)q * ns;
      int rc = sw_rk3_adj(m, xk, s);
      if (rc) return rc;
    }
  }
  return VD_OK;
}

int vd_sw_step_adj(vd_sw_model* m, const double* u0, const double* v0, const double* h0, const double* Fu,
                   const double* Fv, const double* Fh, const double* au1, const double* av1, const double* ah1,
                   double* au0, double* av0, double* ah0, double* aFu, double* aFv, double* aFh, int nstep,
                   void* stream) {
  VD_REQUIRE(m && u0 && v0 && h0 && au1 && av1 && ah1 && au0 && av0 && ah0, VD_ERR_INVALID_ARG, "null argument");
  VD_REQUIRE(nstep >= 0, VD_ERR_INVALID_ARG, "sw: nstep must be >= 0");
  hipStream_t s = static_cast<hipStream_t>(stream);
  const size_t ns = m->ns;
  const bool forced = load_forcing(m, Fu, Fv, Fh, m->F, s);
  const bool want_aF = aFu || aFv || aFh;
  // segments of about sqrt(nstep) steps: their start states from a forward
  // pass, then one segment's step states at a time
  const int seg = std::max(1, (int)std::ceil(std::sqrt((double)nstep)));
  const int nseg = nstep > 0 ? (nstep + seg - 1) / seg : 0;
  if (m->tape_states < (size_t)nseg) {
    if (m->tape) VD_HIP_TRY(hipFree(m->tape));
    m->tape = nullptr;
    m->tape_states = 0;
    VD_HIP_TRY(hipMalloc(&m->tape, std::max(nseg, 1) * ns * sizeof(double)));
    m->tape_states = nseg;
  }
  double* ck = m->tape;
  scale3(m, u0, v0, h0, m->SA, m->SA + m->nu, m->SA + m->nu + m->nv, 0, s);
  for (int k = 0; k < nstep; ++k) {
    if (k % seg == 0) VD_HIP_TRY(hipMemcpyAsync(ck + (size_t)(k / seg) * ns, m->SA, ns * sizeof(double),
                                                hipMemcpyDeviceToDevice, s));
    int rc = sw_rk3<false>(m, m->SA, nullptr, s);
    if (rc) return rc;
    if (forced) add_forcing(m, m->SA, m->F, s);
  }
  // cotangent of the scaled final state: get_physical_uvh divides by dx, dy, dx dy
  scale3(m, au1, av1, ah1, m->L, m->L + m->nu, m->L + m->nu + m->nv, 1, s);
  // the forcing cotangent starts at zero (a missing input scales to zero)
  if (want_aF) scale3(m, nullptr, nullptr, nullptr, m->aF, m->aF + m->nu, m->aF + m->nu + m->nv, 0, s);
  int rc = sw_reverse(m, forced, ck, seg, nstep, want_aF, s);
  if (rc) return rc;
  // set_input_uvh: u_scaled = (u mask_u) dx
  scale3(m, m->L, m->L + m->nu, m->L + m->nu + m->nv, au0, av0, ah0, 2, s);
  if (want_aF) scale3(m, m->aF, m->aF + m->nu, m->aF + m->nu + m->nv, aFu, aFv, aFh, 2, s);
  VD_HIP_TRY(hipGetLastError());
  return VD_OK;
}

int vd_sw_step_keep(vd_sw_model* m, const double* u0, const double* v0, const double* h0, const double* Fu,
                    const double* Fv, const double* Fh, double* u1, double* v1, double* h1, int nstep, int stride,
                    double* keep, void* stream) {
  VD_REQUIRE(m && u0 && v0 && h0 && u1 && v1 && h1 && keep, VD_ERR_INVALID_ARG, "null argument");
  VD_REQUIRE(nstep >= 0 && stride >= 0, VD_ERR_INVALID_ARG, "sw: nstep >= 0 and stride >= 0");
  hipStream_t s = static_cast<hipStream_t>(stream);
  const size_t ns = m->ns;
  const bool forced = load_forcing(m, Fu, Fv, Fh, m->F, s);
  if (stride == 0 && nstep > 0) {
    // every step's start and stage states: the stages write them straight
    // into the tape (slots 3k, 3k+1, 3k+2), the last step's result into SA
    scale3(m, u0, v0, h0, keep, keep + m->nu, keep + m->nu + m->nv, 0, s);
    for (int k = 0; k < nstep; ++k) {
      double* x0 = keep + (size_t)3 * k * ns;
      double* x3 = k + 1 < nstep ? keep + (size_t)3 * (k + 1) * ns : m->SA;
      int rc = sw_rk3<false>(m, x0, nullptr, s, x0 + ns, x0 + 2 * ns, x3);
      if (rc) return rc;
      if (forced) add_forcing(m, x3, m->F, s);
    }
    scale3(m, m->SA, m->SA + m->nu, m->SA + m->nu + m->nv, u1, v1, h1, 1, s);
    VD_HIP_TRY(hipGetLastError());
    return VD_OK;
  }
  scale3(m, u0, v0, h0, m->SA, m->SA + m->nu, m->SA + m->nu + m->nv, 0, s);
  for (int k = 0; k < nstep; ++k) {
    if (stride == 0)
      VD_HIP_TRY(hipMemcpyAsync(keep + (size_t)3 * k * ns, m->SA, ns * sizeof(double), hipMemcpyDeviceToDevice, s));
    else if (k % stride == 0)
      VD_HIP_TRY(hipMemcpyAsync(keep + (size_t)(k / stride) * ns, m->SA, ns * sizeof(double),
                                hipMemcpyDeviceToDevice, s));
    int rc = sw_rk3<false>(m, m->SA, nullptr, s);
    if (rc) return rc;
    if (stride == 0) {
      // the stage states the step left in SB (x1) and SC (x2)
      VD_HIP_TRY(hipMemcpyAsync(keep + ((size_t)3 * k + 1) * ns, m->SB, ns * sizeof(double),
                                hipMemcpyDeviceToDevice, s));
      VD_HIP_TRY(hipMemcpyAsync(keep + ((size_t)3 * k + 2) * ns, m->SC, ns * sizeof(double),
                                hipMemcpyDeviceToDevice, s));
    }
    if (forced) add_forcing(m, m->SA, m->F, s);
  }
  scale3(m, m->SA, m->SA + m->nu, m->SA + m->nu + m->nv, u1, v1, h1, 1, s);
  VD_HIP_TRY(hipGetLastError());
  return VD_OK;
}

int vd_sw_step_adj_kept(vd_sw_model* m, const double* Fu, const double* Fv, const double* Fh, const double* keep,
                        int stride, const double* au1, const double* av1, const double* ah1, double* au0,
                        double* av0, double* ah0, double* aFu, double* aFv, double* aFh, int nstep, void* stream) {
  VD_REQUIRE(m && keep && au1 && av1 && ah1 && au0 && av0 && ah0, VD_ERR_INVALID_ARG, "null argument");
  VD_REQUIRE(nstep >= 0 && stride >= 0, VD_ERR_INVALID_ARG, "sw: nstep >= 0 and stride >= 0");
  hipStream_t s = static_cast<hipStream_t>(stream);
  const bool forced = load_forcing(m, Fu, Fv, Fh, m->F, s);
  const bool want_aF = aFu || aFv || aFh;
  scale3(m, au1, av1, ah1, m->L, m->L + m->nu, m->L + m->nu + m->nv, 1, s);
  if (want_aF) scale3(m, nullptr, nullptr, nullptr, m->aF, m->aF + m->nu, m->aF + m->nu + m->nv, 0, s);
  int rc = sw_reverse(m, forced, keep, stride, nstep, want_aF, s);
  if (rc) return rc;
  scale3(m, m->L, m->L + m->nu, m->L + m->nu + m->nv, au0, av0, ah0, 2, s);
  if (want_aF) scale3(m, m->aF, m->aF + m->nu, m->aF + m->nu + m->nv, aFu, aFv, aFh, 2, s);
  VD_HIP_TRY(hipGetLastError());
  return VD_OK;
}

}  // extern "C"
