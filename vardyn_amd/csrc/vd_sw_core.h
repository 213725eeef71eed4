// Per-point arithmetic of the multilayer shallow-water model
// (mapping/models/model_qgsw/sw.py:67-937, flux.py:39-90,
// reconstruction.py:13-169, finite_diff.py), shared by the forward, tangent
// and adjoint kernels of vd_sw.hip.
//
// The functions are templates over the scalar type S: double for the forward
// model and Dn<1> (a value and one tangent) for the tangent-linear model,
// whose value part therefore runs exactly the forward arithmetic.  The
// adjoint takes the partial derivatives of a face flux with respect to its
// six stencil values from hand-written reverse sweeps of the reconstructions
// (wenoz4_grad, wenoz6_grad, face_recon_grad).
//
// Grids (x before y, y contiguous; one plane per layer):
//   u  (nx+1) x ny      v  nx x (ny+1)      h  nx x ny      w  (nx+1) x (ny+1)
// The model runs on the scaled variables u = u_phys dx, v = v_phys dy,
// h = h_phys dx dy.
#pragma once

#include <cmath>
#include <cstdint>

#ifndef VD_HD
#define VD_HD __host__ __device__
#endif

namespace vd {
namespace sw {

constexpr double kWenoEps = 1e-4;  // reconstruction.py:14

// ------------------------------------------------------------ dual numbers
template <int N>
struct Dn {
  double v;
  double d[N];
};

VD_HD inline double val(double x) { return x; }
template <int N>
VD_HD inline double val(const Dn<N>& x) { return x.v; }

template <typename S>
struct Cst;
template <>
struct Cst<double> {
  VD_HD static double make(double c) { return c; }
};
template <int N>
struct Cst<Dn<N>> {
  VD_HD static Dn<N> make(double c) {
    Dn<N> r;
    r.v = c;
    for (int k = 0; k < N; ++k) r.d[k] = 0.0;
    return r;
  }
};

template <int N>
VD_HD inline Dn<N> operator+(const Dn<N>& a, const Dn<N>& b) {
  Dn<N> r;
  r.v = a.v + b.v;
  for (int k = 0; k < N; ++k) r.d[k] = a.d[k] + b.d[k];
  return r;
}
template <int N>
VD_HD inline Dn<N> operator-(const Dn<N>& a, const Dn<N>& b) {
  Dn<N> r;
  r.v = a.v - b.v;
  for (int k = 0; k < N; ++k) r.d[k] = a.d[k] - b.d[k];
  return r;
}
template <int N>
VD_HD inline Dn<N> operator-(const Dn<N>& a) {
  Dn<N> r;
  r.v = -a.v;
  for (int k = 0; k < N; ++k) r.d[k] = -a.d[k];
  return r;
}
template <int N>
VD_HD inline Dn<N> operator*(const Dn<N>& a, const Dn<N>& b) {
  Dn<N> r;
  r.v = a.v * b.v;
  for (int k = 0; k < N; ++k) r.d[k] = a.d[k] * b.v + a.v * b.d[k];
  return r;
}
template <int N>
VD_HD inline Dn<N> operator/(const Dn<N>& a, const Dn<N>& b) {
  Dn<N> r;
  r.v = a.v / b.v;
  for (int k = 0; k < N; ++k) r.d[k] = (a.d[k] - r.v * b.d[k]) / b.v;
  return r;
}
template <int N>
VD_HD inline Dn<N> operator+(const Dn<N>& a, double c) {
  Dn<N> r = a;
  r.v = a.v + c;
  return r;
}
template <int N>
VD_HD inline Dn<N> operator+(double c, const Dn<N>& a) {
  Dn<N> r = a;
  r.v = c + a.v;
  return r;
}
template <int N>
VD_HD inline Dn<N> operator-(const Dn<N>& a, double c) {
  Dn<N> r = a;
  r.v = a.v - c;
  return r;
}
template <int N>
VD_HD inline Dn<N> operator*(double c, const Dn<N>& a) {
  Dn<N> r;
  r.v = c * a.v;
  for (int k = 0; k < N; ++k) r.d[k] = c * a.d[k];
  return r;
}
template <int N>
VD_HD inline Dn<N> operator*(const Dn<N>& a, double c) {
  Dn<N> r;
  r.v = a.v * c;
  for (int k = 0; k < N; ++k) r.d[k] = a.d[k] * c;
  return r;
}
template <int N>
VD_HD inline Dn<N> operator/(const Dn<N>& a, double c) {
  Dn<N> r;
  r.v = a.v / c;
  for (int k = 0; k < N; ++k) r.d[k] = a.d[k] / c;
  return r;
}
template <int N>
VD_HD inline Dn<N> sqrt(const Dn<N>& a) {
  Dn<N> r;
  r.v = ::sqrt(a.v);
  for (int k = 0; k < N; ++k) r.d[k] = a.d[k] / (2.0 * r.v);
  return r;
}

// softplus(z) = log(1 + e^z) as jax.nn.softplus (max(z, 0) + log1p(e^-|z|)),
// derivative the logistic function
VD_HD inline double softplus(double z) { return (z > 0 ? z : 0.0) + ::log1p(::exp(-::fabs(z))); }
VD_HD inline double logistic(double z) {
  const double e = ::exp(-::fabs(z));
  return z > 0 ? 1.0 / (1.0 + e) : e / (1.0 + e);
}
template <int N>
VD_HD inline Dn<N> softplus(const Dn<N>& a) {
  Dn<N> r;
  r.v = softplus(a.v);
  const double s = logistic(a.v);
  for (int k = 0; k < N; ++k) r.d[k] = s * a.d[k];
  return r;
}

// smooth_clamp (sw.py:25-31): x_min + softplus((x - x_min) s) / s
template <typename S>
VD_HD inline S smooth_clamp(const S& x, double xmin, double s) {
  return xmin + softplus((x - xmin) * s) / s;
}
VD_HD inline double smooth_clamp_deriv(double x, double xmin, double s) { return logistic((x - xmin) * s); }

// ------------------------------------------------------------ reconstructions
using ::sqrt;

template <typename S>
VD_HD inline S smooth_abs(const S& x) {  // reconstruction.py:15-17
  return sqrt(x * x + kWenoEps * kWenoEps);
}

// wenoz4_left(qmm, qm, qp, .) (reconstruction.py:82-104)
template <typename S>
VD_HD inline S wenoz4(const S& qmm, const S& qm, const S& qp) {
  const S qi1 = (-1. / 2.) * qmm + (3. / 2.) * qm;
  const S qi2 = (1. / 2.) * (qm + qp);
  const S b1 = (qm - qmm) * (qm - qmm);
  const S b2 = (qp - qm) * (qp - qm);
  const S tau = smooth_abs(b2 - b1);
  const S w1 = (1. / 3.) * (1. + tau / (b1 + kWenoEps));
  const S w2 = (2. / 3.) * (1. + tau / (b2 + kWenoEps));
  return (w1 * qi1 + w2 * qi2) / (w1 + w2);
}

// wenoz6_left(qmmm, qmm, qm, qp, qpp, .) (reconstruction.py:135-166)
template <typename S>
VD_HD inline S wenoz6(const S& qmmm, const S& qmm, const S& qm, const S& qp, const S& qpp) {
  const S qi1 = (1. / 3.) * qmmm - (7. / 6.) * qmm + (11. / 6.) * qm;
  const S qi2 = (-1. / 6.) * qmm + (5. / 6.) * qm + (1. / 3.) * qp;
  const S qi3 = (1. / 3.) * qm + (5. / 6.) * qp - (1. / 6.) * qpp;
  const double k1 = 13. / 12., k2 = 0.25;
  const S a1 = qmmm - 2. * qmm + qm, c1 = qmmm - 4. * qmm + 3. * qm;
  const S a2 = qmm - 2. * qm + qp, c2 = qmm - qp;
  const S a3 = qm - 2. * qp + qpp, c3 = 3. * qm - 4. * qp + qpp;
  const S b1 = k1 * (a1 * a1) + k2 * (c1 * c1);
  const S b2 = k1 * (a2 * a2) + k2 * (c2 * c2);
  const S b3 = k1 * (a3 * a3) + k2 * (c3 * c3);
  const S tau5 = smooth_abs(b1 - b3);
  const S w1 = 0.1 * (1. + tau5 / (b1 + kWenoEps));
  const S w2 = 0.6 * (1. + tau5 / (b2 + kWenoEps));
  const S w3 = 0.3 * (1. + tau5 / (b3 + kWenoEps));
  return (w1 * qi1 + w2 * qi2 + w3 * qi3) / (w1 + w2 + w3);
}

// Reverse-mode partials of the two reconstructions: return the value (the
// same expressions as wenoz4 / wenoz6 on doubles) and add its partial
// derivatives with respect to the inputs to g*.  One value, few inputs: the
// reverse sweep costs about twice the forward, against seven times for the
// six-component dual numbers.
VD_HD inline double wenoz4_grad(double qmm, double qm, double qp, double& gmm, double& gm, double& gp) {
  const double qi1 = (-1. / 2.) * qmm + (3. / 2.) * qm;
  const double qi2 = (1. / 2.) * (qm + qp);
  const double d1 = qm - qmm, d2 = qp - qm;
  const double b1 = d1 * d1, b2 = d2 * d2;
  const double x = b2 - b1;
  const double tau = ::sqrt(x * x + kWenoEps * kWenoEps);
  const double e1 = b1 + kWenoEps, e2 = b2 + kWenoEps;
  const double r1 = tau / e1, r2 = tau / e2;
  const double w1 = (1. / 3.) * (1. + r1), w2 = (2. / 3.) * (1. + r2);
  const double den = w1 + w2;
  const double q = (w1 * qi1 + w2 * qi2) / den;
  // reverse sweep, seed dq = 1
  const double gN = 1.0 / den, gD = -q / den;
  const double gr1 = (1. / 3.) * (gN * qi1 + gD), gr2 = (2. / 3.) * (gN * qi2 + gD);
  const double gqi1 = gN * w1, gqi2 = gN * w2;
  const double gtau = gr1 / e1 + gr2 / e2;
  const double gx = gtau * x / tau;
  const double gb1 = -gr1 * r1 / e1 - gx, gb2 = -gr2 * r2 / e2 + gx;
  const double gd1 = 2.0 * gb1 * d1, gd2 = 2.0 * gb2 * d2;
  gmm += (-1. / 2.) * gqi1 - gd1;
  gm += (3. / 2.) * gqi1 + (1. / 2.) * gqi2 + gd1 - gd2;
  gp += (1. / 2.) * gqi2 + gd2;
  return q;
}

VD_HD inline double wenoz6_grad(double qmmm, double qmm, double qm, double qp, double qpp, double& gmmm,
                                double& gmm, double& gm, double& gp, double& gpp) {
  const double qi1 = (1. / 3.) * qmmm - (7. / 6.) * qmm + (11. / 6.) * qm;
  const double qi2 = (-1. / 6.) * qmm + (5. / 6.) * qm + (1. / 3.) * qp;
  const double qi3 = (1. / 3.) * qm + (5. / 6.) * qp - (1. / 6.) * qpp;
  const double k1 = 13. / 12., k2 = 0.25;
  const double a1 = qmmm - 2. * qmm + qm, c1 = qmmm - 4. * qmm + 3. * qm;
  const double a2 = qmm - 2. * qm + qp, c2 = qmm - qp;
  const double a3 = qm - 2. * qp + qpp, c3 = 3. * qm - 4. * qp + qpp;
  const double b1 = k1 * (a1 * a1) + k2 * (c1 * c1);
  const double b2 = k1 * (a2 * a2) + k2 * (c2 * c2);
  const double b3 = k1 * (a3 * a3) + k2 * (c3 * c3);
  const double x = b1 - b3;
  const double tau5 = ::sqrt(x * x + kWenoEps * kWenoEps);
  const double e1 = b1 + kWenoEps, e2 = b2 + kWenoEps, e3 = b3 + kWenoEps;
  const double r1 = tau5 / e1, r2 = tau5 / e2, r3 = tau5 / e3;
  const double w1 = 0.1 * (1. + r1), w2 = 0.6 * (1. + r2), w3 = 0.3 * (1. + r3);
  const double den = w1 + w2 + w3;
  const double q = (w1 * qi1 + w2 * qi2 + w3 * qi3) / den;
  const double gN = 1.0 / den, gD = -q / den;
  const double gr1 = 0.1 * (gN * qi1 + gD), gr2 = 0.6 * (gN * qi2 + gD), gr3 = 0.3 * (gN * qi3 + gD);
  const double gqi1 = gN * w1, gqi2 = gN * w2, gqi3 = gN * w3;
  const double gtau = (gr1 / e1 + gr2 / e2) + gr3 / e3;
  const double gx = gtau * x / tau5;
  const double gb1 = -gr1 * r1 / e1 + gx, gb2 = -gr2 * r2 / e2, gb3 = -gr3 * r3 / e3 - gx;
  const double ga1 = 2.0 * k1 * gb1 * a1, gc1 = 2.0 * k2 * gb1 * c1;
  const double ga2 = 2.0 * k1 * gb2 * a2, gc2 = 2.0 * k2 * gb2 * c2;
  const double ga3 = 2.0 * k1 * gb3 * a3, gc3 = 2.0 * k2 * gb3 * c3;
  gmmm += (1. / 3.) * gqi1 + ga1 + gc1;
  gmm += (-7. / 6.) * gqi1 + (-1. / 6.) * gqi2 - 2. * ga1 - 4. * gc1 + ga2 + gc2;
  gm += (11. / 6.) * gqi1 + (5. / 6.) * gqi2 + (1. / 3.) * gqi3 + ga1 + 3. * gc1 - 2. * ga2 + ga3 + 3. * gc3;
  gp += (1. / 3.) * gqi2 + (5. / 6.) * gqi3 + ga2 - gc2 - 2. * ga3 - 4. * gc3;
  gpp += (-1. / 6.) * gqi3 + ga3 + gc3;
  return q;
}

// Stencil classes of a face: which reconstructions the flux adds (the
// reference's 0/1 masks mask_2, mask_4, mask_6, flux.py:80-86).
enum : uint8_t { kSten2 = 1, kSten4 = 2, kSten6 = 4 };

// Velocity-sign biased flux through the face between stencil points a and
// a+1 (flux.py:39-90): s[0..5] = q[a-2 .. a+3], vel the face velocity; the
// 4- and 6-point reconstructions count only where their stencil lies inside
// the n points (the reference zero-pads them, flux.py:60-77), ok4 / ok6.
// Only the reconstruction of the upwind side enters (the other side's weight,
// min(vel, 0) or max(vel, 0), is zero).
template <typename S>
VD_HD inline S face_recon(const S* s, bool pos, uint8_t code, bool ok4, bool ok6) {
  S q = Cst<S>::make(0.0);
  if (code & kSten2) q = q + 0.5 * (s[2] + s[3]);
  if ((code & kSten4) && ok4) q = q + (pos ? wenoz4(s[1], s[2], s[3]) : wenoz4(s[4], s[3], s[2]));
  if ((code & kSten6) && ok6) q = q + (pos ? wenoz6(s[0], s[1], s[2], s[3], s[4]) : wenoz6(s[5], s[4], s[3], s[2], s[1]));
  return q;
}

// face_recon on doubles with its partials: g[k] = d q / d s[k].
VD_HD inline double face_recon_grad(const double* s, bool pos, uint8_t code, bool ok4, bool ok6, double* g) {
  for (int k = 0; k < 6; ++k) g[k] = 0.0;
  double q = 0.0;
  if (code & kSten2) {
    q = q + 0.5 * (s[2] + s[3]);
    g[2] += 0.5;
    g[3] += 0.5;
  }
  if ((code & kSten4) && ok4)
    q = q + (pos ? wenoz4_grad(s[1], s[2], s[3], g[1], g[2], g[3]) : wenoz4_grad(s[4], s[3], s[2], g[4], g[3], g[2]));
  if ((code & kSten6) && ok6)
    q = q + (pos ? wenoz6_grad(s[0], s[1], s[2], s[3], s[4], g[0], g[1], g[2], g[3], g[4])
                 : wenoz6_grad(s[5], s[4], s[3], s[2], s[1], g[5], g[4], g[3], g[2], g[1]));
  return q;
}

// ------------------------------------------------------------ plan data
struct SwGrid {
  int nx, ny, nl;
  double dx, dy, area;
  double dt;
  double h_min_area, sharp;  // smooth clamp of h_ref + h at h_min dx dy
  double drag;               // linear bottom drag on the last layer
  double visc, diff;         // Laplacian viscosity of u, v and diffusion of h (0: none)
  double href[8];            // H_l dx dy
  double gp[8];              // reduced gravities
  // grid-constant fields (device pointers)
  const double* mask_u;  // (nx+1) x ny
  const double* mask_v;  // nx x (ny+1)
  const double* mask_h;  // nx x ny
  const double* wAC;     // vorticity weights on the w grid: d omega / d(-dy u)
  const double* wAB;     // d omega / d(dx v)
  const double* fstar_u;  // f dx dy on the u grid
  const double* fstar_v;  // f dx dy on the v grid
  const uint8_t* code_hx;  // h flux along x, u grid
  const uint8_t* code_hy;  // h flux along y, v grid
  const uint8_t* code_wy;  // vorticity flux along y, u grid
  const uint8_t* code_wx;  // vorticity flux along x, v grid
  const double* wind_u;    // top-layer wind tendency on the u grid, or null
  const double* wind_v;    // the same on the v grid
};

VD_HD inline int iu_(const SwGrid& g, int l, int i, int j) { return (l * (g.nx + 1) + i) * g.ny + j; }
VD_HD inline int iv_(const SwGrid& g, int l, int i, int j) { return (l * g.nx + i) * (g.ny + 1) + j; }
VD_HD inline int ih_(const SwGrid& g, int l, int i, int j) { return (l * g.nx + i) * g.ny + j; }
VD_HD inline int iw_(const SwGrid& g, int l, int i, int j) { return (l * (g.nx + 1) + i) * (g.ny + 1) + j; }

// Loads of S values: a value array and, for Dn<1>, a tangent array.
template <typename S>
struct Ld;
template <>
struct Ld<double> {
  VD_HD static double at(const double* v, const double*, int k) { return v[k]; }
  VD_HD static void put(double* v, double*, int k, double x) { v[k] = x; }
};
template <>
struct Ld<Dn<1>> {
  VD_HD static Dn<1> at(const double* v, const double* t, int k) { return Dn<1>{v[k], {t[k]}}; }
  VD_HD static void put(double* v, double* t, int k, const Dn<1>& x) {
    v[k] = x.v;
    t[k] = x.d[0];
  }
};

// A field as value / tangent pair (tangent null for the forward model).
struct Fld {
  const double* v;
  const double* t;
};
struct FldOut {
  double* v;
  double* t;
};

// Relative vorticity omega on the w grid (sw.py:653-671) at (i, j) of layer l:
// the slip boundary classes enter through wAC = w_valid + (1 - slip) cornerout
// + alpha horizontal_bound and wAB = w_valid + (1 - slip) cornerout + alpha
// vertical_bound.
template <typename S>
VD_HD inline S omega_at(const SwGrid& g, Fld u, Fld v, int l, int i, int j) {
  const S zero = Cst<S>::make(0.0);
  const S vr = i < g.nx ? Ld<S>::at(v.v, v.t, iv_(g, l, i, j)) : zero;
  const S vl = i > 0 ? Ld<S>::at(v.v, v.t, iv_(g, l, i - 1, j)) : zero;
  const S ut = j < g.ny ? Ld<S>::at(u.v, u.t, iu_(g, l, i, j)) : zero;
  const S ub = j > 0 ? Ld<S>::at(u.v, u.t, iu_(g, l, i, j - 1)) : zero;
  const int w = (i * (g.ny + 1) + j);
  return g.wAB[w] * (vr - vl) - g.wAC[w] * (ut - ub);
}

}  // namespace sw
}  // namespace vd
