// Balanced-motion wavelet basis (Basis_bm, mapping/src/basis.py:472-936):
// synthesis of SSH fields from wavelet coefficients (operg) and its transpose
// (operg_transpose), for a batch of basis times at once.
//
// A coefficient is (frequency f, time centre it, space centre P, direction,
// cos/sin); its field at time t is
//   g_f(t, it) * sqrt(2) * w(dx/DX_f) w(dy/DX_f) * cos|sin(kx dx + ky dy)
// on the points of the centre's box (|dx|, |dy| <= DX_f km, unmasked), where
// g_f is the time factor and w(x) = cos^2(pi x / 2) (basis.py:726-820).  The
// host passes the centres with their boxes (a rectilinear grid makes the
// reference's point selection an index box), the per-point lists of centres
// covering each point, and the time factors of the batch.
//
// Matrix-free: the space values are recomputed from the formula in each
// direction (no stored CSC), and every direction is a gather with a fixed
// summation order, so results are deterministic:
//   k_bm_time   C[c][q][t] = sum_it g_f(t, it) X[f, it, P, q]   (q: direction, cos/sin)
//   k_bm_synth  ssh[t][p] = sum over centres c covering p, q of G(p, c, q) C[c][q][t],
//               per 16 x 16 tile of grid points: the coefficients of the centres
//               covering the tile are staged in LDS for 16 times at a time, so
//               that every point reads them as LDS broadcasts
//   k_bm_proj   Dt[e][q][t] = sum over the points of a 16 x 16 tile in the box of
//               centre c of G(p, c, q) adssh[t][p], for every (tile, centre)
//               entry e of the tile lists (the tile's adjoint fields in LDS,
//               16 times per pass)
//   k_bm_proj_sum  D[c][q][t] = sum of c's entries in ascending tile order
//   k_bm_time_adj  adX[f, it, P, q] += sum_t g_f(t, it) D[c][q][t]
// The phase is separable on the rectilinear grid: dx depends on the column
// only and dy on the row only, so cos|sin(kx dx + ky dy) are products of
// per-column and per-row cos/sin (angle addition), and the taper is the
// product of w(dx/DX) and w(dy/DX).  k_bm_tables evaluates those once per
// centre, column or row, and direction when the basis is created; the
// operators then read them (a few loads and multiply-adds per entry instead
// of a sin/cos evaluation).
#include <cmath>
#include <cstring>

#include "vd_internal.h"

namespace vd {

constexpr double kKm2Deg = 1.0 / 110;
constexpr int kBmMaxTheta = 16;  // directions per frequency (9 with the reference defaults)
constexpr int kBmTB = 16;        // times per pass of the synthesis gather and the projection
constexpr int kBmTile = 16;      // grid-point tile side of the synthesis
constexpr int kBmCC = 12;        // centres staged in LDS at a time (48 KiB of coefficients)
constexpr int kBmParts = 4;      // lanes per (centre, direction) pair of the projection
constexpr bool kBmPartCols = true;  // those lanes take interleaved columns (else row bands)

struct BmTables {
  int ny, nx, nf, ntheta, nc;
  const double* lon1;      // nx
  const double* lat1;      // ny
  const uint8_t* mask;     // ny * nx (1 = masked) or null
  const double* dxf;       // nf: DX_f (km)
  const double* kx;        // nf * ntheta
  const double* ky;
  const int* cf;           // nc: frequency of centre c
  const int* cP;           // nc: index of c among its frequency's centres
  const double* clon;      // nc
  const double* clat;
  const double* ccos;      // cos(lat_c pi / 180)
  const int4* cbox;        // nc: i0, i1, j0, j1 (inclusive; i0 > i1 = empty)
  const int64_t* fbase;    // nf: first coefficient of frequency f
  const int* fnp;          // nf: centres of frequency f
  const int* fnt;          // nf: time centres of frequency f
  const int* gtoff;        // nf: offset of frequency f's (T x nt_f) block of time factors / T
  const int64_t* pptr;     // ny * nx + 1: centres covering point p
  const int* plist;
  // separable tables per centre: columns xoff[c] .. xoff[c] + bw - 1 and rows
  // yoff[c] .. yoff[c] + bh - 1 of the global column / row tables
  int ncol, nrow;          // table lengths
  const int* xoff;
  const int* yoff;
  double* wx;              // ncol: w(dx / DX)
  double* wy;              // nrow
  double2* cx;             // ntheta * ncol: (cos, sin)(kx dx), direction-major
  double2* cy;             // ntheta * nrow: (cos, sin)(ky dy)
  int ntx, nty;            // synthesis tiles (kBmTile x kBmTile points)
  const int* tptr;         // ntx * nty + 1: centres whose box meets tile t (ascending)
  const int* tlist;
  int ntl;                 // length of the tile lists
  const int* cptr;         // nc + 1: entries (positions in tlist) of centre c,
  const int* clist;        //   in ascending tile order
};

__device__ __forceinline__ double bm_window(double x) {
  const double c = cos(x * 0.5 * M_PI);
  return c * c;
}

// Separable tables of centre c (one workgroup per centre): the column part
// dx = (lon_j - lon_c) / km2deg * cos(lat_c), the row part
// dy = (lat_i - lat_c) / km2deg, their tapers and (cos, sin)(kx dx), (cos,
// sin)(ky dy) for every direction.
__global__ void k_bm_tables(BmTables B) {
#pragma clang fp contract(off)
  const int c = blockIdx.x, f = B.cf[c];
  const int4 bx = B.cbox[c];
  if (bx.x > bx.y) return;
  const double DX = B.dxf[f];
  const int bw = bx.w - bx.z + 1, bh = bx.y - bx.x + 1;
  for (int jx = threadIdx.x; jx < bw; jx += blockDim.x) {
    const double xx = (B.lon1[bx.z + jx] - B.clon[c]) / kKm2Deg * B.ccos[c];
    const int x = B.xoff[c] + jx;
    B.wx[x] = bm_window(xx / DX);
    for (int th = 0; th < B.ntheta; ++th) {
      double sn, cs;
      sincos(B.kx[f * B.ntheta + th] * xx, &sn, &cs);
      B.cx[(size_t)th * B.ncol + x] = make_double2(cs, sn);
    }
  }
  for (int iy = threadIdx.x; iy < bh; iy += blockDim.x) {
    const double yy = (B.lat1[bx.x + iy] - B.clat[c]) / kKm2Deg;
    const int y = B.yoff[c] + iy;
    B.wy[y] = bm_window(yy / DX);
    for (int th = 0; th < B.ntheta; ++th) {
      double sn, cs;
      sincos(B.ky[f * B.ntheta + th] * yy, &sn, &cs);
      B.cy[(size_t)th * B.nrow + y] = make_double2(cs, sn);
    }
  }
}

// sqrt(2) * taper of centre c at its box point (row iy, column jx)
__device__ __forceinline__ double bm_s2f(const BmTables& B, int x, int y) {
  return M_SQRT2 * (B.wx[x] * B.wy[y]);
}

// (cos, sin)(kx dx + ky dy) of direction th from the separable tables
__device__ __forceinline__ void bm_trig(const BmTables& B, int th, int x, int y, double* s,
                                        double* co) {
  const double2 a = B.cx[(size_t)th * B.ncol + x], b = B.cy[(size_t)th * B.nrow + y];
  *co = a.x * b.x - a.y * b.y;
  *s = a.y * b.x + a.x * b.y;
}

// C[c][q][t] = sum_it g_f(t, it) X[fbase + it * blk + P * 2 ntheta + q]
__global__ void k_bm_time(BmTables B, int T, const double* __restrict__ gt,
                          const double* __restrict__ X, double* __restrict__ C) {
  const int q2 = 2 * B.ntheta;
  const long long n = (long long)B.nc * q2;
  for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < n;
       e += (long long)gridDim.x * blockDim.x) {
    const int c = (int)(e / q2), q = (int)(e % q2);
    const int f = B.cf[c], nt = B.fnt[f];
    const long long blk = (long long)B.fnp[f] * q2;
    const double* g = gt + (long long)B.gtoff[f] * T;
    const double* x = X + B.fbase[f] + (long long)B.cP[c] * q2 + q;
    for (int t = 0; t < T; ++t) {
      double acc = 0;
      for (int it = 0; it < nt; ++it) {
        const double gv = g[t * nt + it];
        if (gv != 0) acc += gv * x[it * blk];
      }
      C[e * T + t] = acc;
    }
  }
}

// ssh[t][p] for times t0 .. t0 + TB - 1 (fewer at the end of the batch): a
// workgroup per 16 x 16 tile of points; the centres meeting the tile, in
// ascending order, are staged kBmCC at a time with their coefficients for the
// pass's times.  Each point adds the centres whose box holds it in that same
// order (so its sum is the one over its own list of centres).
__global__ __launch_bounds__(256) void k_bm_synth(BmTables B, int T, int t0,
                                                  const double* __restrict__ C,
                                                  double* __restrict__ ssh) {
  __shared__ double Cs[kBmCC * 2 * kBmMaxTheta * kBmTB];
  __shared__ int4 Bx[kBmCC];
  __shared__ int Xo[kBmCC], Yo[kBmCC];
  const int tile = blockIdx.x;
  const int ti = tile / B.ntx, tj = tile % B.ntx;
  const int i = ti * kBmTile + threadIdx.x / kBmTile, j = tj * kBmTile + threadIdx.x % kBmTile;
  const bool in = i < B.ny && j < B.nx && !(B.mask && B.mask[(size_t)i * B.nx + j]);
  const int nt = min(kBmTB, T - t0);
  const int q2 = 2 * B.ntheta;
  double acc[kBmTB] = {};
  const int e0 = B.tptr[tile], e1 = B.tptr[tile + 1];
  for (int c0 = e0; c0 < e1; c0 += kBmCC) {
    const int ncc = min(kBmCC, e1 - c0);
    __syncthreads();
    for (int e = threadIdx.x; e < ncc * q2 * kBmTB; e += 256) {
      const int cc = e / (q2 * kBmTB), r = e % (q2 * kBmTB), q = r / kBmTB, t = r % kBmTB;
      const int c = B.tlist[c0 + cc];
      Cs[(cc * q2 + q) * kBmTB + t] = t < nt ? C[((int64_t)c * q2 + q) * T + t0 + t] : 0.0;
    }
    if (threadIdx.x < ncc) {
      const int c = B.tlist[c0 + threadIdx.x];
      Bx[threadIdx.x] = B.cbox[c];
      Xo[threadIdx.x] = B.xoff[c];
      Yo[threadIdx.x] = B.yoff[c];
    }
    __syncthreads();
    if (!in) continue;
    for (int cc = 0; cc < ncc; ++cc) {
      const int4 bx = Bx[cc];
      if (i < bx.x || i > bx.y || j < bx.z || j > bx.w) continue;
      const int x = Xo[cc] + (j - bx.z), y = Yo[cc] + (i - bx.x);
      const double s2f = bm_s2f(B, x, y);
      const double* cs = Cs + cc * q2 * kBmTB;
      for (int th = 0; th < B.ntheta; ++th) {
        double sn, co;
        bm_trig(B, th, x, y, &sn, &co);
        const double vc = s2f * co, vs = s2f * sn;
#pragma unroll
        for (int t = 0; t < kBmTB; ++t)
          acc[t] += vc * cs[(2 * th) * kBmTB + t] + vs * cs[(2 * th + 1) * kBmTB + t];
      }
    }
  }
  if (i < B.ny && j < B.nx) {
    const int np = B.ny * B.nx, p = i * B.nx + j;
#pragma unroll
    for (int t = 0; t < kBmTB; ++t)
      if (t < nt) ssh[(int64_t)(t0 + t) * np + p] = acc[t];
  }
}

// Dt[e][q][t] for times t0 .. t0 + 15 of the tile's entries e: a workgroup per
// 16 x 16 tile of points holds their adjoint fields in LDS (zero on masked
// points and outside the grid).  An (entry, direction) pair is taken by PARTS
// consecutive lanes, each sweeping 16 / PARTS rows of the tile's points inside
// the centre's box in row-major order (one table lookup per point for the
// cosine and sine columns of all 16 times); the lanes' partial sums are combined
// by butterfly shuffles, a fixed tree (the same bits on every run).  With COLS
// the lanes of a pair take interleaved columns instead of row bands, so that
// they read adjacent LDS words.
template <int PARTS, bool COLS>
__global__ __launch_bounds__(256) void k_bm_proj(BmTables B, int T, int t0,
                                                 const double* __restrict__ adssh,
                                                 double* __restrict__ Dt) {
  static_assert(kBmTile % PARTS == 0 && 64 % PARTS == 0, "parts of a pair share a wave");
  constexpr int RP = kBmTile / PARTS;
  __shared__ double As[kBmTB][kBmTile * kBmTile];
  const int tile = blockIdx.x;
  const int ti = tile / B.ntx, tj = tile % B.ntx;
  const int i0 = ti * kBmTile, j0 = tj * kBmTile;
  const int np = B.ny * B.nx, nt = min(kBmTB, T - t0), q2 = 2 * B.ntheta;
  {
    const int i = i0 + threadIdx.x / kBmTile, j = j0 + threadIdx.x % kBmTile;
    const bool in = i < B.ny && j < B.nx && !(B.mask && B.mask[(size_t)i * B.nx + j]);
#pragma unroll
    for (int t = 0; t < kBmTB; ++t)
      As[t][threadIdx.x] = (in && t < nt) ? adssh[(int64_t)(t0 + t) * np + i * B.nx + j] : 0.0;
  }
  __syncthreads();
  const int e0 = B.tptr[tile], ne = B.tptr[tile + 1] - e0;
  const int total = ne * B.ntheta * PARTS;
  // the loop bound is the same for the whole workgroup, so every lane reaches
  // the shuffles
  for (int base = 0; base < total; base += 256) {
    const int w = base + threadIdx.x;
    const int pair = w / PARTS, part = w % PARTS;
    const bool valid = w < total;
    const int e = e0 + pair / B.ntheta, th = pair % B.ntheta;
    double ac[kBmTB] = {}, as[kBmTB] = {};
    if (valid) {
      const int c = B.tlist[e];
      const int4 bx = B.cbox[c];
      const int r0 = COLS ? i0 : i0 + part * RP;
      const int ia = max(bx.x, r0), ib = min(min(bx.y, r0 + (COLS ? kBmTile : RP) - 1), B.ny - 1);
      int ja = max(bx.z, j0);
      const int jb = min(min(bx.w, j0 + kBmTile - 1), B.nx - 1);
      if (COLS) ja += ((part - (ja - j0)) % PARTS + PARTS) % PARTS;
      constexpr int JS = COLS ? PARTS : 1;
      for (int i = ia; i <= ib; ++i) {
        const int y = B.yoff[c] + (i - bx.x);
        const double wy = B.wy[y];
        const double2 b = B.cy[(size_t)th * B.nrow + y];
        for (int j = ja; j <= jb; j += JS) {
          const int x = B.xoff[c] + (j - bx.z);
          const double2 a = B.cx[(size_t)th * B.ncol + x];
          const double s2f = M_SQRT2 * (B.wx[x] * wy);
          const double co = a.x * b.x - a.y * b.y, sn = a.y * b.x + a.x * b.y;
          const double vc = s2f * co, vs = s2f * sn;
          const int lp = (i - i0) * kBmTile + (j - j0);
#pragma unroll
          for (int t = 0; t < kBmTB; ++t) {
            ac[t] += vc * As[t][lp];
            as[t] += vs * As[t][lp];
          }
        }
      }
    }
#pragma unroll
    for (int off = 1; off < PARTS; off <<= 1) {
#pragma unroll
      for (int t = 0; t < kBmTB; ++t) {
        ac[t] += __shfl_xor(ac[t], off);
        as[t] += __shfl_xor(as[t], off);
      }
    }
    if (valid && part == 0) {
      double* d = Dt + ((int64_t)e * q2 + 2 * th) * kBmTB;
#pragma unroll
      for (int t = 0; t < kBmTB; ++t) {
        d[t] = ac[t];
        d[kBmTB + t] = as[t];
      }
    }
  }
}

// D[c][q][t0 + t] = sum over the entries of centre c (ascending tile order) of
// Dt[e][q][t]: eight interleaved partial sums (entries k = u mod 8) combined
// in a fixed tree, so that eight loads are in flight and the result is the
// same on every run
__global__ void k_bm_proj_sum(BmTables B, int T, int t0, const double* __restrict__ Dt,
                              double* __restrict__ D) {
  const int q2 = 2 * B.ntheta, nt = min(kBmTB, T - t0);
  const int64_t n = (int64_t)B.nc * q2 * kBmTB;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int c = (int)(e / ((int64_t)q2 * kBmTB));
    const int r = (int)(e - (int64_t)c * q2 * kBmTB), q = r / kBmTB, t = r % kBmTB;
    if (t >= nt) continue;
    double a[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    int k = B.cptr[c];
    const int kend = B.cptr[c + 1];
    for (; k + 8 <= kend; k += 8) {
#pragma unroll
      for (int u = 0; u < 8; ++u) a[u] += Dt[(int64_t)B.clist[k + u] * q2 * kBmTB + r];
    }
#pragma unroll
    for (int u = 0; u < 8; ++u)
      if (k + u < kend) a[u] += Dt[(int64_t)B.clist[k + u] * q2 * kBmTB + r];
    D[((int64_t)c * q2 + q) * T + t0 + t] = ((a[0] + a[1]) + (a[2] + a[3])) + ((a[4] + a[5]) + (a[6] + a[7]));
  }
}

// adX[fbase + it * blk + P * 2 ntheta + q] += sum_t g_f(t, it) D[c][q][t]
__global__ void k_bm_time_adj(BmTables B, int T, const double* __restrict__ gt,
                              const double* __restrict__ D, double* __restrict__ adX) {
  const int q2 = 2 * B.ntheta;
  const long long n = (long long)B.nc * q2;
  for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < n;
       e += (long long)gridDim.x * blockDim.x) {
    const int c = (int)(e / q2), q = (int)(e % q2);
    const int f = B.cf[c], nt = B.fnt[f];
    const long long blk = (long long)B.fnp[f] * q2;
    const double* g = gt + (long long)B.gtoff[f] * T;
    double* x = adX + B.fbase[f] + (long long)B.cP[c] * q2 + q;
    const double* d = D + e * T;
    for (int it = 0; it < nt; ++it) {
      double acc = 0;
      bool any = false;
      for (int t = 0; t < T; ++t) {
        const double gv = g[t * nt + it];
        if (gv != 0) {
          acc += gv * d[t];
          any = true;
        }
      }
      if (any) x[it * blk] += acc;
    }
  }
}

}  // namespace vd

using namespace vd;

struct vd_bm_basis {
  BmTables B;
  std::vector<int> cptr, clist;
  std::vector<void*> owned;
  long long nbasis;
  int gt_total;  // sum of nt_f
  int device;
  // geostrophic velocities (vd_bm_uv_tables): -g / f and dy on the u grid
  // (ny, nx+1), g / f and dx on the v grid (ny+1, nx); null until set
  const double *au = nullptr, *dyu = nullptr, *bv = nullptr, *dxv = nullptr;
};

// ------------------------------------------------------ geostrophic velocities
// Basis_bm with compute_velocities (basis.py:822-860).  The SSH is padded by
// one edge row and column (np.pad mode 'edge': padded (r, c) = ssh(max(r-1,
// 0), max(c-1, 0))), and
//   u (i, c) = (au (i, c) (P(i+1, c) - P(i, c))) / dyu (i, c),  i < ny, c <= nx
//   v (r, j) = (bv (r, j) (P(r, j+1) - P(r, j))) / dxv (r, j),  r <= ny, j < nx
// in the reference's operation order (no contraction).  The adjoint gathers,
// per padded point, the four terms the reference adds in turn (+pu below,
// -pu at its own row, +pv right, -pv at its own column, pu = (au adu) / dyu,
// pv = (bv adv) / dxv), folds the padded first row, first column and corner
// back onto the SSH's edge in the reference's order, and adds the result to
// adssh.
__global__ void k_bm_uv(int ny, int nx, int T, const double* __restrict__ au, const double* __restrict__ dyu,
                        const double* __restrict__ bv, const double* __restrict__ dxv,
                        const double* __restrict__ ssh, double* __restrict__ u, double* __restrict__ v) {
#pragma clang fp contract(off)
  const size_t nu = (size_t)ny * (nx + 1), nv = (size_t)(ny + 1) * nx, n = nu + nv;
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < n * T; e += (size_t)gridDim.x * blockDim.x) {
    const size_t t = e / n, k = e % n;
    const double* s = ssh + t * ny * nx;
    auto P = [&](int r, int c) { return s[(size_t)max(r - 1, 0) * nx + max(c - 1, 0)]; };
    if (k < nu) {
      const int i = (int)(k / (nx + 1)), c = (int)(k % (nx + 1));
      u[t * nu + k] = (au[k] * (P(i + 1, c) - P(i, c))) / dyu[k];
    } else {
      const size_t kv = k - nu;
      const int r = (int)(kv / nx), j = (int)(kv % nx);
      v[t * nv + kv] = (bv[kv] * (P(r, j + 1) - P(r, j))) / dxv[kv];
    }
  }
}

__global__ void k_bm_uv_adj(int ny, int nx, int T, const double* __restrict__ au, const double* __restrict__ dyu,
                            const double* __restrict__ bv, const double* __restrict__ dxv,
                            const double* __restrict__ adu, const double* __restrict__ adv, double* adssh) {
#pragma clang fp contract(off)
  const size_t nu = (size_t)ny * (nx + 1), nv = (size_t)(ny + 1) * nx, np_ = (size_t)ny * nx;
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < np_ * T; e += (size_t)gridDim.x * blockDim.x) {
    const size_t t = e / np_;
    const int i = (int)((e % np_) / nx), j = (int)(e % nx);
    const double* du = adu + t * nu;
    const double* dv = adv + t * nv;
    // the padded adjoint at (r, c): terms added in the reference's order
    auto Pad = [&](int r, int c) {
      double a = 0.0;
      if (r >= 1) a += (au[(size_t)(r - 1) * (nx + 1) + c] * du[(size_t)(r - 1) * (nx + 1) + c]) / dyu[(size_t)(r - 1) * (nx + 1) + c];
      if (r <= ny - 1) a += -((au[(size_t)r * (nx + 1) + c] * du[(size_t)r * (nx + 1) + c]) / dyu[(size_t)r * (nx + 1) + c]);
      if (c >= 1) a += (bv[(size_t)r * nx + c - 1] * dv[(size_t)r * nx + c - 1]) / dxv[(size_t)r * nx + c - 1];
      if (c <= nx - 1) a += -((bv[(size_t)r * nx + c] * dv[(size_t)r * nx + c]) / dxv[(size_t)r * nx + c]);
      return a;
    };
    double a = Pad(i + 1, j + 1);
    if (i == 0) a += Pad(0, j + 1);
    if (j == 0) a += Pad(i + 1, 0);
    if (i == 0 && j == 0) a += Pad(0, 0);
    const size_t o = t * np_ + (size_t)i * nx + j;
    adssh[o] = adssh[o] + a;
  }
}

template <typename V>
static int bm_upload(vd_bm_basis* b, const V* src, size_t n, const V** dst) {
  if (n == 0) {
    *dst = nullptr;
    return VD_OK;
  }
  void* d = nullptr;
  VD_HIP_TRY(hipMalloc(&d, n * sizeof(V)));
  b->owned.push_back(d);
  VD_HIP_TRY(hipMemcpy(d, src, n * sizeof(V), hipMemcpyHostToDevice));
  *dst = static_cast<const V*>(d);
  return VD_OK;
}

template <typename V>
static int bm_alloc(vd_bm_basis* b, size_t n, V** dst) {
  *dst = nullptr;
  if (n == 0) return VD_OK;
  void* d = nullptr;
  VD_HIP_TRY(hipMalloc(&d, n * sizeof(V)));
  b->owned.push_back(d);
  *dst = static_cast<V*>(d);
  return VD_OK;
}

extern "C" {

int vd_bm_create(const vd_bm_params* prm, vd_bm_basis** out) {
  VD_REQUIRE(prm && out, VD_ERR_INVALID_ARG, "null argument");
  *out = nullptr;
  const vd_bm_params& P = *prm;
  VD_REQUIRE(P.ny > 0 && P.nx > 0 && P.nf > 0 && P.nc >= 0, VD_ERR_SHAPE, "bad basis sizes");
  VD_REQUIRE(P.ntheta > 0 && P.ntheta <= kBmMaxTheta, VD_ERR_UNSUPPORTED,
             "ntheta must be 1..16 (directions per frequency)");
  VD_REQUIRE(P.lon1 && P.lat1 && P.dxf && P.kx && P.ky && P.fbase && P.fnp && P.fnt && P.pptr,
             VD_ERR_INVALID_ARG, "null table");
  VD_REQUIRE(P.nc == 0 || (P.cf && P.cP && P.clon && P.clat && P.ccos && P.cbox && P.plist),
             VD_ERR_INVALID_ARG, "null centre table");
  const size_t np = (size_t)P.ny * P.nx;
  for (int c = 0; c < P.nc; ++c) {
    const int* bx = P.cbox + 4 * c;
    VD_REQUIRE(P.cf[c] >= 0 && P.cf[c] < P.nf && P.cP[c] >= 0 && P.cP[c] < P.fnp[P.cf[c]],
               VD_ERR_INVALID_ARG, "centre frequency / index out of range");
    VD_REQUIRE(bx[0] > bx[1] || (bx[0] >= 0 && bx[1] < P.ny && bx[2] >= 0 && bx[3] < P.nx),
               VD_ERR_INVALID_ARG, "centre box outside the grid");
  }
  VD_REQUIRE(P.pptr[0] == 0, VD_ERR_INVALID_ARG, "point lists must start at 0");
  for (size_t p = 0; p < np; ++p)
    VD_REQUIRE(P.pptr[p + 1] >= P.pptr[p], VD_ERR_INVALID_ARG, "point lists must be ordered");
  for (int64_t e = 0; e < P.pptr[np]; ++e)
    VD_REQUIRE(P.plist[e] >= 0 && P.plist[e] < P.nc, VD_ERR_INVALID_ARG, "bad centre in point list");
  long long nb = 0;
  for (int f = 0; f < P.nf; ++f) {
    VD_REQUIRE(P.fnp[f] >= 0 && P.fnt[f] >= 0, VD_ERR_INVALID_ARG, "bad frequency table");
    VD_REQUIRE(P.fbase[f] == nb, VD_ERR_INVALID_ARG, "frequency blocks must be contiguous");
    nb += 2LL * P.ntheta * P.fnp[f] * P.fnt[f];
  }
  int device = 0;
  VD_HIP_TRY(hipGetDevice(&device));
  vd_bm_basis* b = new vd_bm_basis();
  b->nbasis = nb;
  b->device = device;
  BmTables& B = b->B;
  B.ny = P.ny; B.nx = P.nx; B.nf = P.nf; B.ntheta = P.ntheta; B.nc = P.nc;
  std::vector<int> gtoff(P.nf);
  int acc = 0;
  for (int f = 0; f < P.nf; ++f) {
    gtoff[f] = acc;
    acc += P.fnt[f];
  }
  b->gt_total = acc;
  std::vector<int4> box(P.nc);
  for (int c = 0; c < P.nc; ++c)
    box[c] = make_int4(P.cbox[4 * c], P.cbox[4 * c + 1], P.cbox[4 * c + 2], P.cbox[4 * c + 3]);
  std::vector<int> xoff(P.nc), yoff(P.nc);
  int ncol = 0, nrow = 0;
  for (int c = 0; c < P.nc; ++c) {
    xoff[c] = ncol;
    yoff[c] = nrow;
    if (box[c].x <= box[c].y) {
      ncol += box[c].w - box[c].z + 1;
      nrow += box[c].y - box[c].x + 1;
    }
  }
  B.ncol = ncol;
  B.nrow = nrow;
  // synthesis tiles: the centres whose box meets each 16 x 16 tile, ascending
  B.ntx = (P.nx + kBmTile - 1) / kBmTile;
  B.nty = (P.ny + kBmTile - 1) / kBmTile;
  std::vector<std::vector<int>> per((size_t)B.ntx * B.nty);
  for (int c = 0; c < P.nc; ++c) {
    if (box[c].x > box[c].y) continue;
    for (int ti = box[c].x / kBmTile; ti <= box[c].y / kBmTile; ++ti)
      for (int tj = box[c].z / kBmTile; tj <= box[c].w / kBmTile; ++tj)
        per[(size_t)ti * B.ntx + tj].push_back(c);
  }
  std::vector<int> tptr(per.size() + 1, 0), tlist;
  for (size_t t = 0; t < per.size(); ++t) {
    tlist.insert(tlist.end(), per[t].begin(), per[t].end());
    tptr[t + 1] = (int)tlist.size();
  }
  B.ntl = (int)tlist.size();
  // per centre, its entries of the tile lists in ascending tile order
  {
    std::vector<int> cnt(P.nc + 1, 0);
    for (int c : tlist) ++cnt[c + 1];
    b->cptr.assign(P.nc + 1, 0);
    for (int c = 0; c < P.nc; ++c) b->cptr[c + 1] = b->cptr[c] + cnt[c + 1];
    b->clist.assign(tlist.size(), 0);
    std::vector<int> fill(b->cptr.begin(), b->cptr.end() - 1);
    for (int e = 0; e < (int)tlist.size(); ++e) b->clist[fill[tlist[e]]++] = e;
  }
  int rc = VD_OK;
  if ((rc = bm_upload(b, P.lon1, P.nx, &B.lon1)) || (rc = bm_upload(b, P.lat1, P.ny, &B.lat1)) ||
      (rc = bm_upload(b, P.mask, P.mask ? np : 0, &B.mask)) ||
      (rc = bm_upload(b, P.dxf, P.nf, &B.dxf)) ||
      (rc = bm_upload(b, P.kx, (size_t)P.nf * P.ntheta, &B.kx)) ||
      (rc = bm_upload(b, P.ky, (size_t)P.nf * P.ntheta, &B.ky)) ||
      (rc = bm_upload(b, P.cf, P.nc, &B.cf)) || (rc = bm_upload(b, P.cP, P.nc, &B.cP)) ||
      (rc = bm_upload(b, P.clon, P.nc, &B.clon)) || (rc = bm_upload(b, P.clat, P.nc, &B.clat)) ||
      (rc = bm_upload(b, P.ccos, P.nc, &B.ccos)) ||
      (rc = bm_upload(b, (const int4*)box.data(), P.nc, &B.cbox)) ||
      (rc = bm_upload(b, P.fbase, P.nf, &B.fbase)) || (rc = bm_upload(b, P.fnp, P.nf, &B.fnp)) ||
      (rc = bm_upload(b, P.fnt, P.nf, &B.fnt)) ||
      (rc = bm_upload(b, (const int*)gtoff.data(), P.nf, &B.gtoff)) ||
      (rc = bm_upload(b, P.pptr, np + 1, &B.pptr)) ||
      (rc = bm_upload(b, P.plist, (size_t)P.pptr[np], &B.plist)) ||
      (rc = bm_upload(b, (const int*)b->cptr.data(), b->cptr.size(), &B.cptr)) ||
      (rc = bm_upload(b, (const int*)b->clist.data(), b->clist.size(), &B.clist)) ||
      (rc = bm_upload(b, (const int*)xoff.data(), xoff.size(), &B.xoff)) ||
      (rc = bm_upload(b, (const int*)tptr.data(), tptr.size(), &B.tptr)) ||
      (rc = bm_upload(b, (const int*)tlist.data(), tlist.size(), &B.tlist)) ||
      (rc = bm_upload(b, (const int*)yoff.data(), yoff.size(), &B.yoff)) ||
      (rc = bm_alloc(b, (size_t)ncol, &B.wx)) || (rc = bm_alloc(b, (size_t)nrow, &B.wy)) ||
      (rc = bm_alloc(b, (size_t)ncol * P.ntheta, &B.cx)) ||
      (rc = bm_alloc(b, (size_t)nrow * P.ntheta, &B.cy))) {
    vd_bm_destroy(b);
    return rc;
  }
  if (P.nc > 0) {
    hipLaunchKernelGGL(k_bm_tables, dim3(P.nc), dim3(256), 0, 0, B);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipDeviceSynchronize();
    if (e != hipSuccess) {
      vd_bm_destroy(b);
      VD_HIP_TRY(e);
    }
  }
  *out = b;
  return VD_OK;
}

int vd_bm_destroy(vd_bm_basis* b) {
  if (!b) return VD_OK;
  for (void* p : b->owned) hipFree(p);
  delete b;
  return VD_OK;
}

int vd_bm_info(const vd_bm_basis* b, long long* nbasis, int* gt_total) {
  VD_REQUIRE(b, VD_ERR_INVALID_ARG, "null basis");
  if (nbasis) *nbasis = b->nbasis;
  if (gt_total) *gt_total = b->gt_total;
  return VD_OK;
}

int vd_bm_workspace_size(const vd_bm_basis* b, int T, size_t* bytes) {
  VD_REQUIRE(b && bytes && T > 0, VD_ERR_INVALID_ARG, "null argument or T < 1");
  // C (or D) per centre and time, plus the projection's per-(tile, centre)
  // partial sums of one pass of kBmTB times
  *bytes = ((size_t)b->B.nc * T + (size_t)b->B.ntl * kBmTB) * 2 * b->B.ntheta * sizeof(double);
  return VD_OK;
}

int vd_bm_operg(vd_bm_basis* b, int T, const double* gt, const double* X, double* ssh,
                void* work, void* stream) {
  VD_REQUIRE(b && gt && X && ssh && work && T > 0, VD_ERR_INVALID_ARG, "null argument or T < 1");
  const BmTables& B = b->B;
  hipStream_t s = static_cast<hipStream_t>(stream);
  double* C = static_cast<double*>(work);
  const long long n = (long long)B.nc * 2 * B.ntheta;
  if (n > 0)
    hipLaunchKernelGGL(k_bm_time, dim3((unsigned)std::min<long long>((n + 255) / 256, 65535)),
                       dim3(256), 0, s, B, T, gt, X, C);
  for (int t0 = 0; t0 < T; t0 += kBmTB)
    hipLaunchKernelGGL(k_bm_synth, dim3(B.ntx * B.nty), dim3(256), 0, s, B, T, t0, C, ssh);
  VD_HIP_TRY(hipGetLastError());
  return VD_OK;
}

int vd_bm_operg_adj(vd_bm_basis* b, int T, const double* gt, const double* adssh, double* adX,
                    void* work, void* stream) {
  VD_REQUIRE(b && gt && adssh && adX && work && T > 0, VD_ERR_INVALID_ARG,
             "null argument or T < 1");
  const BmTables& B = b->B;
  hipStream_t s = static_cast<hipStream_t>(stream);
  double* D = static_cast<double*>(work);
  double* Dt = D + (size_t)B.nc * 2 * B.ntheta * T;
  if (B.nc == 0) return VD_OK;
  const int64_t nd = (int64_t)B.nc * 2 * B.ntheta * kBmTB;
  for (int t0 = 0; t0 < T; t0 += kBmTB) {
    hipLaunchKernelGGL((k_bm_proj<kBmParts, kBmPartCols>), dim3(B.ntx * B.nty), dim3(256), 0, s, B, T, t0, adssh, Dt);
    hipLaunchKernelGGL(k_bm_proj_sum, dim3((unsigned)std::min<int64_t>((nd + 255) / 256, 65535)),
                       dim3(256), 0, s, B, T, t0, Dt, D);
  }
  const long long n = (long long)B.nc * 2 * B.ntheta;
  hipLaunchKernelGGL(k_bm_time_adj, dim3((unsigned)std::min<long long>((n + 255) / 256, 65535)),
                     dim3(256), 0, s, B, T, gt, D, adX);
  VD_HIP_TRY(hipGetLastError());
  return VD_OK;
}

int vd_bm_uv_tables(vd_bm_basis* b, const double* au, const double* dyu, const double* bv, const double* dxv) {
  VD_REQUIRE(b && au && dyu && bv && dxv, VD_ERR_INVALID_ARG, "null argument");
  const int ny = b->B.ny, nx = b->B.nx;
  const size_t nu = (size_t)ny * (nx + 1), nv = (size_t)(ny + 1) * nx;
  VD_REQUIRE(b->au == nullptr, VD_ERR_INVALID_ARG, "velocity tables already set");
  int rc;
  if ((rc = bm_upload(b, au, nu, &b->au)) != VD_OK || (rc = bm_upload(b, dyu, nu, &b->dyu)) != VD_OK ||
      (rc = bm_upload(b, bv, nv, &b->bv)) != VD_OK || (rc = bm_upload(b, dxv, nv, &b->dxv)) != VD_OK)
    return rc;
  return VD_OK;
}

static unsigned bm_blocks(size_t n) { return (unsigned)std::max<size_t>(1, std::min<size_t>((n + 255) / 256, 65535)); }

int vd_bm_ssh2uv(vd_bm_basis* b, int T, const double* ssh, double* u, double* v, void* stream) {
  VD_REQUIRE(b && ssh && u && v && T > 0, VD_ERR_INVALID_ARG, "null argument or T < 1");
  VD_REQUIRE(b->au, VD_ERR_INVALID_ARG, "velocity tables not set (vd_bm_uv_tables)");
  const int ny = b->B.ny, nx = b->B.nx;
  const size_t n = ((size_t)ny * (nx + 1) + (size_t)(ny + 1) * nx) * T;
  hipLaunchKernelGGL(k_bm_uv, dim3(bm_blocks(n)), dim3(256), 0, static_cast<hipStream_t>(stream), ny, nx, T, b->au,
                     b->dyu, b->bv, b->dxv, ssh, u, v);
  VD_HIP_TRY(hipGetLastError());
  return VD_OK;
}

int vd_bm_ssh2uv_adj(vd_bm_basis* b, int T, const double* adu, const double* adv, double* adssh, void* stream) {
  VD_REQUIRE(b && adu && adv && adssh && T > 0, VD_ERR_INVALID_ARG, "null argument or T < 1");
  VD_REQUIRE(b->au, VD_ERR_INVALID_ARG, "velocity tables not set (vd_bm_uv_tables)");
  const int ny = b->B.ny, nx = b->B.nx;
  const size_t n = (size_t)ny * nx * T;
  hipLaunchKernelGGL(k_bm_uv_adj, dim3(bm_blocks(n)), dim3(256), 0, static_cast<hipStream_t>(stream), ny, nx, T,
                     b->au, b->dyu, b->bv, b->dxv, adu, adv, adssh);
  VD_HIP_TRY(hipGetLastError());
  return VD_OK;
}

}  // extern "C"
