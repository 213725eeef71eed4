// QG model kernels for gfx950: Helmholtz inversion as sine-matrix GEMMs on
// MFMA, fused RK stage stencils (forward, tangent-linear, adjoint).
//
// Reference algorithm: mapping/models/model_qg1l/jqgm.py (datlas-ocean/VarDyn)
//   h2uv :238-262, h2pv :264-293, pv2h :295-319, rhs :321-375, adv :377-396,
//   inverse_elliptic_dst :83-91 (dstI2D :75-81).
#include "vd_internal.h"

#include <cmath>

namespace vd {

typedef float floatx4 __attribute__((ext_vector_type(4)));
typedef double doublex4 __attribute__((ext_vector_type(4)));

// 16x16x4 MFMA tiles.  Inputs, both precisions: lane l supplies A[l%16][l/16]
// and B[l/16][l%16].  Result register r of lane l holds column l%16 and row
// out_row(l, r), which differs between the two instructions (measured on
// gfx950: the f64 form interleaves the four registers across row groups).
template <typename T> struct Mfma;
template <> struct Mfma<float> {
  typedef floatx4 acc_t;
  static __device__ __forceinline__ acc_t run(float a, float b, acc_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int out_row(int lane, int r) { return 4 * (lane >> 4) + r; }
};
template <> struct Mfma<double> {
  typedef doublex4 acc_t;
  static __device__ __forceinline__ acc_t run(double a, double b, acc_t c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int out_row(int lane, int r) { return 4 * r + (lane >> 4); }
};

// ---------------------------------------------------------------------------
// Sine-matrix GEMMs of the Helmholtz inversion.
//
// C = A * B on 32x32 output tiles.  A workgroup holds KS groups of four
// waves; the four waves of a group own the four 16x16 quadrants of the tile,
// and group g accumulates the K chunks g, g+KS, g+2KS, ...  (32 wide, staged
// through LDS with a one-chunk register prefetch).  The KS partial tiles are
// summed through LDS in a fixed order, so results are deterministic.  More
// groups give the CU more waves to hide L2 latency with; the product stays
// one launch.
//
// Operands: A dense (AOP_DENSE) or column-folded (AOP_FOLDCOL); B dense,
// the interior of a field minus an optional second field (BOP_FIELD), or
// row-folded (BOP_FOLDROW_*).  All dense operands are zero-padded to
// multiples of 32, so the K loop has no bounds checks.
//
// Folding: for an even interior size m, S[j][m-1-k] = (-1)^j S[j][k] and S
// is symmetric, so with h = m/2 and p = blockIdx.z in {0, 1}
//   (S X)[2r+p][:] = sum_{k<h} S[2r+p][k] (X[k][:] + (-1)^p X[m-1-k][:])
//   (X S)[:][2c+p] = sum_{l<h} (X[:][l] + (-1)^p X[:][m-1-l]) S[l][2c+p]
// and each m x m sine product becomes two h x h products; the loaders form
// the sums/differences and the epilogues interleave outputs back into
// natural order.
// ---------------------------------------------------------------------------
enum { AOP_DENSE = 0, AOP_FOLDCOL = 1 };
enum { BOP_DENSE = 0, BOP_FIELD = 1, BOP_FOLDROW_FIELD = 2, BOP_FOLDROW_DENSE = 3 };
enum { EK_STORE = 0, EK_SCALE = 1, EK_FIELD_SET = 2, EK_FIELD_ACC = 3 };
enum { MAP_PLAIN = 0, MAP_ROWFOLD = 1, MAP_COLFOLD = 2 };

template <typename T>
struct GemmArgs {
  const T* A;
  int lda;
  size_t strideA;  // block of parity p at A + p * strideA (AOP_DENSE)
  const T* B;
  int ldb;
  size_t strideB;  // block of parity p at B + p * strideB (BOP_DENSE)
  const T* F;      // field operand (BOP_FIELD, BOP_FOLDROW_FIELD)
  const T* Fsub;   // optional subtrahend field
  int ny, nx, my, mx, hy, hx;
  int K;
  T* C;
  int ldc;
  const T* scale;  // EK_SCALE factor, ld = ldc
  const T* hb;     // EK_FIELD_SET offset / ring field (nullptr = 0)
  T* out;          // EK_FIELD_*: (ny, nx) destination
};

template <typename T, int AOP, int BOP, int KIND, int MAP, int KS>
__global__ __launch_bounds__(256 * KS) void k_gemm(GemmArgs<T> g) {
  typedef typename Mfma<T>::acc_t acc_t;
  constexpr int LDT = kTile + 1;
  __shared__ T sm[KS][2][kTile][LDT];
  const int tid = threadIdx.x;
  const int grp = tid >> 8;          // K group
  const int gt = tid & 255;          // thread within the group
  const int lane = tid & 63;
  const int wq = (tid >> 6) & 3;     // quadrant of the 32x32 tile
  const int wm = wq >> 1, wn = wq & 1;
  const int par = blockIdx.z;
  const T sgn = par ? T(-1) : T(1);
  const int m0 = blockIdx.y * kTile, n0 = blockIdx.x * kTile;
  const int lr = gt >> 3, lc = (gt & 7) * 4;
  const T* A = g.A + (AOP == AOP_DENSE ? par * g.strideA : 0);
  const T* B = g.B + (BOP == BOP_DENSE ? par * g.strideB : 0);
  T(*As)[LDT] = sm[grp][0];
  T(*Bs)[LDT] = sm[grp][1];

  T ra[4], rb[4];
  auto load = [&](int k0) {
    const int i = m0 + lr;
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int l = k0 + lc + u;
      if (AOP == AOP_DENSE) {
        ra[u] = A[(size_t)i * g.lda + l];
      } else {
        ra[u] = l < g.hx ? A[(size_t)i * g.lda + l] + sgn * A[(size_t)i * g.lda + (g.mx - 1 - l)]
                         : T(0);
      }
    }
    const int k = k0 + lr;
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int n = n0 + lc + u;
      if (BOP == BOP_DENSE) {
        rb[u] = B[(size_t)k * g.ldb + n];
      } else if (BOP == BOP_FOLDROW_DENSE) {
        rb[u] = k < g.hy ? B[(size_t)k * g.ldb + n] + sgn * B[(size_t)(g.my - 1 - k) * g.ldb + n]
                         : T(0);
      } else if (BOP == BOP_FIELD) {
        T v = T(0);
        if (k < g.my && n < g.mx) {
          const size_t idx = (size_t)(k + 1) * g.nx + (n + 1);
          v = g.F[idx];
          if (g.Fsub) v -= g.Fsub[idx];
        }
        rb[u] = v;
      } else {
        T v = T(0);
        if (k < g.hy && n < g.mx) {
          const size_t i1 = (size_t)(k + 1) * g.nx + (n + 1);
          const size_t i2 = (size_t)(g.my - k) * g.nx + (n + 1);
          T a = g.F[i1], b = g.F[i2];
          if (g.Fsub) { a -= g.Fsub[i1]; b -= g.Fsub[i2]; }
          v = a + sgn * b;
        }
        rb[u] = v;
      }
    }
  };

  acc_t acc4[4];
#pragma unroll
  for (int c = 0; c < 4; ++c)
#pragma unroll
    for (int r = 0; r < 4; ++r) acc4[c][r] = T(0);

  // every group runs the same number of chunks (K is a multiple of 32*KS)
  const int nchunk = g.K / (kTile * KS);
  load(grp * kTile);
  for (int it = 0; it < nchunk; ++it) {
    __syncthreads();
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      As[lr][lc + u] = ra[u];
      Bs[lr][lc + u] = rb[u];
    }
    __syncthreads();
    if (it + 1 < nchunk) load(((it + 1) * KS + grp) * kTile);
#pragma unroll
    for (int kk = 0; kk < kTile / 4; ++kk) {
      const T a = As[wm * 16 + (lane & 15)][kk * 4 + (lane >> 4)];
      const T b = Bs[kk * 4 + (lane >> 4)][wn * 16 + (lane & 15)];
      acc4[kk & 3] = Mfma<T>::run(a, b, acc4[kk & 3]);
    }
  }
  T val[4];
#pragma unroll
  for (int r = 0; r < 4; ++r) val[r] = (acc4[0][r] + acc4[1][r]) + (acc4[2][r] + acc4[3][r]);

  if (KS > 1) {
    // groups 1..KS-1 park their partial tiles in their (now idle) LDS slots
    __syncthreads();
    T* park = &sm[grp][0][0][0];
    if (grp > 0) {
#pragma unroll
      for (int r = 0; r < 4; ++r) park[(wq * 64 + lane) * 4 + r] = val[r];
    }
    __syncthreads();
    if (grp != 0) return;
#pragma unroll
    for (int q = 1; q < KS; ++q) {
      const T* src = &sm[q][0][0][0];
#pragma unroll
      for (int r = 0; r < 4; ++r) val[r] += src[(wq * 64 + lane) * 4 + r];
    }
  }

  const int col = n0 + wn * 16 + (lane & 15);
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const int row = m0 + wm * 16 + Mfma<T>::out_row(lane, r);
    const T v = val[r];
    if (MAP == MAP_ROWFOLD) {
      if (row < g.hy) g.C[(size_t)(2 * row + par) * g.ldc + col] = v;
      continue;
    }
    const int oc = MAP == MAP_COLFOLD ? 2 * col + par : col;
    if (KIND == EK_STORE || KIND == EK_SCALE) {
      if (MAP == MAP_COLFOLD && col >= g.hx) continue;
      const size_t o = (size_t)row * g.ldc + oc;
      g.C[o] = KIND == EK_SCALE ? v * g.scale[o] : v;
      continue;
    }
    const bool ok = row < g.my && (MAP == MAP_COLFOLD ? col < g.hx : col < g.mx);
    if (!ok) continue;
    const int nx = g.nx, ny = g.ny;
    const size_t i = row + 1, j = oc + 1;
    if (KIND == EK_FIELD_ACC) {
      g.out[i * nx + j] += v;
    } else {
      const T* hb = g.hb;
      g.out[i * nx + j] = (hb ? hb[i * nx + j] : T(0)) + v;
      // ring: h = hb on rows 0, ny-1 and columns 0, nx-1 (jqgm.py:316-317)
      auto ring = [&](size_t ii, size_t jj) { g.out[ii * nx + jj] = hb ? hb[ii * nx + jj] : T(0); };
      const bool top = row == 0, bot = row == g.my - 1, lft = oc == 0, rgt = oc == g.mx - 1;
      if (top) ring(0, j);
      if (bot) ring(ny - 1, j);
      if (lft) ring(i, 0);
      if (rgt) ring(i, nx - 1);
      if (top && lft) ring(0, 0);
      if (top && rgt) ring(0, nx - 1);
      if (bot && lft) ring(ny - 1, 0);
      if (bot && rgt) ring(ny - 1, nx - 1);
    }
  }
}

template <typename T, int AOP, int BOP, int KIND, int MAP>
static void gemm(const GemmArgs<T>& a, int M, int N, int nz, int ksplit, hipStream_t s) {
  const int nchunk = a.K / kTile;
  int ks = 1;
  if (ksplit >= 4 && nchunk % 4 == 0) ks = 4;
  else if (ksplit >= 2 && nchunk % 2 == 0) ks = 2;
  dim3 grid(N / kTile, M / kTile, nz);
  if (ks == 4)
    hipLaunchKernelGGL((k_gemm<T, AOP, BOP, KIND, MAP, 4>), grid, dim3(1024), 0, s, a);
  else if (ks == 2)
    hipLaunchKernelGGL((k_gemm<T, AOP, BOP, KIND, MAP, 2>), grid, dim3(512), 0, s, a);
  else
    hipLaunchKernelGGL((k_gemm<T, AOP, BOP, KIND, MAP, 1>), grid, dim3(256), 0, s, a);
}

template <typename T>
void launch_inversion(const Plan& p, const T* F, const T* Fsub, const T* hb, T* out, int mode,
                      hipStream_t s) {
  T* T1 = static_cast<T*>(p.gT[0]);
  T* T2 = static_cast<T*>(p.gT[1]);
  const int ks = p.gemm_ksplit;
  GemmArgs<T> a{};
  a.ny = p.ny; a.nx = p.nx; a.my = p.my; a.mx = p.mx; a.hy = p.hy; a.hx = p.hx;
  a.scale = static_cast<const T*>(p.lam_inv);
  if (p.fold) {
    const T* SyF = static_cast<const T*>(p.SyF);
    const T* SxF = static_cast<const T*>(p.SxF);
    a.strideA = (size_t)p.HY * p.HY;
    a.strideB = (size_t)p.HX * p.HX;
    // T1 = S_y (F - Fsub)
    a.A = SyF; a.lda = p.HY; a.F = F; a.Fsub = Fsub; a.K = p.HY; a.C = T1; a.ldc = p.MX;
    gemm<T, AOP_DENSE, BOP_FOLDROW_FIELD, EK_STORE, MAP_ROWFOLD>(a, p.HY, p.MX, 2, ks, s);
    // T2 = (T1 S_x) ./ Lambda
    a.A = T1; a.lda = p.MX; a.B = SxF; a.ldb = p.HX; a.K = p.HX; a.C = T2; a.ldc = p.MX;
    gemm<T, AOP_FOLDCOL, BOP_DENSE, EK_SCALE, MAP_COLFOLD>(a, p.MY, p.HX, 2, ks, s);
    // T1 = S_y T2
    a.A = SyF; a.lda = p.HY; a.B = T2; a.ldb = p.MX; a.K = p.HY; a.C = T1; a.ldc = p.MX;
    gemm<T, AOP_DENSE, BOP_FOLDROW_DENSE, EK_STORE, MAP_ROWFOLD>(a, p.HY, p.MX, 2, ks, s);
    // out = epilogue(T1 S_x)
    a.A = T1; a.lda = p.MX; a.B = SxF; a.ldb = p.HX; a.K = p.HX; a.out = out; a.hb = hb;
    if (mode == INV_ACC)
      gemm<T, AOP_FOLDCOL, BOP_DENSE, EK_FIELD_ACC, MAP_COLFOLD>(a, p.MY, p.HX, 2, ks, s);
    else
      gemm<T, AOP_FOLDCOL, BOP_DENSE, EK_FIELD_SET, MAP_COLFOLD>(a, p.MY, p.HX, 2, ks, s);
    return;
  }
  const T* Sy = static_cast<const T*>(p.Sy);
  const T* Sx = static_cast<const T*>(p.Sx);
  // T1 = S_y (F - Fsub)_interior
  a.A = Sy; a.lda = p.MY; a.F = F; a.Fsub = Fsub; a.K = p.MY; a.C = T1; a.ldc = p.MX;
  gemm<T, AOP_DENSE, BOP_FIELD, EK_STORE, MAP_PLAIN>(a, p.MY, p.MX, 1, ks, s);
  // T2 = (T1 S_x) ./ Lambda
  a.A = T1; a.lda = p.MX; a.B = Sx; a.ldb = p.MX; a.K = p.MX; a.C = T2; a.ldc = p.MX;
  gemm<T, AOP_DENSE, BOP_DENSE, EK_SCALE, MAP_PLAIN>(a, p.MY, p.MX, 1, ks, s);
  // T1 = S_y T2
  a.A = Sy; a.lda = p.MY; a.B = T2; a.ldb = p.MX; a.K = p.MY; a.C = T1; a.ldc = p.MX;
  gemm<T, AOP_DENSE, BOP_DENSE, EK_STORE, MAP_PLAIN>(a, p.MY, p.MX, 1, ks, s);
  // out = epilogue(T1 S_x)
  a.A = T1; a.lda = p.MX; a.B = Sx; a.ldb = p.MX; a.K = p.MX; a.out = out; a.hb = hb;
  if (mode == INV_ACC)
    gemm<T, AOP_DENSE, BOP_DENSE, EK_FIELD_ACC, MAP_PLAIN>(a, p.MY, p.MX, 1, ks, s);
  else
    gemm<T, AOP_DENSE, BOP_DENSE, EK_FIELD_SET, MAP_PLAIN>(a, p.MY, p.MX, 1, ks, s);
}

// ---------------------------------------------------------------------------
// Stencil helpers.  Point (i, j) is row i (y), column j (x).
// ---------------------------------------------------------------------------
template <typename T>
__device__ __forceinline__ T nz(T v) { return isnan(v) ? T(0) : v; }

// u[i][j] = cu (h[i+1][j-1] + h[i+1][j] - h[i-1][j] - h[i-1][j-1]), NaN -> 0
template <typename T>
__device__ __forceinline__ T u_at(const T* h, int nx, int i, int j, T cu) {
  return nz(cu * (h[(i + 1) * nx + j - 1] + h[(i + 1) * nx + j] - h[(i - 1) * nx + j] -
                  h[(i - 1) * nx + j - 1]));
}
// v[i][j] = cv (h[i][j+1] + h[i-1][j+1] - h[i-1][j-1] - h[i][j-1]), NaN -> 0
template <typename T>
__device__ __forceinline__ T v_at(const T* h, int nx, int i, int j, T cv) {
  return nz(cv * (h[i * nx + j + 1] + h[(i - 1) * nx + j + 1] - h[(i - 1) * nx + j - 1] -
                  h[i * nx + j - 1]));
}
template <typename T>
__device__ __forceinline__ bool u_isnan(const T* h, int nx, int i, int j) {
  return isnan(h[(i + 1) * nx + j - 1] + h[(i + 1) * nx + j] + h[(i - 1) * nx + j] +
               h[(i - 1) * nx + j - 1]);
}
template <typename T>
__device__ __forceinline__ bool v_isnan(const T* h, int nx, int i, int j) {
  return isnan(h[i * nx + j + 1] + h[(i - 1) * nx + j + 1] + h[(i - 1) * nx + j - 1] +
               h[i * nx + j - 1]);
}

// Upwind difference sums of adv (jqgm.py:387-394), without the 1/(6 d) factor.
template <typename T>
struct Diffs {
  T xp, xm, yp, ym;
};
template <typename T>
__device__ __forceinline__ Diffs<T> diffs(const T* q, int nx, int i, int j) {
  const T c = q[i * nx + j];
  const T e1 = q[i * nx + j + 1], e2 = q[i * nx + j + 2];
  const T w1 = q[i * nx + j - 1], w2 = q[i * nx + j - 2];
  const T n1 = q[(i + 1) * nx + j], n2 = q[(i + 2) * nx + j];
  const T s1 = q[(i - 1) * nx + j], s2 = q[(i - 2) * nx + j];
  Diffs<T> d;
  d.xp = 2 * e1 + 3 * c - 6 * w1 + w2;
  d.xm = e2 - 6 * e1 + 3 * c + 2 * w1;
  d.yp = 2 * n1 + 3 * c - 6 * s1 + s2;
  d.ym = n2 - 6 * n1 + 3 * c + 2 * s1;
  return d;
}

template <typename T>
struct Vel {
  T uT, vT;
};
template <typename T>
__device__ __forceinline__ Vel<T> vel_T(const T* h, int nx, int i, int j, const Phys<T>& P) {
  Vel<T> r;
  r.uT = T(0.5) * (u_at(h, nx, i, j, P.cu) + u_at(h, nx, i, j + 1, P.cu));
  r.vT = T(0.5) * (v_at(h, nx, i, j, P.cv) + v_at(h, nx, i + 1, j, P.cv));
  return r;
}

// rhs (jqgm.py:321-375) at an active point (mask == 3, so 2 <= i,j <= n-4).
template <typename T>
__device__ __forceinline__ T rhs_at(const T* h, const T* q, int i, int j, const Phys<T>& P) {
  const int nx = P.nx;
  const Vel<T> V = vel_T(h, nx, i, j, P);
  const T up = V.uT < 0 ? T(0) : V.uT, um = V.uT > 0 ? T(0) : V.uT;
  const T vp = V.vT < 0 ? T(0) : V.vT, vm = V.vT > 0 ? T(0) : V.vT;
  const Diffs<T> D = diffs(q, nx, i, j);
  T k = (um * D.xm - up * D.xp) * P.inv6dx + (vm * D.ym - vp * D.yp) * P.inv6dy;
  if (P.fbathy) {
    const Diffs<T> B = diffs(P.fbathy, nx, i, j);
    k += (um * B.xm - up * B.xp) * P.inv6dx + (vm * B.ym - vp * B.yp) * P.inv6dy;
  }
  if (P.beta) k -= P.beta[i * nx + j] * V.vT;
  if (P.has_kdiff) {
    const T c = q[i * nx + j];
    k += P.kx * (q[i * nx + j + 1] + q[i * nx + j - 1] - 2 * c) +
         P.ky * (q[(i + 1) * nx + j] + q[(i - 1) * nx + j] - 2 * c);
  }
  return nz(k);
}

#define VD_POINT_2D                                    \
  const int j = blockIdx.x * blockDim.x + threadIdx.x; \
  const int i = blockIdx.y * blockDim.y + threadIdx.y; \
  if (i >= P.ny || j >= P.nx) return;                  \
  const int nx = P.nx;                                 \
  const size_t o = (size_t)i * nx + j;

static inline dim3 grid2d(int ny, int nx) { return dim3((nx + 63) / 64, (ny + 3) / 4); }
static const dim3 kBlock2d(64, 4);

// ---------------------------------------------------------------------------
// Forward kernels
// ---------------------------------------------------------------------------
// h2pv (jqgm.py:264-293) at point (i, j) of field h with ring values from hb.
template <typename T>
__device__ __forceinline__ T h2pv_at(const T* h, const T* hb, int i, int j, uint8_t m,
                                     const Phys<T>& P) {
  const int nx = P.nx;
  if (m == 0) return T(0);
  if (m != 3) return -P.gfc2 * hb[i * nx + j];
  const T c = h[i * nx + j];
  const T lap = P.gf_dy2 * (h[(i + 1) * nx + j] + h[(i - 1) * nx + j] - 2 * c) +
                P.gf_dx2 * (h[i * nx + j + 1] + h[i * nx + j - 1] - 2 * c);
  return nz(lap - P.gfc2 * c);
}

template <typename T>
__global__ void k_prep(Phys<T> P, const T* __restrict__ h0, const T* __restrict__ hb,
                       const T* __restrict__ mdt, T* __restrict__ hcur, T* __restrict__ hbm,
                       T* __restrict__ q0, T* __restrict__ qb) {
  VD_POINT_2D
  const uint8_t m = P.mask[o];
  if (mdt) {
    // q needs h + mdt at neighbours: evaluate the stencil on the fly.
    auto H = [&](int ii, int jj) { return h0[ii * nx + jj] + mdt[ii * nx + jj]; };
    auto HB = [&](int ii, int jj) { return hb[ii * nx + jj] + mdt[ii * nx + jj]; };
    hcur[o] = H(i, j);
    hbm[o] = HB(i, j);
    T qv = 0, qbv = 0;
    if (m == 3) {
      const T c = H(i, j), cb = HB(i, j);
      qv = nz(P.gf_dy2 * (H(i + 1, j) + H(i - 1, j) - 2 * c) +
              P.gf_dx2 * (H(i, j + 1) + H(i, j - 1) - 2 * c) - P.gfc2 * c);
      qbv = nz(P.gf_dy2 * (HB(i + 1, j) + HB(i - 1, j) - 2 * cb) +
               P.gf_dx2 * (HB(i, j + 1) + HB(i, j - 1) - 2 * cb) - P.gfc2 * cb);
    } else if (m != 0) {
      qv = -P.gfc2 * HB(i, j);
      qbv = qv;
    }
    q0[o] = qv;
    qb[o] = qbv;
  } else {
    hcur[o] = h0[o];
    hbm[o] = hb[o];
    q0[o] = h2pv_at(h0, hb, i, j, m, P);
    qb[o] = h2pv_at(hb, hb, i, j, m, P);
  }
}

template <typename T>
__global__ void k_h2pv(Phys<T> P, const T* __restrict__ h, const T* __restrict__ hb,
                       T* __restrict__ q) {
  VD_POINT_2D
  q[o] = h2pv_at(h, hb, i, j, P.mask[o], P);
}

template <typename T>
__global__ void k_fill_land(Phys<T> P, T* __restrict__ f, T value) {
  VD_POINT_2D
  if (P.mask[o] == 0) f[o] = value;
}

template <typename T>
__global__ void k_finish(Phys<T> P, const T* __restrict__ h, const T* __restrict__ mdt,
                         T* __restrict__ h1) {
  VD_POINT_2D
  h1[o] = P.mask[o] == 0 ? T(NAN) : (mdt ? h[o] - mdt[o] : h[o]);
}

template <typename T>
__global__ void k_stage(Phys<T> P, const T* __restrict__ h, const T* __restrict__ q,
                        const T* __restrict__ qbase, const T* __restrict__ k0,
                        const T* __restrict__ k1, T c0, T c1, T c2, T* __restrict__ qout,
                        T* __restrict__ kstore) {
  VD_POINT_2D
  T k = T(0);
  if (P.mask[o] == 3) k = rhs_at(h, q, i, j, P);
  if (kstore) kstore[o] = k;
  T r = qbase[o];
  if (k0) r += c0 * k0[o];
  if (k1) r += c1 * k1[o];
  qout[o] = r + c2 * k;
}

// ---------------------------------------------------------------------------
// Tangent-linear kernels
// ---------------------------------------------------------------------------
template <typename T>
__device__ __forceinline__ T du_at(const T* h, const T* dh, int nx, int i, int j, T cu) {
  if (u_isnan(h, nx, i, j)) return T(0);
  return cu * (dh[(i + 1) * nx + j - 1] + dh[(i + 1) * nx + j] - dh[(i - 1) * nx + j] -
               dh[(i - 1) * nx + j - 1]);
}
template <typename T>
__device__ __forceinline__ T dv_at(const T* h, const T* dh, int nx, int i, int j, T cv) {
  if (v_isnan(h, nx, i, j)) return T(0);
  return cv * (dh[i * nx + j + 1] + dh[(i - 1) * nx + j + 1] - dh[(i - 1) * nx + j - 1] -
               dh[i * nx + j - 1]);
}

template <typename T>
__device__ __forceinline__ T rhs_tl_at(const T* h, const T* q, const T* dh, const T* dq, int i,
                                       int j, const Phys<T>& P) {
  const int nx = P.nx;
  const Vel<T> V = vel_T(h, nx, i, j, P);
  const T up = V.uT < 0 ? T(0) : V.uT, um = V.uT > 0 ? T(0) : V.uT;
  const T vp = V.vT < 0 ? T(0) : V.vT, vm = V.vT > 0 ? T(0) : V.vT;
  const T duT = T(0.5) * (du_at(h, dh, nx, i, j, P.cu) + du_at(h, dh, nx, i, j + 1, P.cu));
  const T dvT = T(0.5) * (dv_at(h, dh, nx, i, j, P.cv) + dv_at(h, dh, nx, i + 1, j, P.cv));
  const T dup = V.uT < 0 ? T(0) : duT, dum = V.uT > 0 ? T(0) : duT;
  const T dvp = V.vT < 0 ? T(0) : dvT, dvm = V.vT > 0 ? T(0) : dvT;
  const Diffs<T> D = diffs(q, nx, i, j);
  const Diffs<T> dD = diffs(dq, nx, i, j);
  T k = (dum * D.xm - dup * D.xp + um * dD.xm - up * dD.xp) * P.inv6dx +
        (dvm * D.ym - dvp * D.yp + vm * dD.ym - vp * dD.yp) * P.inv6dy;
  if (P.fbathy) {
    const Diffs<T> B = diffs(P.fbathy, nx, i, j);
    k += (dum * B.xm - dup * B.xp) * P.inv6dx + (dvm * B.ym - dvp * B.yp) * P.inv6dy;
  }
  if (P.beta) k -= P.beta[i * nx + j] * dvT;
  if (P.has_kdiff) {
    const T c = dq[i * nx + j];
    k += P.kx * (dq[i * nx + j + 1] + dq[i * nx + j - 1] - 2 * c) +
         P.ky * (dq[(i + 1) * nx + j] + dq[(i - 1) * nx + j] - 2 * c);
  }
  return k;
}

template <typename T>
__global__ void k_stage_tl(Phys<T> P, const T* __restrict__ h, const T* __restrict__ q,
                           const T* __restrict__ dh, const T* __restrict__ dq,
                           const T* __restrict__ dqbase, const T* __restrict__ dk0,
                           const T* __restrict__ dk1, T c0, T c1, T c2, T* __restrict__ dqout,
                           T* __restrict__ dkstore) {
  VD_POINT_2D
  T k = T(0);
  if (P.mask[o] == 3) k = rhs_tl_at(h, q, dh, dq, i, j, P);
  if (dkstore) dkstore[o] = k;
  T r = dqbase[o];
  if (dk0) r += c0 * dk0[o];
  if (dk1) r += c1 * dk1[o];
  dqout[o] = r + c2 * k;
}

template <typename T>
__device__ __forceinline__ bool q_isnan(const T* h, int nx, int i, int j) {
  return isnan(h[(i + 1) * nx + j] + h[(i - 1) * nx + j] + h[i * nx + j + 1] +
               h[i * nx + j - 1] + h[i * nx + j]);
}

template <typename T>
__global__ void k_prep_tl(Phys<T> P, const T* __restrict__ dh0, const T* __restrict__ hcur,
                          T* __restrict__ dhcur, T* __restrict__ dq0) {
  VD_POINT_2D
  dhcur[o] = dh0[o];
  T r = T(0);
  if (P.mask[o] == 3 && !q_isnan(hcur, nx, i, j)) {
    const T c = dh0[o];
    r = P.gf_dy2 * (dh0[o + nx] + dh0[o - nx] - 2 * c) +
        P.gf_dx2 * (dh0[o + 1] + dh0[o - 1] - 2 * c) - P.gfc2 * c;
  }
  dq0[o] = r;
}

template <typename T>
__global__ void k_finish_tl(Phys<T> P, const T* __restrict__ dh, T* __restrict__ dh1) {
  VD_POINT_2D
  dh1[o] = P.mask[o] == 0 ? T(0) : dh[o];
}

// ---------------------------------------------------------------------------
// Adjoint kernels (deterministic two-pass gather, no atomics).
// Pass A, per T point t (active only where mask == 3):
//   ak = sum_s a_s src_s[t]
//   W0 = d rhs/d uT contribution (lambda_uT), W1 = lambda_vT (incl. beta),
//   W2..W5 = up ak/(6dx), um ak/(6dx), vp ak/(6dy), vm ak/(6dy), W6 = ak.
// Pass B, per point p: lambda_q(p) from W2..W6 of the T points whose stencil
//   reads q(p); lambda_h(p) from lambda_u / lambda_v at the velocity points
//   that read h(p).
// ---------------------------------------------------------------------------
template <typename T>
__global__ void k_adj_a(Phys<T> P, const T* __restrict__ h, const T* __restrict__ q,
                        const T* __restrict__ s0, const T* __restrict__ s1,
                        const T* __restrict__ s2, T a0, T a1, T a2, T* __restrict__ W0,
                        T* __restrict__ W1, T* __restrict__ W2, T* __restrict__ W3,
                        T* __restrict__ W4, T* __restrict__ W5, T* __restrict__ W6) {
  VD_POINT_2D
  T w0 = 0, w1 = 0, w2 = 0, w3 = 0, w4 = 0, w5 = 0, w6 = 0;
  if (P.mask[o] == 3) {
    T ak = a0 * s0[o];
    if (s1) ak += a1 * s1[o];
    if (s2) ak += a2 * s2[o];
    const Vel<T> V = vel_T(h, nx, i, j, P);
    const Diffs<T> D = diffs(q, nx, i, j);
    T dxp = D.xp, dxm = D.xm, dyp = D.yp, dym = D.ym;
    if (P.fbathy) {
      const Diffs<T> B = diffs(P.fbathy, nx, i, j);
      dxp += B.xp; dxm += B.xm; dyp += B.yp; dym += B.ym;
    }
    const T aux = ak * P.inv6dx, auy = ak * P.inv6dy;
    w0 = (V.uT < 0 ? T(0) : -aux * dxp) + (V.uT > 0 ? T(0) : aux * dxm);
    w1 = (V.vT < 0 ? T(0) : -auy * dyp) + (V.vT > 0 ? T(0) : auy * dym);
    if (P.beta) w1 -= P.beta[o] * ak;
    w2 = (V.uT < 0 ? T(0) : V.uT) * aux;
    w3 = (V.uT > 0 ? T(0) : V.uT) * aux;
    w4 = (V.vT < 0 ? T(0) : V.vT) * auy;
    w5 = (V.vT > 0 ? T(0) : V.vT) * auy;
    w6 = ak;
  }
  W0[o] = w0; W1[o] = w1; W2[o] = w2; W3[o] = w3; W4[o] = w4; W5[o] = w5; W6[o] = w6;
}

template <typename T>
__global__ void k_adj_b(Phys<T> P, const T* __restrict__ h, const T* __restrict__ W0,
                        const T* __restrict__ W1, const T* __restrict__ W2,
                        const T* __restrict__ W3, const T* __restrict__ W4,
                        const T* __restrict__ W5, const T* __restrict__ W6,
                        const T* acc0, const T* acc1, const T* acc2, T* lq_out,
                        T* __restrict__ lh_out) {
  // acc* may alias lq_out (in-place accumulation); each thread reads and
  // writes only its own point of those fields.
  VD_POINT_2D
  const int ny = P.ny;
  // W fields vanish outside 2..n-4, so only array bounds need checking.
  auto W = [&](const T* f, int ii, int jj) -> T {
    return (ii >= 0 && ii < ny && jj >= 0 && jj < nx) ? f[ii * nx + jj] : T(0);
  };
  T lq = 0;
  if (acc0) lq += acc0[o];
  if (acc1) lq += acc1[o];
  if (acc2) lq += acc2[o];
  // transpose of the x upwind stencils: T point t = p + d reads q(p)
  lq += (-2) * W(W2, i, j - 1) + (-6) * W(W3, i, j - 1) + (-3) * W(W2, i, j) +
        3 * W(W3, i, j) + 6 * W(W2, i, j + 1) + 2 * W(W3, i, j + 1) - W(W2, i, j + 2) +
        W(W3, i, j - 2);
  lq += (-2) * W(W4, i - 1, j) + (-6) * W(W5, i - 1, j) + (-3) * W(W4, i, j) +
        3 * W(W5, i, j) + 6 * W(W4, i + 1, j) + 2 * W(W5, i + 1, j) - W(W4, i + 2, j) +
        W(W5, i - 2, j);
  if (P.has_kdiff) {
    const T c = W(W6, i, j);
    lq += P.kx * (W(W6, i, j + 1) + W(W6, i, j - 1) - 2 * c) +
          P.ky * (W(W6, i + 1, j) + W(W6, i - 1, j) - 2 * c);
  }
  lq_out[o] = lq;

  // lambda_u[a][b] = 0.5 (W0[a][b] + W0[a][b-1]) on the u domain, zero where u was NaN
  auto lu = [&](int a, int b) -> T {
    if (a < 1 || a > ny - 2 || b < 1 || b > nx - 1) return T(0);
    if (u_isnan(h, nx, a, b)) return T(0);
    return T(0.5) * (W0[a * nx + b] + W0[a * nx + b - 1]);
  };
  auto lv = [&](int a, int b) -> T {
    if (a < 1 || a > ny - 1 || b < 1 || b > nx - 2) return T(0);
    if (v_isnan(h, nx, a, b)) return T(0);
    return T(0.5) * (W1[a * nx + b] + W1[(a - 1) * nx + b]);
  };
  const T lh = P.cu * (lu(i - 1, j + 1) + lu(i - 1, j) - lu(i + 1, j) - lu(i + 1, j + 1)) +
               P.cv * (lv(i, j - 1) + lv(i + 1, j - 1) - lv(i + 1, j + 1) - lv(i, j + 1));
  lh_out[o] = lh;
}

template <typename T>
__global__ void k_adj_init(Phys<T> P, const T* __restrict__ adh1, T* __restrict__ lh,
                           T* __restrict__ lq) {
  VD_POINT_2D
  lh[o] = P.mask[o] == 0 ? T(0) : adh1[o];
  lq[o] = T(0);
}

template <typename T>
__global__ void k_adj_final(Phys<T> P, const T* __restrict__ hcur, const T* __restrict__ lh,
                            const T* __restrict__ lq, T* __restrict__ adh0) {
  VD_POINT_2D
  const int ny = P.ny;
  // a(t) = lq(t) where h2pv used its Laplacian formula (mask 3, q not NaN)
  auto A = [&](int a, int b) -> T {
    if (a < 1 || a > ny - 2 || b < 1 || b > nx - 2) return T(0);
    const size_t t = (size_t)a * nx + b;
    if (P.mask[t] != 3 || q_isnan(hcur, nx, a, b)) return T(0);
    return lq[t];
  };
  const T r = P.gf_dy2 * (A(i + 1, j) + A(i - 1, j)) + P.gf_dx2 * (A(i, j + 1) + A(i, j - 1)) +
              (-2 * P.gf_dy2 - 2 * P.gf_dx2 - P.gfc2) * A(i, j);
  adh0[o] = lh[o] + r;
}

// ---------------------------------------------------------------------------
// Launchers
// ---------------------------------------------------------------------------
template <typename T>
Phys<T> make_phys(const Plan& p) {
  Phys<T> P;
  P.ny = p.ny;
  P.nx = p.nx;
  P.cu = T(-p.g / p.f0 / (4 * p.dy));
  P.cv = T(p.g / p.f0 / (4 * p.dx));
  P.inv6dx = T(1.0 / (6 * p.dx));
  P.inv6dy = T(1.0 / (6 * p.dy));
  P.kx = T(p.has_kdiff ? p.kdiffus / (p.dx * p.dx) : 0.0);
  P.ky = T(p.has_kdiff ? p.kdiffus / (p.dy * p.dy) : 0.0);
  P.gfc2 = T(p.g * p.f0 / (p.c * p.c));
  P.gf_dx2 = T(p.g / p.f0 / (p.dx * p.dx));
  P.gf_dy2 = T(p.g / p.f0 / (p.dy * p.dy));
  P.mask = p.mask;
  P.beta = p.has_beta ? static_cast<const T*>(p.beta) : nullptr;
  P.fbathy = p.has_bathy ? static_cast<const T*>(p.fbathy) : nullptr;
  P.has_kdiff = p.has_kdiff;
  return P;
}

template <typename T>
void launch_prep(const Plan& p, const T* h0, const T* hb, T* hcur, T* hbm, T* q0, T* qb,
                 hipStream_t s) {
  Phys<T> P = make_phys<T>(p);
  const T* mdt = p.has_mdt ? static_cast<const T*>(p.mdt) : nullptr;
  hipLaunchKernelGGL(k_prep<T>, grid2d(p.ny, p.nx), kBlock2d, 0, s, P, h0, hb, mdt, hcur, hbm,
                     q0, qb);
}

template <typename T>
void launch_h2pv(const Plan& p, const T* h, const T* hb, T* q, hipStream_t s) {
  Phys<T> P = make_phys<T>(p);
  hipLaunchKernelGGL(k_h2pv<T>, grid2d(p.ny, p.nx), kBlock2d, 0, s, P, h, hb, q);
}

template <typename T>
void launch_fill_land(const Plan& p, T* f, T value, hipStream_t s) {
  Phys<T> P = make_phys<T>(p);
  hipLaunchKernelGGL(k_fill_land<T>, grid2d(p.ny, p.nx), kBlock2d, 0, s, P, f, value);
}

template <typename T>
void launch_finish(const Plan& p, const T* h, T* h1, hipStream_t s) {
  Phys<T> P = make_phys<T>(p);
  const T* mdt = p.has_mdt ? static_cast<const T*>(p.mdt) : nullptr;
  hipLaunchKernelGGL(k_finish<T>, grid2d(p.ny, p.nx), kBlock2d, 0, s, P, h, mdt, h1);
}

template <typename T>
void launch_stage(const Plan& p, const T* h, const T* q, const T* qbase, const T* kprev0,
                  const T* kprev1, T c0, T c1, T c2, T* qout, T* kstore, hipStream_t s) {
  Phys<T> P = make_phys<T>(p);
  hipLaunchKernelGGL(k_stage<T>, grid2d(p.ny, p.nx), kBlock2d, 0, s, P, h, q, qbase, kprev0,
                     kprev1, c0, c1, c2, qout, kstore);
}

template <typename T>
void launch_stage_tl(const Plan& p, const T* h, const T* q, const T* dh, const T* dq,
                     const T* dqbase, const T* dkprev0, const T* dkprev1, T c0, T c1, T c2,
                     T* dqout, T* dkstore, hipStream_t s) {
  Phys<T> P = make_phys<T>(p);
  hipLaunchKernelGGL(k_stage_tl<T>, grid2d(p.ny, p.nx), kBlock2d, 0, s, P, h, q, dh, dq, dqbase,
                     dkprev0, dkprev1, c0, c1, c2, dqout, dkstore);
}

template <typename T>
void launch_prep_tl(const Plan& p, const T* dh0, const T* hcur, T* dhcur, T* dq0, hipStream_t s) {
  Phys<T> P = make_phys<T>(p);
  hipLaunchKernelGGL(k_prep_tl<T>, grid2d(p.ny, p.nx), kBlock2d, 0, s, P, dh0, hcur, dhcur, dq0);
}

template <typename T>
void launch_finish_tl(const Plan& p, const T* dh, T* dh1, hipStream_t s) {
  Phys<T> P = make_phys<T>(p);
  hipLaunchKernelGGL(k_finish_tl<T>, grid2d(p.ny, p.nx), kBlock2d, 0, s, P, dh, dh1);
}

template <typename T>
void launch_stage_adj(const Plan& p, const T* h, const T* q, const T* s0, const T* s1,
                      const T* s2, T a0, T a1, T a2, const T* acc0, const T* acc1,
                      const T* acc2, T* lq_out, T* lh_out, T* W[7], hipStream_t s) {
  Phys<T> P = make_phys<T>(p);
  hipLaunchKernelGGL(k_adj_a<T>, grid2d(p.ny, p.nx), kBlock2d, 0, s, P, h, q, s0, s1, s2, a0, a1,
                     a2, W[0], W[1], W[2], W[3], W[4], W[5], W[6]);
  hipLaunchKernelGGL(k_adj_b<T>, grid2d(p.ny, p.nx), kBlock2d, 0, s, P, h, W[0], W[1], W[2], W[3],
                     W[4], W[5], W[6], acc0, acc1, acc2, lq_out, lh_out);
}

template <typename T>
void launch_adj_init(const Plan& p, const T* adh1, T* lh, T* lq, hipStream_t s) {
  Phys<T> P = make_phys<T>(p);
  hipLaunchKernelGGL(k_adj_init<T>, grid2d(p.ny, p.nx), kBlock2d, 0, s, P, adh1, lh, lq);
}

template <typename T>
void launch_adj_final(const Plan& p, const T* hcur, const T* lh, const T* lq, T* adh0,
                      hipStream_t s) {
  Phys<T> P = make_phys<T>(p);
  hipLaunchKernelGGL(k_adj_final<T>, grid2d(p.ny, p.nx), kBlock2d, 0, s, P, hcur, lh, lq, adh0);
}

#define VD_INSTANTIATE(T)                                                                     \
  template Phys<T> make_phys<T>(const Plan&);                                                \
  template void launch_inversion<T>(const Plan&, const T*, const T*, const T*, T*, int,       \
                                    hipStream_t);                                             \
  template void launch_prep<T>(const Plan&, const T*, const T*, T*, T*, T*, T*, hipStream_t); \
  template void launch_h2pv<T>(const Plan&, const T*, const T*, T*, hipStream_t);             \
  template void launch_finish<T>(const Plan&, const T*, T*, hipStream_t);                     \
  template void launch_fill_land<T>(const Plan&, T*, T, hipStream_t);                         \
  template void launch_stage<T>(const Plan&, const T*, const T*, const T*, const T*,          \
                                const T*, T, T, T, T*, T*, hipStream_t);                      \
  template void launch_stage_tl<T>(const Plan&, const T*, const T*, const T*, const T*,       \
                                   const T*, const T*, const T*, T, T, T, T*, T*,             \
                                   hipStream_t);                                              \
  template void launch_prep_tl<T>(const Plan&, const T*, const T*, T*, T*, hipStream_t);      \
  template void launch_finish_tl<T>(const Plan&, const T*, T*, hipStream_t);                  \
  template void launch_stage_adj<T>(const Plan&, const T*, const T*, const T*, const T*,      \
                                    const T*, T, T, T, const T*, const T*, const T*, T*, T*,  \
                                    T* W[7], hipStream_t);                                    \
  template void launch_adj_init<T>(const Plan&, const T*, T*, T*, hipStream_t);               \
  template void launch_adj_final<T>(const Plan&, const T*, const T*, const T*, T*,            \
                                    hipStream_t);

VD_INSTANTIATE(float)
VD_INSTANTIATE(double)

}  // namespace vd
