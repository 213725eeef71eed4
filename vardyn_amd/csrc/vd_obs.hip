// L3 observation operator and FP64 reductions for the 4D-Var cost.
//
// Reference: Obsop_interp_l3_jax (mapping/src/obsop.py:375-678) and
// Variational.cost_and_grad (mapping/src/tools_4Dvar.py:252-400),
// datlas-ocean/VarDyn.
#include "vd_internal.h"

#include <algorithm>
#include <cmath>
#include <mutex>
#include <vector>

namespace vd {

constexpr int kRedThreads = 256;
constexpr int kRedBlocks = VD_REDUCE_SCRATCH;

// Deterministic block sum, valid in thread 0: each 64-lane wavefront adds
// its values by a fixed shuffle tree (lane l takes lane l + 32, 16, ..., 1),
// then thread 0 adds the four wavefront sums in order.  Every block writes
// its partial to part[block]; a single-block pass adds those in index order.
constexpr int kWaves = kRedThreads / 64;
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  return v;
}

__device__ __forceinline__ double block_sum(double v, double* sh) {
  v = wave_sum(v);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
  __syncthreads();
  double s = 0;
  if (threadIdx.x == 0) {
    s = sh[0];
#pragma unroll
    for (int w = 1; w < kWaves; ++w) s += sh[w];
  }
  return s;
}

__global__ __launch_bounds__(kRedThreads) void k_sum_partials(const double* __restrict__ part,
                                                              int n, double alpha,
                                                              double* __restrict__ out) {
  __shared__ double sh[kWaves];
  double v = 0;
  for (int i = threadIdx.x; i < n; i += kRedThreads) v += part[i];
  const double s = block_sum(v, sh);
  if (threadIdx.x == 0) *out += alpha * s;
}

static int red_blocks(int64_t n) {
  int64_t b = (n + kRedThreads - 1) / kRedThreads;
  if (b < 1) b = 1;
  if (b > kRedBlocks) b = kRedBlocks;
  return (int)b;
}

// misfit[o] = (sum_e w[e] x[col[e]] - y[o]) / err[o], NaN -> 0  (obsop.py:610-628)
template <typename TX>
__global__ __launch_bounds__(kRedThreads) void k_obs_misfit(
    int n_obs, const int32_t* __restrict__ row_ptr, const int32_t* __restrict__ col,
    const double* __restrict__ w, const double* __restrict__ y, const double* __restrict__ err,
    const TX* __restrict__ x, double* __restrict__ misfit, uint8_t* __restrict__ valid,
    double* __restrict__ part) {
  __shared__ double sh[kWaves];
  double m2 = 0;
  for (int o = blockIdx.x * kRedThreads + threadIdx.x; o < n_obs; o += gridDim.x * kRedThreads) {
    double hx = 0;
    const int e1 = row_ptr[o + 1];
    for (int e = row_ptr[o]; e < e1; ++e) hx += w[e] * (double)x[col[e]];
    const double d = hx - y[o];
    const bool ok = !isnan(d);
    double inverr = 1.0 / err[o];
    if (isnan(inverr)) inverr = 0.0;
    const double m = inverr * (ok ? d : 0.0);
    misfit[o] = m;
    valid[o] = ok ? 1 : 0;
    m2 += m * m;
  }
  if (part) {
    const double s = block_sum(m2, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = s;
  }
}

// adx[tcol[c]] += sum_{e in c} tw[e] * valid[trow[e]] * cot[trow[e]] / err[trow[e]]
// (the vjp of the misfit, obsop.py:630-654), one thread per grid point.
template <typename TX>
__global__ void k_obs_adjoint(int n_cols, const int32_t* __restrict__ tcol,
                              const int32_t* __restrict__ tptr, const int32_t* __restrict__ trow,
                              const double* __restrict__ tw, const double* __restrict__ err,
                              const double* __restrict__ cot, const uint8_t* __restrict__ valid,
                              TX* __restrict__ adx) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n_cols) return;
  double s = 0;
  const int e1 = tptr[c + 1];
  for (int e = tptr[c]; e < e1; ++e) {
    const int o = trow[e];
    if (!valid[o]) continue;
    double inverr = 1.0 / err[o];
    if (isnan(inverr)) inverr = 0.0;
    s += tw[e] * (inverr * cot[o]);
  }
  const int g = tcol[c];
  adx[g] = adx[g] + (TX)s;
}

__global__ __launch_bounds__(kRedThreads) void k_dot(int64_t n, const double* __restrict__ x,
                                                     const double* __restrict__ y,
                                                     double* __restrict__ part) {
  __shared__ double sh[kWaves];
  double v = 0;
  for (int64_t i = (int64_t)blockIdx.x * kRedThreads + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * kRedThreads)
    v += x[i] * y[i];
  const double s = block_sum(v, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

template <typename T>
__global__ void k_axpby(int64_t n, T a, const T* __restrict__ x, T b, T* __restrict__ y) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    y[i] = a * x[i] + b * y[i];
}

template <typename TS, typename TD>
__global__ void k_convert(int64_t n, const TS* __restrict__ src, TD* __restrict__ dst, int nz) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    TS v = src[i];
    if (nz && isnan(v)) v = TS(0);
    dst[i] = (TD)v;
  }
}

static int grid1d(int64_t n) {
  int64_t b = (n + 255) / 256;
  if (b < 1) b = 1;
  if (b > 65535) b = 65535;
  return (int)b;
}

// Device scratch of vd_dot_f64_noalloc: one partial sum per block of a call,
// at most kRedBlocks of them (8 KiB).  In host code `scratch` is a host-side
// stand-in, not a device address: kernels get the address that
// hipGetSymbolAddress returns for the current device (each device holds its
// own copy of the array).
static __device__ __align__(256) double scratch[VD_REDUCE_SCRATCH];

constexpr int kMaxDevices = 64;
static double* g_scratch[kMaxDevices];
static std::mutex g_scratch_mu;

// The current device's address of `scratch`, looked up on the first call on
// that device and kept.
static int device_scratch(double** out) {
  int dev = -1;
  VD_HIP_TRY(hipGetDevice(&dev));
  VD_REQUIRE(dev >= 0 && dev < kMaxDevices, VD_ERR_UNSUPPORTED, "device index out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch[dev]) {
    void* p = nullptr;
    VD_HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    VD_REQUIRE(p != nullptr, VD_ERR_HIP, "hipGetSymbolAddress returned no address for the "
                                         "dot-product scratch");
    g_scratch[dev] = static_cast<double*>(p);
  }
  *out = g_scratch[dev];
  return VD_OK;
}

// ------------------------------------------------------------ operator construction
// The Npix nearest grid points of each observation that lie within dmax
// (Obsop_interp_l3_jax._sparse_op, obsop.py:417-460: a scipy KDTree query for
// the Npix nearest points, then the cut D <= dmax; the result is the Npix
// nearest among the points within dmax).  The grid's cartesian points are
// binned into cubes of side L >= dmax and sorted by cube key once, on the
// host, at creation; a query visits the 3 x 3 x 3 cubes around the
// observation's cube, which hold every point within dmax, as nine key ranges
// (the three cubes along z have consecutive keys).  Distances are formed as
// scipy's KDTree forms them, sqrt((dx^2 + dy^2) + dz^2) without contraction,
// and ranked by their square with the grid index breaking exact ties.
constexpr int kKnnMax = 8;
constexpr int kKeyBits = 21;
constexpr int64_t kKeyOff = int64_t(1) << (kKeyBits - 1);

__host__ __device__ inline int64_t cube_key(int64_t cx, int64_t cy, int64_t cz) {
  return ((cx + kKeyOff) << (2 * kKeyBits)) | ((cy + kKeyOff) << kKeyBits) | (cz + kKeyOff);
}

struct ObsGridDev {
  int64_t n;
  double dmax, cube;
  const int64_t* key;  // sorted cube keys
  const double* x;     // cartesian km of the points, in key order
  const double* y;
  const double* z;
  const int32_t* idx;  // grid index of each point
};

__device__ __forceinline__ int64_t lower_bound_key(const int64_t* __restrict__ key, int64_t n, int64_t v) {
  int64_t lo = 0, hi = n;
  while (lo < hi) {
    const int64_t mid = (lo + hi) >> 1;
    if (key[mid] < v) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

// One thread per observation; xyz is n_obs x 3 (geo2cart of the
// observation).  Slots past the neighbours found get col -1, weight 0 and
// distance +inf.  Weights: exp(-D^2 / (2 (dmax / 2)^2)) normalised over the
// observation's neighbours (obsop.py:436-447).
__global__ __launch_bounds__(256) void k_obs_knn(ObsGridDev g, const double* __restrict__ xyz, int n_obs, int npix,
                                                 int32_t* __restrict__ col, double* __restrict__ w,
                                                 double* __restrict__ dist) {
#pragma clang fp contract(off)
  const int o = blockIdx.x * blockDim.x + threadIdx.x;
  if (o >= n_obs) return;
  double bd[kKnnMax];
  int32_t bi[kKnnMax];
#pragma unroll
  for (int s = 0; s < kKnnMax; ++s) {
    bd[s] = __builtin_inf();
    bi[s] = -1;
  }
  const double px = xyz[3 * (size_t)o], py = xyz[3 * (size_t)o + 1], pz = xyz[3 * (size_t)o + 2];
  if (isfinite(px) && isfinite(py) && isfinite(pz) && fabs(px) < 1e5 && fabs(py) < 1e5 && fabs(pz) < 1e5) {
    const int64_t cx = (int64_t)floor(px / g.cube), cy = (int64_t)floor(py / g.cube),
                  cz = (int64_t)floor(pz / g.cube);
    for (int ab = 0; ab < 9; ++ab) {
      const int a = ab / 3 - 1, b = ab % 3 - 1;
      // the z-column of three cubes: keys [key(cz - 1), key(cz + 1)]
      const int64_t first = lower_bound_key(g.key, g.n, cube_key(cx + a, cy + b, cz - 1));
      const int64_t last = lower_bound_key(g.key, g.n, cube_key(cx + a, cy + b, cz + 1) + 1);
      {
        for (int64_t e = first; e < last; ++e) {
          const double dx = px - g.x[e], dy = py - g.y[e], dz = pz - g.z[e];
          double cd = (dx * dx + dy * dy) + dz * dz;
          if (!(sqrt(cd) <= g.dmax)) continue;
          int32_t ci = g.idx[e];
          // insert (cd, ci) into the ascending list (bubble the larger one on)
#pragma unroll
          for (int s = 0; s < kKnnMax; ++s) {
            const bool lt = bi[s] < 0 || cd < bd[s] || (cd == bd[s] && ci < bi[s]);
            if (lt && ci >= 0) {
              const double td = bd[s];
              const int32_t ti = bi[s];
              bd[s] = cd;
              bi[s] = ci;
              cd = td;
              ci = ti;
            }
          }
        }
      }
    }
  }
  const double h = 0.5 * g.dmax, den = 2.0 * (h * h);
  double ws[kKnnMax], sum = 0.0;
#pragma unroll
  for (int s = 0; s < kKnnMax; ++s) {
    ws[s] = 0.0;
    if (s < npix && bi[s] >= 0) {
      const double D = sqrt(bd[s]);
      ws[s] = exp(-(D * D) / den);
    }
    if (s < npix) sum = sum + ws[s];
  }
#pragma unroll
  for (int s = 0; s < kKnnMax; ++s) {
    if (s >= npix) break;
    const bool ok = bi[s] >= 0;
    col[(size_t)o * npix + s] = ok ? bi[s] : -1;
    w[(size_t)o * npix + s] = (ok && sum > 0.0) ? ws[s] / sum : 0.0;
    if (dist) dist[(size_t)o * npix + s] = ok ? sqrt(bd[s]) : __builtin_inf();
  }
}

}  // namespace vd

struct vd_obs_grid {
  int dev = 0;
  vd::ObsGridDev g{};
  std::vector<void*> owned;
};

using namespace vd;

extern "C" {

int vd_obs_grid_create(const double* xyz, int64_t n_grid, double dmax, vd_obs_grid** out) {
  VD_REQUIRE(xyz && out, VD_ERR_INVALID_ARG, "null argument");
  *out = nullptr;
  VD_REQUIRE(n_grid > 0 && n_grid < (int64_t(1) << 31), VD_ERR_INVALID_ARG,
             "obs grid: the number of grid points must be in [1, 2^31)");
  VD_REQUIRE(std::isfinite(dmax) && dmax > 0, VD_ERR_INVALID_ARG, "obs grid: dmax must be positive and finite");
  // cubes of side >= dmax (a point within dmax is at most one cube away along
  // each axis; the margin covers the rounding of x / L), and large enough for
  // every key to fit 21 bits per axis within |x| < 1e5 km
  const double cube = std::max(dmax * (1.0 + 1e-9), 1e5 / double(kKeyOff - 4));
  std::vector<int64_t> key(n_grid);
  for (int64_t i = 0; i < n_grid; ++i) {
    const double* p = xyz + 3 * i;
    VD_REQUIRE(std::isfinite(p[0]) && std::isfinite(p[1]) && std::isfinite(p[2]) && std::fabs(p[0]) < 1e5 &&
                   std::fabs(p[1]) < 1e5 && std::fabs(p[2]) < 1e5,
               VD_ERR_INVALID_ARG, "obs grid: grid coordinates must be finite and within 1e5 km");
    key[i] = cube_key((int64_t)std::floor(p[0] / cube), (int64_t)std::floor(p[1] / cube),
                      (int64_t)std::floor(p[2] / cube));
  }
  std::vector<int32_t> order(n_grid);
  for (int64_t i = 0; i < n_grid; ++i) order[i] = (int32_t)i;
  std::stable_sort(order.begin(), order.end(), [&](int32_t a, int32_t b) { return key[a] < key[b]; });
  std::vector<int64_t> skey(n_grid);
  std::vector<double> sx(n_grid), sy(n_grid), sz(n_grid);
  for (int64_t k = 0; k < n_grid; ++k) {
    const int32_t i = order[k];
    skey[k] = key[i];
    sx[k] = xyz[3 * (size_t)i];
    sy[k] = xyz[3 * (size_t)i + 1];
    sz[k] = xyz[3 * (size_t)i + 2];
  }
  vd_obs_grid* m = new vd_obs_grid();
  (void)hipGetDevice(&m->dev);
  auto up = [&](const void* src, size_t bytes, void** dst) -> hipError_t {
    hipError_t e = hipMalloc(dst, bytes);
    if (e != hipSuccess) return e;
    m->owned.push_back(*dst);
    return hipMemcpy(*dst, src, bytes, hipMemcpyHostToDevice);
  };
  void *dk, *dx, *dy, *dz, *di;
  hipError_t e = up(skey.data(), n_grid * sizeof(int64_t), &dk);
  if (e == hipSuccess) e = up(sx.data(), n_grid * sizeof(double), &dx);
  if (e == hipSuccess) e = up(sy.data(), n_grid * sizeof(double), &dy);
  if (e == hipSuccess) e = up(sz.data(), n_grid * sizeof(double), &dz);
  if (e == hipSuccess) e = up(order.data(), n_grid * sizeof(int32_t), &di);
  if (e != hipSuccess) {
    vd_obs_grid_destroy(m);
    VD_HIP_TRY(e);
  }
  m->g = ObsGridDev{n_grid, dmax, cube, (const int64_t*)dk, (const double*)dx, (const double*)dy,
                    (const double*)dz, (const int32_t*)di};
  *out = m;
  return VD_OK;
}

int vd_obs_grid_destroy(vd_obs_grid* m) {
  if (!m) return VD_OK;
  for (void* p : m->owned) hipFree(p);
  delete m;
  return VD_OK;
}

int vd_obs_grid_query(const vd_obs_grid* m, const double* xyz, int n_obs, int npix, int32_t* col, double* w,
                      double* dist, void* stream) {
  VD_REQUIRE(m && col && w && (xyz || n_obs == 0), VD_ERR_INVALID_ARG, "null argument");
  VD_REQUIRE(n_obs >= 0, VD_ERR_INVALID_ARG, "obs grid: negative observation count");
  VD_REQUIRE(npix >= 1 && npix <= kKnnMax, VD_ERR_INVALID_ARG, "obs grid: npix must be in [1, 8]");
  VD_REQUIRE((int64_t)n_obs * npix < (int64_t(1) << 31), VD_ERR_INVALID_ARG, "obs grid: too many observations");
  if (n_obs == 0) return VD_OK;
  hipStream_t s = static_cast<hipStream_t>(stream);
  hipLaunchKernelGGL(k_obs_knn, dim3((n_obs + 255) / 256), dim3(256), 0, s, m->g, xyz, n_obs, npix, col, w, dist);
  VD_HIP_TRY(hipGetLastError());
  return VD_OK;
}

int vd_obs_misfit(const vd_obs_table* t, int precision, const void* x, double* misfit,
                  uint8_t* valid, double* jo, double inv_sigma_r2, void* stream) {
  VD_REQUIRE(t && x && misfit && valid, VD_ERR_INVALID_ARG, "null argument");
  VD_REQUIRE(precision == VD_F32 || precision == VD_F64, VD_ERR_INVALID_ARG, "bad precision");
  VD_REQUIRE(!jo || t->partial, VD_ERR_INVALID_ARG, "jo requires table scratch (partial)");
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (t->n_obs <= 0) return VD_OK;
  const int nb = red_blocks(t->n_obs);
  double* part = jo ? t->partial : nullptr;
  if (precision == VD_F64)
    hipLaunchKernelGGL(k_obs_misfit<double>, dim3(nb), dim3(kRedThreads), 0, s, t->n_obs,
                       t->row_ptr, t->col, t->w, t->y, t->err, (const double*)x, misfit, valid,
                       part);
  else
    hipLaunchKernelGGL(k_obs_misfit<float>, dim3(nb), dim3(kRedThreads), 0, s, t->n_obs,
                       t->row_ptr, t->col, t->w, t->y, t->err, (const float*)x, misfit, valid,
                       part);
  if (jo)
    hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(kRedThreads), 0, s, t->partial, nb,
                       inv_sigma_r2, jo);
  VD_HIP_TRY(hipGetLastError());
  return VD_OK;
}

int vd_obs_adjoint(const vd_obs_table* t, int precision, const double* cot, const uint8_t* valid,
                   void* adx, void* stream) {
  VD_REQUIRE(t && cot && valid && adx, VD_ERR_INVALID_ARG, "null argument");
  VD_REQUIRE(precision == VD_F32 || precision == VD_F64, VD_ERR_INVALID_ARG, "bad precision");
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (t->n_cols <= 0) return VD_OK;
  const int nb = (t->n_cols + 255) / 256;
  if (precision == VD_F64)
    hipLaunchKernelGGL(k_obs_adjoint<double>, dim3(nb), dim3(256), 0, s, t->n_cols, t->tcol,
                       t->tptr, t->trow, t->tw, t->err, cot, valid, (double*)adx);
  else
    hipLaunchKernelGGL(k_obs_adjoint<float>, dim3(nb), dim3(256), 0, s, t->n_cols, t->tcol,
                       t->tptr, t->trow, t->tw, t->err, cot, valid, (float*)adx);
  VD_HIP_TRY(hipGetLastError());
  return VD_OK;
}

int vd_dot_f64(int64_t n, const double* x, const double* y, double alpha, double* out,
               double* scratch, void* stream) {
  VD_REQUIRE(x && y && out && scratch, VD_ERR_INVALID_ARG, "null argument");
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (n <= 0) return VD_OK;
  const int nb = red_blocks(n);
  hipLaunchKernelGGL(k_dot, dim3(nb), dim3(kRedThreads), 0, s, n, x, y, scratch);
  hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(kRedThreads), 0, s, scratch, nb, alpha, out);
  VD_HIP_TRY(hipGetLastError());
  return VD_OK;
}

int vd_dot_f64_noalloc(int64_t n, const double* x, const double* y, double alpha, double* out,
                       void* stream) {
  VD_REQUIRE(x && y && out, VD_ERR_INVALID_ARG, "null argument");
  VD_REQUIRE(n >= 0 && n <= VD_DOT_NOALLOC_MAX_N, VD_ERR_INVALID_ARG,
             "vd_dot_f64_noalloc: n must be in [0, VD_DOT_NOALLOC_MAX_N]");
  double* part = nullptr;
  int rc = device_scratch(&part);
  if (rc != VD_OK) return rc;
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (n == 0) return VD_OK;
  // the same two launches as vd_dot_f64; every partial the second one reads
  // is written by the first, so nothing in the scratch needs clearing
  const int nb = red_blocks(n);
  hipLaunchKernelGGL(k_dot, dim3(nb), dim3(kRedThreads), 0, s, n, x, y, part);
  hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(kRedThreads), 0, s, part, nb, alpha, out);
  VD_HIP_TRY(hipGetLastError());
  return VD_OK;
}

int vd_axpby(int precision, int64_t n, double a, const void* x, double b, void* y, void* stream) {
  VD_REQUIRE(x && y, VD_ERR_INVALID_ARG, "null argument");
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (n <= 0) return VD_OK;
  if (precision == VD_F64)
    hipLaunchKernelGGL(k_axpby<double>, dim3(grid1d(n)), dim3(256), 0, s, n, a, (const double*)x,
                       b, (double*)y);
  else if (precision == VD_F32)
    hipLaunchKernelGGL(k_axpby<float>, dim3(grid1d(n)), dim3(256), 0, s, n, (float)a,
                       (const float*)x, (float)b, (float*)y);
  else
    VD_REQUIRE(false, VD_ERR_INVALID_ARG, "bad precision");
  VD_HIP_TRY(hipGetLastError());
  return VD_OK;
}

int vd_convert(int sp, const void* src, int dp, void* dst, int64_t n, int nan_to_zero,
               void* stream) {
  VD_REQUIRE(src && dst, VD_ERR_INVALID_ARG, "null argument");
  VD_REQUIRE((sp == VD_F32 || sp == VD_F64) && (dp == VD_F32 || dp == VD_F64),
             VD_ERR_INVALID_ARG, "bad precision");
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (n <= 0) return VD_OK;
  const dim3 g(grid1d(n)), b(256);
  if (sp == VD_F64 && dp == VD_F64)
    hipLaunchKernelGGL((k_convert<double, double>), g, b, 0, s, n, (const double*)src,
                       (double*)dst, nan_to_zero);
  else if (sp == VD_F64 && dp == VD_F32)
    hipLaunchKernelGGL((k_convert<double, float>), g, b, 0, s, n, (const double*)src, (float*)dst,
                       nan_to_zero);
  else if (sp == VD_F32 && dp == VD_F64)
    hipLaunchKernelGGL((k_convert<float, double>), g, b, 0, s, n, (const float*)src, (double*)dst,
                       nan_to_zero);
  else
    hipLaunchKernelGGL((k_convert<float, float>), g, b, 0, s, n, (const float*)src, (float*)dst,
                       nan_to_zero);
  VD_HIP_TRY(hipGetLastError());
  return VD_OK;
}

}  // extern "C"
