// Merge of subwindow analyses onto the target grid (compute_weights_map,
// merge_output_date and merge_time_windows_outputs,
// mapping/src/run_assimilation.py:351-412, 537-607, 732-830).
//
// Each subwindow s has a rectilinear grid (ascending 1-D latitude and
// longitude axes), or an irregular 2-D grid with host-built Delaunay
// triangle tables (the reference's LinearNDInterpolator branch), and
// carries, per date, a field on it.  Per target point p
// and subwindow s the merge uses scipy's RegularGridInterpolator (linear,
// bounds_error False, fill NaN) semantics: the cell [a_i, a_{i+1}) holding
// the coordinate on each axis (the last cell closed on the right), NaN
// outside the axis range, and the bilinear combination of the cell's four
// corners, so a NaN corner makes the value NaN.
//   k_merge_setup   per (s, p): the cell and the fractions, and the interpolated
//                   raised-cosine weight w_s(p) of the subwindow (the separable
//                   Hann window 0.5 (1 - cos) over linspace(0, 2 pi, n) on each
//                   axis; ones when there is a single subwindow)
//   k_merge_wsum    wsum(p) = sum over s, ascending, of the w_s(p) that are not NaN
//   k_merge_apply   out[t][p] = sum over s, ascending, of (w_s(p) v_s(t, p)) / wsum(p)
//                   over the subwindows whose interpolated value v_s is not NaN;
//                   NaN on masked points.  One launch for all dates: blockIdx.y
//                   takes 8 of them, which share the tables.
//   k_merge_blend   out = W1 a + W2 b (the raised-cosine blend of two time windows)
#include <cmath>

#include "vd_internal.h"

namespace vd {

constexpr int kMgTB = 8;  // dates per pass of the merge

struct MergeTables {
  int ny, nx, nsub;
  const double* lon;      // ny * nx target longitudes
  const double* lat;
  const uint8_t* mask;    // ny * nx or null
  const int* sny;         // nsub
  const int* snx;
  const int* yoff;        // nsub: offset of subwindow s's latitude axis in slat
  const int* xoff;
  const long long* foff;  // nsub: points before subwindow s (sum of sny * snx)
  const double* slat;
  const double* slon;
  int2* cell;             // nsub * np: (row, column) of the cell's lower corner, row -1 outside
  double2* frac;          // nsub * np: (fy, fx)
  double* w;              // nsub * np: interpolated weight (NaN outside)
  double* wsum;           // np
  const int* tri_slot;    // nsub: block of an irregular subwindow's triangle tables, -1 rectilinear
  const int* tri_idx;     // blocks of np * 3 vertices (-1 outside)
  const double* tri_w;    // blocks of np * 3 barycentric weights
};

// LinearNDInterpolator's evaluation on one triangle (scipy interpnd:
// out = 0; out += c_j v_j for j = 0, 1, 2), unfused like its C loop
__device__ __forceinline__ double mg_tri(const double* v, const int* id, const double* c) {
#pragma clang fp contract(off)
  double out = 0.0;
  out = out + c[0] * v[id[0]];
  out = out + c[1] * v[id[1]];
  out = out + c[2] * v[id[2]];
  return out;
}

// scipy's interval search on an ascending axis a[0..n-1]: the i with
// a[i] <= v < a[i+1], n-2 for v == a[n-1], -1 outside or for NaN
__device__ __forceinline__ int mg_interval(const double* a, int n, double v) {
  if (!(v >= a[0] && v <= a[n - 1])) return -1;
  int lo = 0, hi = n - 1;  // a[lo] <= v, and v < a[hi] unless v == a[n-1]
  if (v >= a[n - 1]) return n - 2;
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (a[mid] <= v) lo = mid;
    else hi = mid;
  }
  return lo;
}

__device__ __forceinline__ double mg_bilinear(double v00, double v01, double v10, double v11, double fy,
                                              double fx) {
  return v00 * (1 - fy) * (1 - fx) + v01 * (1 - fy) * fx + v10 * fy * (1 - fx) + v11 * fy * fx;
}

// raised-cosine window 0.5 (1 - cos(t_i)) of numpy's linspace(0, 2 pi, n)
__device__ __forceinline__ double mg_hann(int i, int n) {
  const double step = 2 * M_PI / (n - 1);
  const double t = i == n - 1 ? 2 * M_PI : i * step;
  return 0.5 * (1.0 - cos(t));
}

__global__ void k_merge_setup(MergeTables M) {
  const int np = M.ny * M.nx;
  const long long n = (long long)M.nsub * np;
  for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < n;
       e += (long long)gridDim.x * blockDim.x) {
    const int s = (int)(e / np), p = (int)(e % np);
    const int ny = M.sny[s], nx = M.snx[s];
    if (M.tri_slot && M.tri_slot[s] >= 0) {
      // irregular grid: the window (ones for a single subwindow) interpolated on the triangle
      const size_t t3 = ((size_t)M.tri_slot[s] * np + p) * 3;
      const int* id = M.tri_idx + t3;
      M.cell[e] = make_int2(-1, -1);
      M.frac[e] = make_double2(0, 0);
      if (id[0] < 0) {
        M.w[e] = NAN;
        continue;
      }
      double wv[3];
      const int loc[3] = {0, 1, 2};
      for (int k = 0; k < 3; ++k)
        wv[k] = M.nsub == 1 ? 1.0 : mg_hann(id[k] / nx, ny) * mg_hann(id[k] % nx, nx);
      M.w[e] = mg_tri(wv, loc, M.tri_w + t3);
      continue;
    }
    const double* ay = M.slat + M.yoff[s];
    const double* ax = M.slon + M.xoff[s];
    const double la = M.lat[p], lo = M.lon[p];
    const int i = mg_interval(ay, ny, la), j = mg_interval(ax, nx, lo);
    if (i < 0 || j < 0) {
      M.cell[e] = make_int2(-1, -1);
      M.frac[e] = make_double2(0, 0);
      M.w[e] = NAN;
      continue;
    }
    const double fy = (la - ay[i]) / (ay[i + 1] - ay[i]);
    const double fx = (lo - ax[j]) / (ax[j + 1] - ax[j]);
    M.cell[e] = make_int2(i, j);
    M.frac[e] = make_double2(fy, fx);
    if (M.nsub == 1) {
      M.w[e] = mg_bilinear(1, 1, 1, 1, fy, fx);
    } else {
      const double y0 = mg_hann(i, ny), y1 = mg_hann(i + 1, ny);
      const double x0 = mg_hann(j, nx), x1 = mg_hann(j + 1, nx);
      M.w[e] = mg_bilinear(y0 * x0, y0 * x1, y1 * x0, y1 * x1, fy, fx);
    }
  }
}

__global__ void k_merge_wsum(MergeTables M) {
  const int np = M.ny * M.nx;
  for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < np; p += gridDim.x * blockDim.x) {
    double acc = 0;
    for (int s = 0; s < M.nsub; ++s) {
      const double w = M.w[(size_t)s * np + p];
      if (!isnan(w)) acc += w;
    }
    M.wsum[p] = acc;
  }
}

// fields: subwindow s's T fields (sny[s] x snx[s] each) start at T * foff[s]
__global__ __launch_bounds__(256) void k_merge_apply(MergeTables M, int T,
                                                     const double* __restrict__ fields,
                                                     double* __restrict__ out) {
  const int np = M.ny * M.nx;
  const int t0 = blockIdx.y * kMgTB;
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= np) return;
  const int nt = min(kMgTB, T - t0);
  double acc[kMgTB] = {};
  const double ws = M.wsum[p];
  for (int s = 0; s < M.nsub; ++s) {
    const size_t e = (size_t)s * np + p;
    if (M.tri_slot && M.tri_slot[s] >= 0) {
      const size_t t3 = ((size_t)M.tri_slot[s] * np + p) * 3;
      const int* id = M.tri_idx + t3;
      if (id[0] < 0) continue;
      const double* c = M.tri_w + t3;
      const double w = M.w[e];
      const long long plane = (long long)M.sny[s] * M.snx[s];
      const double* base = fields + (long long)T * M.foff[s];
#pragma unroll
      for (int t = 0; t < kMgTB; ++t) {
        if (t >= nt) break;
        const double val = mg_tri(base + (long long)(t0 + t) * plane, id, c);
        if (!isnan(val)) acc[t] += w * val / ws;
      }
      continue;
    }
    const int2 c = M.cell[e];
    if (c.x < 0) continue;
    const double2 f = M.frac[e];
    const double w = M.w[e];
    const int nx = M.snx[s];
    const long long plane = (long long)M.sny[s] * nx;
    const double* base = fields + (long long)T * M.foff[s] + (long long)c.x * nx + c.y;
#pragma unroll
    for (int t = 0; t < kMgTB; ++t) {
      if (t >= nt) break;
      const double* v = base + (long long)(t0 + t) * plane;
      const double val = mg_bilinear(v[0], v[1], v[nx], v[nx + 1], f.x, f.y);
      if (!isnan(val)) acc[t] += w * val / ws;
    }
  }
  const bool masked = M.mask && M.mask[p];
#pragma unroll
  for (int t = 0; t < kMgTB; ++t)
    if (t < nt) out[(long long)(t0 + t) * np + p] = masked ? NAN : acc[t];
}

__global__ void k_merge_blend(long long n, double w1, double w2, const double* __restrict__ a,
                              const double* __restrict__ b, double* __restrict__ out) {
#pragma clang fp contract(off)  // two roundings and a sum, as W1 * ds1 + W2 * ds2 in numpy
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    out[i] = w1 * a[i] + w2 * b[i];
}

}  // namespace vd

using namespace vd;

struct vd_merge {
  MergeTables M;
  std::vector<void*> owned;
  long long total;  // sum of sny * snx
};

namespace {

template <class V>
int mg_upload(vd_merge* m, const V* src, size_t n, const V** dst) {
  *dst = nullptr;
  if (n == 0) return VD_OK;
  void* d = nullptr;
  VD_HIP_TRY(hipMalloc(&d, n * sizeof(V)));
  m->owned.push_back(d);
  VD_HIP_TRY(hipMemcpy(d, src, n * sizeof(V), hipMemcpyHostToDevice));
  *dst = static_cast<const V*>(d);
  return VD_OK;
}

template <class V>
int mg_alloc(vd_merge* m, size_t n, V** dst) {
  void* d = nullptr;
  VD_HIP_TRY(hipMalloc(&d, n * sizeof(V)));
  m->owned.push_back(d);
  *dst = static_cast<V*>(d);
  return VD_OK;
}

unsigned mg_blocks(long long n) { return (unsigned)std::min<long long>((n + 255) / 256, 65535); }

}  // namespace

int vd_merge_create(const vd_merge_params* params, vd_merge** out) {
  VD_REQUIRE(params && out, VD_ERR_INVALID_ARG, "null argument");
  *out = nullptr;
  const vd_merge_params& P = *params;
  VD_REQUIRE(P.ny > 0 && P.nx > 0 && P.nsub > 0 && P.lon && P.lat && P.sny && P.snx && P.slat && P.slon,
             VD_ERR_INVALID_ARG, "merge: empty target grid, no subwindow or null table");
  VD_REQUIRE((long long)P.ny * P.nx < (1LL << 31), VD_ERR_INVALID_ARG, "merge: target grid too large");
  std::vector<int> yoff(P.nsub), xoff(P.nsub);
  std::vector<long long> foff(P.nsub);
  long long ny_tot = 0, nx_tot = 0, tot = 0;
  int n_irr = 0;
  const size_t np_target = (size_t)P.ny * P.nx;
  for (int s = 0; s < P.nsub; ++s) {
    VD_REQUIRE(P.sny[s] >= 2 && P.snx[s] >= 2, VD_ERR_INVALID_ARG,
               "merge: a subwindow axis needs at least 2 points");
    const bool irr = P.tri_slot && P.tri_slot[s] >= 0;
    const double* ay = P.slat + ny_tot;
    const double* ax = P.slon + nx_tot;
    if (irr) {
      VD_REQUIRE(P.tri_slot[s] == n_irr, VD_ERR_INVALID_ARG,
                 "merge: irregular subwindows take triangle blocks 0, 1, ... in order");
      VD_REQUIRE(P.tri_idx && P.tri_w, VD_ERR_INVALID_ARG, "merge: null triangle tables");
      const long long npts = (long long)P.sny[s] * P.snx[s];
      const int32_t* id = P.tri_idx + (size_t)n_irr * np_target * 3;
      for (size_t q = 0; q < np_target * 3; ++q)
        VD_REQUIRE(id[q] >= -1 && id[q] < npts && ((q % 3 == 0) || (id[q] < 0) == (id[q - q % 3] < 0)),
                   VD_ERR_INVALID_ARG, "merge: triangle vertex outside its subwindow");
      ++n_irr;
    } else {
      for (int i = 0; i + 1 < P.sny[s]; ++i)
        VD_REQUIRE(ay[i] < ay[i + 1], VD_ERR_INVALID_ARG,
                   "merge: subwindow latitudes must be strictly ascending");
      for (int j = 0; j + 1 < P.snx[s]; ++j)
        VD_REQUIRE(ax[j] < ax[j + 1], VD_ERR_INVALID_ARG,
                   "merge: subwindow longitudes must be strictly ascending");
    }
    yoff[s] = (int)ny_tot;
    xoff[s] = (int)nx_tot;
    foff[s] = tot;
    ny_tot += P.sny[s];
    nx_tot += P.snx[s];
    tot += (long long)P.sny[s] * P.snx[s];
  }
  const size_t np = (size_t)P.ny * P.nx;
  vd_merge* m = new vd_merge();
  m->total = tot;
  MergeTables& M = m->M;
  M.ny = P.ny;
  M.nx = P.nx;
  M.nsub = P.nsub;
  int rc = VD_OK;
  if ((rc = mg_upload(m, P.lon, np, &M.lon)) || (rc = mg_upload(m, P.lat, np, &M.lat)) ||
      (rc = mg_upload(m, P.mask, P.mask ? np : 0, &M.mask)) ||
      (rc = mg_upload(m, (const int*)P.sny, (size_t)P.nsub, &M.sny)) ||
      (rc = mg_upload(m, (const int*)P.snx, (size_t)P.nsub, &M.snx)) ||
      (rc = mg_upload(m, (const int*)yoff.data(), yoff.size(), &M.yoff)) ||
      (rc = mg_upload(m, (const int*)xoff.data(), xoff.size(), &M.xoff)) ||
      (rc = mg_upload(m, (const long long*)foff.data(), foff.size(), &M.foff)) ||
      (rc = mg_upload(m, P.slat, (size_t)ny_tot, &M.slat)) ||
      (rc = mg_upload(m, P.slon, (size_t)nx_tot, &M.slon)) ||
      (rc = mg_alloc(m, np * P.nsub, &M.cell)) || (rc = mg_alloc(m, np * P.nsub, &M.frac)) ||
      (rc = mg_alloc(m, np * P.nsub, &M.w)) || (rc = mg_alloc(m, np, &M.wsum)) ||
      (rc = mg_upload(m, (const int*)P.tri_slot, n_irr ? (size_t)P.nsub : 0, &M.tri_slot)) ||
      (rc = mg_upload(m, (const int*)P.tri_idx, (size_t)n_irr * np * 3, &M.tri_idx)) ||
      (rc = mg_upload(m, P.tri_w, (size_t)n_irr * np * 3, &M.tri_w))) {
    vd_merge_destroy(m);
    return rc;
  }
  hipLaunchKernelGGL(k_merge_setup, dim3(mg_blocks((long long)np * P.nsub)), dim3(256), 0, 0, M);
  hipLaunchKernelGGL(k_merge_wsum, dim3(mg_blocks((long long)np)), dim3(256), 0, 0, M);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipDeviceSynchronize();
  if (e != hipSuccess) {
    vd_merge_destroy(m);
    VD_HIP_TRY(e);
  }
  *out = m;
  return VD_OK;
}

int vd_merge_destroy(vd_merge* m) {
  if (!m) return VD_OK;
  for (void* p : m->owned) hipFree(p);
  delete m;
  return VD_OK;
}

int vd_merge_info(const vd_merge* m, long long* sub_points) {
  VD_REQUIRE(m && sub_points, VD_ERR_INVALID_ARG, "null argument");
  *sub_points = m->total;
  return VD_OK;
}

int vd_merge_weights(const vd_merge* m, double* weights, double* weights_sum, void* stream) {
  VD_REQUIRE(m, VD_ERR_INVALID_ARG, "null merge");
  hipStream_t s = static_cast<hipStream_t>(stream);
  const size_t np = (size_t)m->M.ny * m->M.nx;
  if (weights)
    VD_HIP_TRY(hipMemcpyAsync(weights, m->M.w, np * m->M.nsub * sizeof(double), hipMemcpyDeviceToDevice, s));
  if (weights_sum)
    VD_HIP_TRY(hipMemcpyAsync(weights_sum, m->M.wsum, np * sizeof(double), hipMemcpyDeviceToDevice, s));
  return VD_OK;
}

int vd_merge_apply(vd_merge* m, int T, const double* fields, double* out, void* stream) {
  VD_REQUIRE(m && fields && out && T > 0, VD_ERR_INVALID_ARG, "null argument or T < 1");
  VD_REQUIRE(out != fields, VD_ERR_INVALID_ARG, "merge: output must not alias the fields");
  hipStream_t s = static_cast<hipStream_t>(stream);
  const int np = m->M.ny * m->M.nx;
  VD_REQUIRE((T + kMgTB - 1) / kMgTB <= 65535, VD_ERR_INVALID_ARG, "merge: too many dates in one call");
  hipLaunchKernelGGL(k_merge_apply, dim3((np + 255) / 256, (T + kMgTB - 1) / kMgTB), dim3(256), 0, s,
                     m->M, T, fields, out);
  VD_HIP_TRY(hipGetLastError());
  return VD_OK;
}

int vd_merge_blend(long long n, double w1, double w2, const double* a, const double* b, double* out,
                   void* stream) {
  VD_REQUIRE(n >= 0 && (n == 0 || (a && b && out)), VD_ERR_INVALID_ARG, "null argument or n < 0");
  if (n == 0) return VD_OK;
  hipLaunchKernelGGL(k_merge_blend, dim3(mg_blocks(n)), dim3(256), 0, static_cast<hipStream_t>(stream),
                     n, w1, w2, a, b, out);
  VD_HIP_TRY(hipGetLastError());
  return VD_OK;
}
