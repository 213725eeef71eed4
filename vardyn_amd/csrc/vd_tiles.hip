// Overlapping tiles of QgmWithTiles (mapping/models/model_qg1l/jqgm.py:1033-1258):
// gather the tiles of a field, blend tile fields back with their window
// weights, and the transposes of both (for the adjoint).
//
// Tile t = ty * ntx + tx starts at (ys[ty], xs[tx]) and spans ts x ts points;
// the tiles are the tensor product of the y and x starts, in the reference's
// order (jqgm.py:1098-1113: y outer, x inner).  Every output point loops over
// the tiles that cover it in that order, so results are deterministic and no
// atomics are needed.
#include "vd_internal.h"

namespace vd {

template <typename T>
__global__ void k_tiles_gather(int ny, int nx, int ts, int nty, int ntx, const int* __restrict__ ys,
                               const int* __restrict__ xs, const T* __restrict__ full,
                               const T* __restrict__ add, T* __restrict__ flat) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x, i = blockIdx.y * blockDim.y + threadIdx.y;
  const int t = blockIdx.z;
  if (i >= ts || j >= ts) return;
  const size_t p = (size_t)(ys[t / ntx] + i) * nx + xs[t % ntx] + j;
  flat[((size_t)t * ts + i) * ts + j] = add ? full[p] + add[p] : full[p];
}

// full(p) = sum_t w(p - start_t) flat_t(p - start_t) / wsum(p) - sub(p)
// (jqgm.py:1238-1243 with the MDT removed, jqgm.py:1241-1243).
template <typename T>
__global__ void k_tiles_blend(int ny, int nx, int ts, int nty, int ntx, const int* __restrict__ ys,
                              const int* __restrict__ xs, const T* __restrict__ w,
                              const T* __restrict__ wsum, const T* __restrict__ flat,
                              const T* __restrict__ sub, T* __restrict__ full) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x, i = blockIdx.y * blockDim.y + threadIdx.y;
  if (i >= ny || j >= nx) return;
  T acc = 0;
  for (int ty = 0; ty < nty; ++ty) {
    const int a = i - ys[ty];
    if (a < 0 || a >= ts) continue;
    for (int tx = 0; tx < ntx; ++tx) {
      const int b = j - xs[tx];
      if (b < 0 || b >= ts) continue;
      const int t = ty * ntx + tx;
      acc += w[a * ts + b] * flat[((size_t)t * ts + a) * ts + b];
    }
  }
  const size_t p = (size_t)i * nx + j;
  full[p] = sub ? acc / wsum[p] - sub[p] : acc / wsum[p];
}

// Transpose of the blend: flat_t(a, b) = w(a, b) full(p) / wsum(p), p = start_t + (a, b).
template <typename T>
__global__ void k_tiles_blend_adj(int ny, int nx, int ts, int nty, int ntx, const int* __restrict__ ys,
                                  const int* __restrict__ xs, const T* __restrict__ w,
                                  const T* __restrict__ wsum, const T* __restrict__ full,
                                  T* __restrict__ flat) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x, a = blockIdx.y * blockDim.y + threadIdx.y;
  const int t = blockIdx.z;
  if (a >= ts || b >= ts) return;
  const size_t p = (size_t)(ys[t / ntx] + a) * nx + xs[t % ntx] + b;
  flat[((size_t)t * ts + a) * ts + b] = w[a * ts + b] * full[p] / wsum[p];
}

// Transpose of the gather: full(p) = sum_t flat_t(p - start_t).
template <typename T>
__global__ void k_tiles_gather_adj(int ny, int nx, int ts, int nty, int ntx, const int* __restrict__ ys,
                                   const int* __restrict__ xs, const T* __restrict__ flat,
                                   T* __restrict__ full) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x, i = blockIdx.y * blockDim.y + threadIdx.y;
  if (i >= ny || j >= nx) return;
  T acc = 0;
  for (int ty = 0; ty < nty; ++ty) {
    const int a = i - ys[ty];
    if (a < 0 || a >= ts) continue;
    for (int tx = 0; tx < ntx; ++tx) {
      const int b = j - xs[tx];
      if (b < 0 || b >= ts) continue;
      acc += flat[((size_t)(ty * ntx + tx) * ts + a) * ts + b];
    }
  }
  full[(size_t)i * nx + j] = acc;
}

}  // namespace vd

using namespace vd;

#define VD_TILES_CHECK()                                                                         \
  VD_REQUIRE(ny >= 1 && nx >= 1 && ts >= 1 && ts <= ny && ts <= nx && nty >= 1 && ntx >= 1,      \
             VD_ERR_SHAPE, "tiles: need 1 <= tile_size <= ny, nx and at least one tile");         \
  VD_REQUIRE(precision == VD_F32 || precision == VD_F64, VD_ERR_INVALID_ARG,                     \
             "precision must be VD_F32 or VD_F64")

static inline dim3 tile_grid(int n1, int n0, int nz) { return dim3((n0 + 63) / 64, (n1 + 3) / 4, nz); }

extern "C" {

int vd_tiles_gather(int precision, int ny, int nx, int ts, int nty, int ntx, const int* ys,
                    const int* xs, const void* full, const void* add, void* flat, void* stream) {
  VD_TILES_CHECK();
  VD_REQUIRE(ys && xs && full && flat, VD_ERR_INVALID_ARG, "null argument");
  hipStream_t s = static_cast<hipStream_t>(stream);
  const dim3 g = tile_grid(ts, ts, nty * ntx), b(64, 4);
  if (precision == VD_F64)
    hipLaunchKernelGGL(k_tiles_gather<double>, g, b, 0, s, ny, nx, ts, nty, ntx, ys, xs,
                       (const double*)full, (const double*)add, (double*)flat);
  else
    hipLaunchKernelGGL(k_tiles_gather<float>, g, b, 0, s, ny, nx, ts, nty, ntx, ys, xs,
                       (const float*)full, (const float*)add, (float*)flat);
  VD_HIP_TRY(hipGetLastError());
  return VD_OK;
}

int vd_tiles_blend(int precision, int ny, int nx, int ts, int nty, int ntx, const int* ys,
                   const int* xs, const void* w, const void* wsum, const void* flat, const void* sub,
                   void* full, void* stream) {
  VD_TILES_CHECK();
  VD_REQUIRE(ys && xs && w && wsum && flat && full, VD_ERR_INVALID_ARG, "null argument");
  hipStream_t s = static_cast<hipStream_t>(stream);
  const dim3 g = tile_grid(ny, nx, 1), b(64, 4);
  if (precision == VD_F64)
    hipLaunchKernelGGL(k_tiles_blend<double>, g, b, 0, s, ny, nx, ts, nty, ntx, ys, xs,
                       (const double*)w, (const double*)wsum, (const double*)flat,
                       (const double*)sub, (double*)full);
  else
    hipLaunchKernelGGL(k_tiles_blend<float>, g, b, 0, s, ny, nx, ts, nty, ntx, ys, xs,
                       (const float*)w, (const float*)wsum, (const float*)flat, (const float*)sub,
                       (float*)full);
  VD_HIP_TRY(hipGetLastError());
  return VD_OK;
}

int vd_tiles_blend_adj(int precision, int ny, int nx, int ts, int nty, int ntx, const int* ys,
                       const int* xs, const void* w, const void* wsum, const void* full, void* flat,
                       void* stream) {
  VD_TILES_CHECK();
  VD_REQUIRE(ys && xs && w && wsum && full && flat, VD_ERR_INVALID_ARG, "null argument");
  hipStream_t s = static_cast<hipStream_t>(stream);
  const dim3 g = tile_grid(ts, ts, nty * ntx), b(64, 4);
  if (precision == VD_F64)
    hipLaunchKernelGGL(k_tiles_blend_adj<double>, g, b, 0, s, ny, nx, ts, nty, ntx, ys, xs,
                       (const double*)w, (const double*)wsum, (const double*)full, (double*)flat);
  else
    hipLaunchKernelGGL(k_tiles_blend_adj<float>, g, b, 0, s, ny, nx, ts, nty, ntx, ys, xs,
                       (const float*)w, (const float*)wsum, (const float*)full, (float*)flat);
  VD_HIP_TRY(hipGetLastError());
  return VD_OK;
}

int vd_tiles_gather_adj(int precision, int ny, int nx, int ts, int nty, int ntx, const int* ys,
                        const int* xs, const void* flat, void* full, void* stream) {
  VD_TILES_CHECK();
  VD_REQUIRE(ys && xs && flat && full, VD_ERR_INVALID_ARG, "null argument");
  hipStream_t s = static_cast<hipStream_t>(stream);
  const dim3 g = tile_grid(ny, nx, 1), b(64, 4);
  if (precision == VD_F64)
    hipLaunchKernelGGL(k_tiles_gather_adj<double>, g, b, 0, s, ny, nx, ts, nty, ntx, ys, xs,
                       (const double*)flat, (double*)full);
  else
    hipLaunchKernelGGL(k_tiles_gather_adj<float>, g, b, 0, s, ny, nx, ts, nty, ntx, ys, xs,
                       (const float*)flat, (float*)full);
  VD_HIP_TRY(hipGetLastError());
  return VD_OK;
}

}  // extern "C"
