// 3-D wavelet basis (Basis_wavelet3d, mapping/src/basis.py:2640-2985): the
// synthesis of fields from wavelet coefficients (operg) and its transpose
// (operg_transpose) as dense FP64 MFMA products over tiles of the grid.
//
// A wave is (space frequency f, space centre P, direction, cos/sin); a
// coefficient adds a time block b (time frequency, time centre) in front:
//   field_t(p) = sum_waves G(p, w) sum_b g_f(t, b) X[f, b, w]
// where G(p, w) = sqrt(2) w(dx/DX_f) w(dy/DX_f) cos|sin(kx dx + ky dy) on the
// points of the centre's box (|dx|, |dy| <= DX_f km, unmasked) and g_f(t, b)
// the Gaspari-Cohn time factor (basis.py:2850-2934).  The space part is a
// sparse matrix in the reference (CSC, one per frequency, applied per time
// with scipy).  Here the grid is cut into 16 x 16-point tiles and every tile
// keeps, as one dense block, the columns of G of the waves whose box meets it
// (K_tau waves, padded to a multiple of 16): the operators are then dense
// products over a batch of up to 16 basis times,
//   synthesis   F_tau (16 times x 256 points) = C_tau (16 x K) G_tau (K x 256)
//   transpose   P_tau (K x 16 times)          = G_tau (K x 256) A_tau^T (256 x 16)
// on v_mfma_f64_16x16x4f64, with the coefficients contracted over the time
// blocks before (k_w3_time) and the tile partials of each wave summed in
// ascending tile order after (k_w3_proj_sum, then k_w3_time_adj).  Every sum
// has a fixed order, so results are deterministic.
//
// G_tau is built once, on the device, when the basis is created
// (k_w3_build): the host passes each tile's (family, 256-bit point
// membership) slots, where a family is one space centre with its 2 ntheta
// waves; membership is decided on the host with the reference's box test, so
// the selected points are exactly the reference's.  Storage order inside a
// tile: [p / 4][k][p % 4], so that both products read whole 128-byte lines:
// the synthesis reads 4 k x 16 consecutive points, the transpose 16 k x 4
// consecutive points.
#include <cmath>
#include <cstring>

#include "vd_gemm.h"

namespace vd {

constexpr double kW3Km2Deg = 1.0 / 110;
constexpr int kW3Tile = 16;    // grid-point tile side
constexpr int kW3Pts = 256;    // points per tile
constexpr int kW3T = 16;       // basis times per pass (the MFMA's 16 rows or columns)
constexpr int kW3Chunk = 64;   // tile entries per projection workgroup (16 per wave)
constexpr int kW3SA = 260;     // LDS row stride (doubles) of the staged adjoint fields

struct W3Tables {
  int ny, nx, nf, ntheta, ntb, nfam, ntx, nty, nwave;
  long long nbasis;
  const double* lon;       // ny * nx
  const double* lat;
  const double* dxs;       // nf
  const double* kx;        // nf * ntheta
  const double* ky;
  const int* fam_f;        // nfam
  const double* fam_lon;
  const double* fam_lat;
  const double* fam_cos;
  const long long* fbase;  // nf: first coefficient of frequency f
  const int* ffam0;        // nf: first family of frequency f
  const int* fnfam;        // nf: families of frequency f
  const long long* tslot;  // ntiles + 1: first slot of each tile
  const int* sfam;         // slots: family
  const unsigned long long* sbits;  // slots x 4: points of the tile in the family's support
  const long long* toff;   // ntiles + 1: first entry of each tile (multiples of 16)
  const int* ewave;        // entries: wave (-1: padding)
  const long long* wptr;   // nwave + 1: entries of each wave, ascending tile
  const long long* went;
  const int* jtile;        // projection jobs: tile
  const int* jk0;          //                  first entry (tile-relative)
  double* G;               // entries x 256
};

// mywindow (basis.py:4138-4140), evaluated in the reference's order
__device__ __forceinline__ double w3_window(double x) {
  const double c = cos(x * 0.5 * M_PI);
  return c * c;
}

// G_tau for every tile: one workgroup per tile, threads over (entry, point).
__global__ __launch_bounds__(256) void k_w3_build(W3Tables B) {
  const int tile = blockIdx.x;
  const int ty = tile / B.ntx, tx = tile % B.ntx;
  const long long e0 = B.toff[tile];
  const int kp = (int)(B.toff[tile + 1] - e0);
  const long long s0 = B.tslot[tile];
  const int nslot = (int)(B.tslot[tile + 1] - s0);
  const int q2 = 2 * B.ntheta;
  double* Gt = B.G + e0 * kW3Pts;
  // storage order [p / 4][k][p % 4]: consecutive threads write consecutive words
  for (int idx = threadIdx.x; idx < kp * kW3Pts; idx += blockDim.x) {
    const int p = (idx / (4 * kp)) * 4 + (idx & 3), k = (idx >> 2) % kp;
    double val = 0.0;
    const int slot = k / q2;
    if (slot < nslot) {
      const unsigned long long bits = B.sbits[(s0 + slot) * 4 + (p >> 6)];
      const int y = ty * kW3Tile + (p >> 4), x = tx * kW3Tile + (p & 15);
      if ((bits >> (p & 63)) & 1ull) {
        const int fam = B.sfam[s0 + slot];
        const int f = B.fam_f[fam];
        const int q = k % q2, th = q >> 1;
        const size_t g = (size_t)y * B.nx + x;
        const double xx = (B.lon[g] - B.fam_lon[fam]) / kW3Km2Deg * B.fam_cos[fam];
        const double yy = (B.lat[g] - B.fam_lat[fam]) / kW3Km2Deg;
        const double d = B.dxs[f];
        const double facs = w3_window(xx / d) * w3_window(yy / d);
        const double ph = B.kx[f * B.ntheta + th] * xx + B.ky[f * B.ntheta + th] * yy;
        val = (M_SQRT2 * facs) * ((q & 1) ? sin(ph) : cos(ph));
      }
    }
    Gt[idx] = val;
  }
}

// Wave of tile entry e -> (frequency, column inside the frequency's blocks).
__device__ __forceinline__ void w3_wave_col(const W3Tables& B, int w, int* f, long long* col, long long* nxf) {
  const int q2 = 2 * B.ntheta;
  const int fam = w / q2;
  const int ff = B.fam_f[fam];
  *f = ff;
  *col = (long long)(fam - B.ffam0[ff]) * q2 + (w % q2);
  *nxf = (long long)q2 * B.fnfam[ff];
}

// cg[e][t] = sum_b g_f(t0 + t, b) X[f, b, w(e)] for t < nt, else 0
// (the time contraction of operg, basis.py:2946-2952).
__global__ __launch_bounds__(256) void k_w3_time(W3Tables B, long long nent, int nt, const double* gt,
                                                 const double* X, double* cg) {
  const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= nent * kW3T) return;
  const long long e = idx / kW3T;
  const int t = (int)(idx % kW3T);
  const int w = B.ewave[e];
  double s = 0.0;
  if (w >= 0 && t < nt) {
    int f;
    long long col, nxf;
    w3_wave_col(B, w, &f, &col, &nxf);
    const double* g = gt + ((size_t)t * B.nf + f) * B.ntb;
    const double* x = X + B.fbase[f] + col;
    for (int b = 0; b < B.ntb; ++b) {
      const double gb = g[b];
      if (gb != 0.0) s += gb * x[(long long)b * nxf];
    }
  }
  cg[idx] = s;
}

// Synthesis: out[t][y][x] for t < nt.  Workgroup = (tile, 64 points: four
// rows of the tile); its four waves split the tile's entries in four K
// ranges and their 16 x 64 partials are added through LDS in wave order.
__global__ __launch_bounds__(256) void k_w3_synth(W3Tables B, int nt, const double* cg, double* out) {
  typedef Mfma<double>::acc_t acc_t;
  __shared__ double red[4][kW3T][64];
  const int tile = blockIdx.x, pg = blockIdx.y;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const long long e0 = B.toff[tile];
  const int kp = (int)(B.toff[tile + 1] - e0);
  const double* Gt = B.G + e0 * kW3Pts;
  const int kq = kp / 4 / 4 * 4;           // entries per wave: a multiple of 4 (kp is a multiple of 16)
  const int k_lo = wv * kq, k_hi = wv == 3 ? kp : k_lo + kq;
  acc_t acc[4];
#pragma unroll
  for (int j = 0; j < 4; ++j) acc[j] = acc_t{0.0, 0.0, 0.0, 0.0};
  const int r16 = lane & 15, q4 = lane >> 4;
  for (int k0 = k_lo; k0 < k_hi; k0 += 4) {
    const double a = cg[(e0 + k0 + q4) * kW3T + r16];       // A[t = r16][k = q4]
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int p = pg * 64 + j * 16 + r16;                   // B[k = q4][point = r16]
      const double b = Gt[((size_t)(p >> 2) * kp + k0 + q4) * 4 + (p & 3)];
      acc[j] = Mfma<double>::run(a, b, acc[j]);
    }
  }
#pragma unroll
  for (int j = 0; j < 4; ++j)
#pragma unroll
    for (int r = 0; r < 4; ++r) red[wv][Mfma<double>::out_row(lane, r)][j * 16 + r16] = acc[j][r];
  __syncthreads();
  const int ty = tile / B.ntx, tx = tile % B.ntx;
  for (int o = threadIdx.x; o < kW3T * 64; o += blockDim.x) {
    const int t = o >> 6, pl = o & 63;
    if (t >= nt) continue;
    const int p = pg * 64 + pl;
    const int y = ty * kW3Tile + (p >> 4), x = tx * kW3Tile + (p & 15);
    if (y >= B.ny || x >= B.nx) continue;
    const double v = ((red[0][t][pl] + red[1][t][pl]) + red[2][t][pl]) + red[3][t][pl];
    out[((size_t)t * B.ny + y) * B.nx + x] = v;
  }
}

// Transpose, per job (tile, 64 entries): P[e][t] = sum_p G(e, p) ad[t][p].
// The tile's adjoint fields of the pass (16 times x 256 points, zero where
// t >= nt or outside the grid) are staged in LDS; each wave takes 16 entries.
__global__ __launch_bounds__(256) void k_w3_proj(W3Tables B, int nt, const double* ad, double* P) {
  typedef Mfma<double>::acc_t acc_t;
  __shared__ double sa[kW3T * kW3SA];
  const int tile = B.jtile[blockIdx.x];
  const int kbase = B.jk0[blockIdx.x];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int ty = tile / B.ntx, tx = tile % B.ntx;
  for (int o = threadIdx.x; o < kW3T * kW3Pts; o += blockDim.x) {
    const int t = o >> 8, p = o & 255;
    const int y = ty * kW3Tile + (p >> 4), x = tx * kW3Tile + (p & 15);
    double v = 0.0;
    if (t < nt && y < B.ny && x < B.nx) v = ad[((size_t)t * B.ny + y) * B.nx + x];
    sa[t * kW3SA + p] = v;
  }
  __syncthreads();
  const long long e0 = B.toff[tile];
  const int kp = (int)(B.toff[tile + 1] - e0);
  const int k0 = kbase + wv * 16;
  if (k0 >= kp) return;
  const double* Gt = B.G + e0 * kW3Pts;
  const int r16 = lane & 15, q4 = lane >> 4;
  acc_t acc = acc_t{0.0, 0.0, 0.0, 0.0};
  for (int p0 = 0; p0 < kW3Pts; p0 += 4) {
    const double a = Gt[((size_t)(p0 >> 2) * kp + k0 + r16) * 4 + q4];  // A[k = r16][p = p0 + q4]
    const double b = sa[r16 * kW3SA + p0 + q4];                          // B[p = p0 + q4][t = r16]
    acc = Mfma<double>::run(a, b, acc);
  }
#pragma unroll
  for (int r = 0; r < 4; ++r) P[(e0 + k0 + Mfma<double>::out_row(lane, r)) * kW3T + r16] = acc[r];
}

// A[w][t] = sum of w's tile partials in ascending tile order.
__global__ __launch_bounds__(256) void k_w3_proj_sum(W3Tables B, const double* P, double* A) {
  const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (long long)B.nwave * kW3T) return;
  const long long w = idx / kW3T;
  const int t = (int)(idx % kW3T);
  double s = 0.0;
  for (long long j = B.wptr[w]; j < B.wptr[w + 1]; ++j) s += P[B.went[j] * kW3T + t];
  A[idx] = s;
}

// adX[f, b, col] += sum_{t < nt} g_f(t, b) A[w][t] (basis.py:2962-2985).
__global__ __launch_bounds__(256) void k_w3_time_adj(W3Tables B, int nt, const double* gt, const double* A,
                                                     double* adX) {
  const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= B.nbasis) return;
  int f = 0;
  while (f + 1 < B.nf && B.fbase[f + 1] <= c) ++f;
  const int q2 = 2 * B.ntheta;
  const long long nxf = (long long)q2 * B.fnfam[f];
  const long long r = c - B.fbase[f];
  const int b = (int)(r / nxf);
  const long long col = r % nxf;
  const long long w = (long long)(B.ffam0[f] + col / q2) * q2 + col % q2;
  double s = 0.0;
  for (int t = 0; t < nt; ++t) {
    const double g = gt[((size_t)t * B.nf + f) * B.ntb + b];
    if (g != 0.0) s += g * A[w * kW3T + t];
  }
  adX[c] += s;
}

}  // namespace vd

using namespace vd;

struct vd_w3_basis {
  int device = 0;
  W3Tables B{};
  long long nent = 0;
  int njobs = 0;
  std::vector<void*> allocs;
  ~vd_w3_basis() {
    for (void* p : allocs) (void)hipFree(p);
  }
};

namespace {

template <typename T>
int w3_upload(vd_w3_basis* b, const T* h, size_t n, const T** out) {
  void* p = nullptr;
  VD_HIP_TRY(hipMalloc(&p, std::max<size_t>(n, 1) * sizeof(T)));
  b->allocs.push_back(p);
  if (n) VD_HIP_TRY(hipMemcpy(p, h, n * sizeof(T), hipMemcpyHostToDevice));
  *out = static_cast<const T*>(p);
  return VD_OK;
}

// Workspace of one pass of 16 times: cg or P (entries x 16), then A (waves x 16).
size_t w3_ws_doubles(const vd_w3_basis* b) {
  return (size_t)b->nent * kW3T + (size_t)b->B.nwave * kW3T;
}

}  // namespace

extern "C" {

int vd_w3_create(const vd_w3_params* prm, vd_w3_basis** out) {
  VD_REQUIRE(prm && out, VD_ERR_INVALID_ARG, "null argument");
  *out = nullptr;
  const vd_w3_params& P = *prm;
  VD_REQUIRE(P.ny > 0 && P.nx > 0 && P.nf > 0 && P.ntheta > 0 && P.ntb > 0 && P.nfam >= 0, VD_ERR_SHAPE,
             "bad basis sizes");
  VD_REQUIRE(P.lon && P.lat && P.dxs && P.kx && P.ky && P.fbase && P.ffam0 && P.fnfam && P.tslot,
             VD_ERR_INVALID_ARG, "null table");
  VD_REQUIRE(P.nfam == 0 || (P.fam_f && P.fam_lon && P.fam_lat && P.fam_cos), VD_ERR_INVALID_ARG,
             "null family table");
  const int ntx = (P.nx + kW3Tile - 1) / kW3Tile, nty = (P.ny + kW3Tile - 1) / kW3Tile;
  const int ntiles = ntx * nty;
  const long long nslot = P.tslot[ntiles];
  VD_REQUIRE(P.tslot[0] == 0 && (nslot == 0 || (P.sfam && P.sbits)), VD_ERR_INVALID_ARG, "bad slot table");
  const int q2 = 2 * P.ntheta;
  // families of a frequency are consecutive; coefficient blocks contiguous
  long long nb = 0;
  int fam_next = 0;
  for (int f = 0; f < P.nf; ++f) {
    VD_REQUIRE(P.fnfam[f] >= 0 && P.ffam0[f] == fam_next, VD_ERR_INVALID_ARG, "families must be frequency-major");
    VD_REQUIRE(P.fbase[f] == nb, VD_ERR_INVALID_ARG, "frequency blocks must be contiguous");
    if (P.fnfam[f] > 0) nb += (long long)P.ntb * q2 * P.fnfam[f];
    fam_next += P.fnfam[f];
  }
  VD_REQUIRE(fam_next == P.nfam, VD_ERR_INVALID_ARG, "family count mismatch");
  for (int c = 0; c < P.nfam; ++c)
    VD_REQUIRE(P.fam_f[c] >= 0 && P.fam_f[c] < P.nf && c >= P.ffam0[P.fam_f[c]] &&
                   c < P.ffam0[P.fam_f[c]] + P.fnfam[P.fam_f[c]],
               VD_ERR_INVALID_ARG, "family frequency out of range");
  for (int t = 0; t < ntiles; ++t) {
    VD_REQUIRE(P.tslot[t + 1] >= P.tslot[t], VD_ERR_INVALID_ARG, "slot table must be ordered");
    for (long long s = P.tslot[t]; s < P.tslot[t + 1]; ++s) {
      VD_REQUIRE(P.sfam[s] >= 0 && P.sfam[s] < P.nfam, VD_ERR_INVALID_ARG, "bad family in slot table");
      VD_REQUIRE(s == P.tslot[t] || P.sfam[s] > P.sfam[s - 1], VD_ERR_INVALID_ARG,
                 "families of a tile must ascend");
    }
  }
  const long long nwave = (long long)P.nfam * q2;
  VD_REQUIRE(nwave < (1LL << 31), VD_ERR_SHAPE, "too many waves");
  // tile entries (padded to 16), the waves of each entry, projection jobs and
  // the per-wave entry lists in ascending tile order
  std::vector<long long> toff(ntiles + 1, 0);
  for (int t = 0; t < ntiles; ++t)
    toff[t + 1] = toff[t] + (long long)round_up((int)((P.tslot[t + 1] - P.tslot[t]) * q2), 16);
  const long long nent = toff[ntiles];
  std::vector<int> ewave((size_t)nent, -1), jtile, jk0;
  std::vector<long long> wcount((size_t)nwave + 1, 0);
  for (int t = 0; t < ntiles; ++t) {
    long long e = toff[t];
    for (long long s = P.tslot[t]; s < P.tslot[t + 1]; ++s)
      for (int q = 0; q < q2; ++q, ++e) {
        ewave[e] = P.sfam[s] * q2 + q;
        ++wcount[(size_t)ewave[e] + 1];
      }
    for (int k = 0; k < (int)(toff[t + 1] - toff[t]); k += kW3Chunk) {
      jtile.push_back(t);
      jk0.push_back(k);
    }
  }
  std::vector<long long> wptr((size_t)nwave + 1, 0);
  for (long long w = 0; w < nwave; ++w) wptr[w + 1] = wptr[w] + wcount[w + 1];
  std::vector<long long> went((size_t)wptr[nwave], 0);
  {
    std::vector<long long> fill(wptr.begin(), wptr.end() - 1);
    for (long long e = 0; e < nent; ++e)  // tiles ascend with e
      if (ewave[e] >= 0) went[fill[ewave[e]]++] = e;
  }
  int device = 0;
  VD_HIP_TRY(hipGetDevice(&device));
  vd_w3_basis* b = new vd_w3_basis();
  b->device = device;
  b->nent = nent;
  b->njobs = (int)jtile.size();
  W3Tables& B = b->B;
  B.ny = P.ny; B.nx = P.nx; B.nf = P.nf; B.ntheta = P.ntheta; B.ntb = P.ntb; B.nfam = P.nfam;
  B.ntx = ntx; B.nty = nty; B.nwave = (int)nwave; B.nbasis = nb;
  const size_t np = (size_t)P.ny * P.nx;
  int rc = VD_OK;
  void* g = nullptr;
  if ((rc = w3_upload(b, P.lon, np, &B.lon)) || (rc = w3_upload(b, P.lat, np, &B.lat)) ||
      (rc = w3_upload(b, P.dxs, (size_t)P.nf, &B.dxs)) ||
      (rc = w3_upload(b, P.kx, (size_t)P.nf * P.ntheta, &B.kx)) ||
      (rc = w3_upload(b, P.ky, (size_t)P.nf * P.ntheta, &B.ky)) ||
      (rc = w3_upload(b, P.fam_f, (size_t)P.nfam, &B.fam_f)) ||
      (rc = w3_upload(b, P.fam_lon, (size_t)P.nfam, &B.fam_lon)) ||
      (rc = w3_upload(b, P.fam_lat, (size_t)P.nfam, &B.fam_lat)) ||
      (rc = w3_upload(b, P.fam_cos, (size_t)P.nfam, &B.fam_cos)) ||
      (rc = w3_upload(b, reinterpret_cast<const long long*>(P.fbase), (size_t)P.nf, &B.fbase)) ||
      (rc = w3_upload(b, P.ffam0, (size_t)P.nf, &B.ffam0)) || (rc = w3_upload(b, P.fnfam, (size_t)P.nf, &B.fnfam)) ||
      (rc = w3_upload(b, reinterpret_cast<const long long*>(P.tslot), (size_t)ntiles + 1, &B.tslot)) ||
      (rc = w3_upload(b, P.sfam, (size_t)nslot, &B.sfam)) ||
      (rc = w3_upload(b, reinterpret_cast<const unsigned long long*>(P.sbits), (size_t)nslot * 4, &B.sbits)) ||
      (rc = w3_upload(b, toff.data(), toff.size(), &B.toff)) ||
      (rc = w3_upload(b, ewave.data(), ewave.size(), &B.ewave)) ||
      (rc = w3_upload(b, wptr.data(), wptr.size(), &B.wptr)) ||
      (rc = w3_upload(b, went.data(), went.size(), &B.went)) ||
      (rc = w3_upload(b, jtile.data(), jtile.size(), &B.jtile)) ||
      (rc = w3_upload(b, jk0.data(), jk0.size(), &B.jk0))) {
    delete b;
    return rc;
  }
  if (hipMalloc(&g, std::max<size_t>((size_t)nent * kW3Pts, 1) * sizeof(double)) != hipSuccess) {
    delete b;
    set_error("vd_w3_create: cannot allocate the tile blocks of G");
    return VD_ERR_HIP;
  }
  b->allocs.push_back(g);
  B.G = static_cast<double*>(g);
  if (ntiles > 0) hipLaunchKernelGGL(k_w3_build, dim3(ntiles), dim3(256), 0, 0, B);
  if (hipGetLastError() != hipSuccess || hipDeviceSynchronize() != hipSuccess) {
    delete b;
    set_error("vd_w3_create: building the tile blocks failed");
    return VD_ERR_HIP;
  }
  *out = b;
  return VD_OK;
}

int vd_w3_destroy(vd_w3_basis* basis) {
  delete basis;
  return VD_OK;
}

int vd_w3_info(const vd_w3_basis* b, long long* nbasis, long long* nentries, long long* nwave) {
  VD_REQUIRE(b, VD_ERR_INVALID_ARG, "null basis");
  if (nbasis) *nbasis = b->B.nbasis;
  if (nentries) *nentries = b->nent;
  if (nwave) *nwave = b->B.nwave;
  return VD_OK;
}

int vd_w3_workspace_size(const vd_w3_basis* b, size_t* bytes) {
  VD_REQUIRE(b && bytes, VD_ERR_INVALID_ARG, "null argument");
  *bytes = w3_ws_doubles(b) * sizeof(double);
  return VD_OK;
}

int vd_w3_operg(vd_w3_basis* b, int T, const double* gt, const double* X, double* out, void* work,
                void* stream) {
  VD_REQUIRE(b && gt && X && out && work, VD_ERR_INVALID_ARG, "null argument");
  VD_REQUIRE(T >= 1, VD_ERR_INVALID_ARG, "T must be >= 1");
  hipStream_t s = static_cast<hipStream_t>(stream);
  const W3Tables& B = b->B;
  double* cg = static_cast<double*>(work);
  const int ntiles = B.ntx * B.nty;
  for (int t0 = 0; t0 < T; t0 += kW3T) {
    const int nt = std::min(kW3T, T - t0);
    const long long n = b->nent * kW3T;
    if (n > 0) {
      hipLaunchKernelGGL(k_w3_time, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, B, b->nent, nt,
                         gt + (size_t)t0 * B.nf * B.ntb, X, cg);
    }
    hipLaunchKernelGGL(k_w3_synth, dim3(ntiles, kW3Pts / 64), dim3(256), 0, s, B, nt, cg,
                       out + (size_t)t0 * B.ny * B.nx);
  }
  VD_HIP_TRY(hipGetLastError());
  return VD_OK;
}

int vd_w3_operg_adj(vd_w3_basis* b, int T, const double* gt, const double* ad, double* adX, void* work,
                    void* stream) {
  VD_REQUIRE(b && gt && ad && adX && work, VD_ERR_INVALID_ARG, "null argument");
  VD_REQUIRE(T >= 1, VD_ERR_INVALID_ARG, "T must be >= 1");
  hipStream_t s = static_cast<hipStream_t>(stream);
  const W3Tables& B = b->B;
  double* Pp = static_cast<double*>(work);
  double* A = Pp + (size_t)b->nent * kW3T;
  for (int t0 = 0; t0 < T; t0 += kW3T) {
    const int nt = std::min(kW3T, T - t0);
    if (b->njobs > 0)
      hipLaunchKernelGGL(k_w3_proj, dim3(b->njobs), dim3(256), 0, s, B, nt, ad + (size_t)t0 * B.ny * B.nx, Pp);
    const long long nw = (long long)B.nwave * kW3T;
    if (nw > 0)
      hipLaunchKernelGGL(k_w3_proj_sum, dim3((unsigned)((nw + 255) / 256)), dim3(256), 0, s, B, Pp, A);
    if (B.nbasis > 0)
      hipLaunchKernelGGL(k_w3_time_adj, dim3((unsigned)((B.nbasis + 255) / 256)), dim3(256), 0, s, B, nt,
                         gt + (size_t)t0 * B.nf * B.ntb, A, adX);
  }
  VD_HIP_TRY(hipGetLastError());
  return VD_OK;
}

}  // extern "C"
