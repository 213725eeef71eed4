// QG model kernels for gfx950: Helmholtz inversion as sine-matrix GEMMs on
// MFMA, fused RK stage stencils (forward, tangent-linear, adjoint).
//
// Reference algorithm: mapping/models/model_qg1l/jqgm.py (datlas-ocean/VarDyn)
//   h2uv :238-262, h2pv :264-293, pv2h :295-319, rhs :321-375, adv :377-396,
//   inverse_elliptic_dst :83-91 (dstI2D :75-81).
#include "vd_gemm.h"

#include <cmath>

namespace vd {

// LDS-staged kernel: each K chunk of 32 is loaded by the whole group into
// LDS (two chunks of register prefetch) and shared by its four waves.
template <typename T, int AOP, int BOP, int KIND, int MAP, int BM, int BN, int KS>
__device__ __forceinline__ void gemm_tile(const GemmArgs<T>& g, int z, int tile) {
  typedef typename Mfma<T>::acc_t acc_t;
  constexpr int LDT = kTile + 1;
  constexpr int MT = BM / 32, NT = BN / 32;          // 16x16 MFMA blocks per wave
  constexpr int ASZ = BM * LDT, BSZ = kTile * (BN + 1);
  // split-K partials of groups 1..KS-1 are parked over the whole staging block
  static_assert(KS == 1 || (KS - 1) * BM * BN <= KS * (ASZ + BSZ),
                "split-K partials must fit the staging area");
  __shared__ T sm[KS][ASZ + BSZ];
  const int tid = threadIdx.x;
  const int grp = tid >> 8;          // K group
  const int gt = tid & 255;          // thread within the group
  const int lane = tid & 63;
  const int wq = (tid >> 6) & 3;     // quadrant of the tile owned by this wave
  const int wm = wq >> 1, wn = wq & 1;
  const GemmPtrs<T> P = gemm_ptrs<T, AOP, BOP, MAP>(g, z);
  const int m0 = (tile / g.tn) * BM, n0 = (tile % g.tn) * BN;
  T* As = &sm[grp][0];               // [BM][LDT]
  T* Bs = &sm[grp][ASZ];             // [kTile][BN + 1]

  // A chunk: BM x 32, four consecutive k per thread, 32 rows per pass.
  // B chunk: 32 x BN, four consecutive n per thread, 1024/BN rows per pass.
  constexpr int APASS = BM / 32, BPASS = BN / 32, BROWS = 1024 / BN;
  const int ar = gt >> 3, ac = (gt & 7) * 4;
  const int br = gt / (BN / 4), bc = (gt % (BN / 4)) * 4;
  constexpr int NA = a_words<AOP>(), NB = b_words<BOP>();
  typedef T ABuf[APASS][4][NA];
  typedef T BBuf[BPASS][4][NB];
  // Prefetch depth: two chunks in flight when their registers take less than
  // half of what a wave may hold in a group of 256 KS threads (512 / KS
  // VGPRs), otherwise one -- deeper prefetch would spill.
  constexpr int WORDS = APASS * 4 * NA + BPASS * 4 * NB;
  constexpr int PD = 2 * WORDS * (int)(sizeof(T) / 4) < 256 / KS ? 2 : 1;
  ABuf ra[PD];
  BBuf rb[PD];
  auto fetch = [&](ABuf& a, BBuf& b, int k0) {
#pragma unroll
    for (int ps = 0; ps < APASS; ++ps)
#pragma unroll
      for (int u = 0; u < 4; ++u) fetch_a<T, AOP>(g, P.A, m0 + ar + 32 * ps, k0 + ac + u, a[ps][u]);
#pragma unroll
    for (int ps = 0; ps < BPASS; ++ps)
#pragma unroll
      for (int u = 0; u < 4; ++u)
        fetch_b<T, BOP>(g, P.B, P.Fld, P.Fsub, k0 + br + BROWS * ps, n0 + bc + u, b[ps][u]);
  };
  auto stage = [&](const ABuf& a, const BBuf& b, int k0) {
    __syncthreads();
#pragma unroll
    for (int ps = 0; ps < APASS; ++ps)
#pragma unroll
      for (int u = 0; u < 4; ++u)
        As[(ar + 32 * ps) * LDT + ac + u] = combine_a<T, AOP>(g, k0 + ac + u, a[ps][u], P.sgn);
#pragma unroll
    for (int ps = 0; ps < BPASS; ++ps)
#pragma unroll
      for (int u = 0; u < 4; ++u)
        Bs[(br + BROWS * ps) * (BN + 1) + bc + u] =
            combine_b<T, BOP>(g, k0 + br + BROWS * ps, n0 + bc + u, b[ps][u], P.sgn);
    __syncthreads();
  };

  // accumulators: MT x NT blocks; a single block keeps four K-interleaved
  // accumulators so that consecutive MFMAs do not depend on each other
  constexpr int NACC = (MT * NT == 1) ? 4 : 1;
  acc_t acc[MT][NT][NACC];
#pragma unroll
  for (int a_ = 0; a_ < MT; ++a_)
#pragma unroll
    for (int b_ = 0; b_ < NT; ++b_)
#pragma unroll
      for (int c_ = 0; c_ < NACC; ++c_)
#pragma unroll
        for (int r = 0; r < 4; ++r) acc[a_][b_][c_][r] = T(0);
  auto multiply = [&]() {
#pragma unroll
    for (int kk = 0; kk < kTile / 4; ++kk) {
      T av[MT], bv[NT];
#pragma unroll
      for (int a_ = 0; a_ < MT; ++a_)
        av[a_] = As[(wm * (BM / 2) + a_ * 16 + (lane & 15)) * LDT + kk * 4 + (lane >> 4)];
#pragma unroll
      for (int b_ = 0; b_ < NT; ++b_)
        bv[b_] = Bs[(kk * 4 + (lane >> 4)) * (BN + 1) + wn * (BN / 2) + b_ * 16 + (lane & 15)];
#pragma unroll
      for (int a_ = 0; a_ < MT; ++a_)
#pragma unroll
        for (int b_ = 0; b_ < NT; ++b_) {
          const int c_ = NACC == 4 ? (kk & 3) : 0;
          acc[a_][b_][c_] = Mfma<T>::run(av[a_], bv[b_], acc[a_][b_][c_]);
        }
    }
  };

  // Group grp takes the K chunks grp, grp+KS, grp+2KS, ...; every group runs
  // the same number of them (K is a multiple of 32*KS).  PD chunks are in
  // flight in registers while the staged one is multiplied; with PD = 2, four
  // K groups and K = 256, both of a group's chunks load before its first
  // barrier.
  const int nchunk = g.K / (kTile * KS);
  auto kof = [&](int it) { return (it * KS + grp) * kTile; };
  if constexpr (PD == 1) {
    fetch(ra[0], rb[0], kof(0));
    for (int it = 0; it < nchunk; ++it) {
      stage(ra[0], rb[0], kof(it));
      if (it + 1 < nchunk) fetch(ra[0], rb[0], kof(it + 1));
      multiply();
    }
  } else {
    fetch(ra[0], rb[0], kof(0));
    if (nchunk > 1) fetch(ra[PD - 1], rb[PD - 1], kof(1));
    for (int it = 0; it < nchunk; it += 2) {
      stage(ra[0], rb[0], kof(it));
      if (it + 2 < nchunk) fetch(ra[0], rb[0], kof(it + 2));
      multiply();
      if (it + 1 < nchunk) {
        stage(ra[PD - 1], rb[PD - 1], kof(it + 1));
        if (it + 3 < nchunk) fetch(ra[PD - 1], rb[PD - 1], kof(it + 3));
        multiply();
      }
    }
  }
  T val[MT][NT][4];
#pragma unroll
  for (int a_ = 0; a_ < MT; ++a_)
#pragma unroll
    for (int b_ = 0; b_ < NT; ++b_)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        T v = acc[a_][b_][0][r];
        if (NACC == 4) v = (acc[a_][b_][0][r] + acc[a_][b_][NACC > 1 ? 1 : 0][r]) +
                           (acc[a_][b_][NACC > 2 ? 2 : 0][r] + acc[a_][b_][NACC > 3 ? 3 : 0][r]);
        val[a_][b_][r] = v;
      }

  if (KS > 1) {
    // groups 1..KS-1 park their partial tiles in the (now idle) staging block;
    // group 0 adds them in group order
    __syncthreads();
    T* flat = &sm[0][0];
    constexpr int PER = MT * NT * 4;
    if (grp > 0) {
      T* park = flat + (size_t)(grp - 1) * BM * BN;
#pragma unroll
      for (int a_ = 0; a_ < MT; ++a_)
#pragma unroll
        for (int b_ = 0; b_ < NT; ++b_)
#pragma unroll
          for (int r = 0; r < 4; ++r)
            park[(wq * 64 + lane) * PER + (a_ * NT + b_) * 4 + r] = val[a_][b_][r];
    }
    __syncthreads();
    if (grp != 0) return;
#pragma unroll
    for (int q = 1; q < KS; ++q) {
      const T* src = flat + (size_t)(q - 1) * BM * BN;
#pragma unroll
      for (int a_ = 0; a_ < MT; ++a_)
#pragma unroll
        for (int b_ = 0; b_ < NT; ++b_)
#pragma unroll
          for (int r = 0; r < 4; ++r)
            val[a_][b_][r] += src[(wq * 64 + lane) * PER + (a_ * NT + b_) * 4 + r];
    }
  }

  // epilogue inputs of every element first, so that their loads are in flight together
  EpiIn<T> in[MT][NT][4];
#pragma unroll
  for (int a_ = 0; a_ < MT; ++a_)
#pragma unroll
    for (int b_ = 0; b_ < NT; ++b_)
#pragma unroll
      for (int r = 0; r < 4; ++r)
        in[a_][b_][r] = epi_input<T, KIND, MAP>(g, P, m0 + wm * (BM / 2) + a_ * 16 + Mfma<T>::out_row(lane, r),
                                                n0 + wn * (BN / 2) + b_ * 16 + (lane & 15));
#pragma unroll
  for (int a_ = 0; a_ < MT; ++a_)
#pragma unroll
    for (int b_ = 0; b_ < NT; ++b_)
#pragma unroll
      for (int r = 0; r < 4; ++r)
        epi_store<T, KIND, MAP>(g, P, m0 + wm * (BM / 2) + a_ * 16 + Mfma<T>::out_row(lane, r),
                                n0 + wn * (BN / 2) + b_ * 16 + (lane & 15), val[a_][b_][r], in[a_][b_][r]);
#pragma unroll
  for (int a_ = 0; a_ < MT; ++a_)
#pragma unroll
    for (int b_ = 0; b_ < NT; ++b_)
#pragma unroll
      for (int r = 0; r < 4; ++r)
        epi_ring<T, KIND, MAP>(g, P, m0 + wm * (BM / 2) + a_ * 16 + Mfma<T>::out_row(lane, r),
                               n0 + wn * (BN / 2) + b_ * 16 + (lane & 15));
}

template <typename T, int AOP, int BOP, int KIND, int MAP, int BM, int BN, int KS>
__global__ __launch_bounds__(256 * KS) void k_gemm(GemmArgs<T> g) {
  int z, tile;
  work_item(g, &z, &tile);
  gemm_tile<T, AOP, BOP, KIND, MAP, BM, BN, KS>(g, z, tile);
}

// Ragged groups: workgroup b of the launch takes tile table entry b = (member,
// product z, tile); the member's static operands come from its descriptor and
// the per-call fields from `call` at the member's offset.
template <typename T, int AOP, int BOP, int KIND, int MAP, int BM, int BN, int KS>
__global__ __launch_bounds__(256 * KS) void k_gemm_rg(RgCall<T> call, const RgGemm<T>* __restrict__ rg,
                                                      const int4* __restrict__ tiles) {
  const int4 t = tiles[blockIdx.x];
  GemmArgs<T> g = rg[t.x].g;
  const size_t fo = rg[t.x].foff;
  g.F = call.F ? call.F + fo : nullptr;
  g.Fsub = call.Fsub ? call.Fsub + fo : nullptr;
  g.hb = call.hb ? call.hb + fo : nullptr;
  g.out = call.out ? call.out + fo : nullptr;
  gemm_tile<T, AOP, BOP, KIND, MAP, BM, BN, KS>(g, t.y, t.z);
}

// Double-buffered kernel.  Per K chunk of 32 there is one barrier: while
// the four waves of a group multiply the chunk in one LDS buffer, the next
// chunk (fetched a chunk earlier) is combined into the other buffer and the
// one after is fetched.  LDS holds A row-major (As[m][k]) and B transposed
// (Bt[n][k]) with a row stride of 34 doubles / 36 floats, so that lane
// (r = l%16, q = l/16) reads its four consecutive k of a 16-wide group with
// 16-byte loads and no bank conflicts; MFMA step s pairs k = 4q + s across the
// four lane quarters on both operands.  A is fetched four consecutive k per
// thread along its rows, B one column per lane (four k per thread), so every
// global load instruction reads contiguous row segments.
template <typename T, int AOP, int BOP, int KIND, int MAP, int BM, int BN, int KS>
__global__ __launch_bounds__(256 * KS) void k_gemm2(GemmArgs<T> g) {
  typedef typename Mfma<T>::acc_t acc_t;
  constexpr int LDK = sizeof(T) == 8 ? 34 : 36;
  constexpr int MT = BM / 32, NT = BN / 32;
  constexpr int ASZ = BM * LDK, BSZ = BN * LDK;
  constexpr int STAGE = 2 * (ASZ + BSZ);  // two buffers of A and B per group
  static_assert(KS == 1 || (KS - 1) * BM * BN <= KS * STAGE, "split-K partials must fit");
  __shared__ __align__(16) T sm[KS * STAGE];
  const int tid = threadIdx.x;
  const int grp = tid >> 8, gt = tid & 255, lane = tid & 63;
  const int wq = (tid >> 6) & 3, wm = wq >> 1, wn = wq & 1;
  const int r16 = lane & 15, q4 = lane >> 4;
  int z, tile;
  work_item(g, &z, &tile);
  const GemmPtrs<T> P = gemm_ptrs<T, AOP, BOP, MAP>(g, z);
  const int m0 = (tile / g.tn) * BM, n0 = (tile % g.tn) * BN;
  T* base = sm + grp * STAGE;
  auto As = [&](int b) { return base + b * (ASZ + BSZ); };
  auto Bs = [&](int b) { return base + b * (ASZ + BSZ) + ASZ; };

  // fetch mapping: A row ar + 32 ps, k = ac..ac+3; B column bn, k = bk + 16 v + 0..3
  constexpr int APASS = BM / 32, BV = BN / 32, BKG = 256 / BN;
  const int ar = gt >> 3, ac = (gt & 7) * 4;
  const int bn = gt % BN, bk = (gt / BN) * 4;
  static_assert(4 * BKG * BV == kTile, "the B fetch mapping covers the 32 k of a chunk");
  constexpr int NA = a_words<AOP>(), NB = b_words<BOP>();
  T ra[APASS][4][NA], rb[BV][4][NB];
  auto fetch = [&](int k0) {
#pragma unroll
    for (int ps = 0; ps < APASS; ++ps)
#pragma unroll
      for (int u = 0; u < 4; ++u) fetch_a<T, AOP>(g, P.A, m0 + ar + 32 * ps, k0 + ac + u, ra[ps][u]);
#pragma unroll
    for (int v = 0; v < BV; ++v)
#pragma unroll
      for (int u = 0; u < 4; ++u)
        fetch_b<T, BOP>(g, P.B, P.Fld, P.Fsub, k0 + bk + 4 * BKG * v + u, n0 + bn, rb[v][u]);
  };
  auto store = [&](int b, int k0) {
    T* a = As(b);
    T* bt = Bs(b);
#pragma unroll
    for (int ps = 0; ps < APASS; ++ps)
#pragma unroll
      for (int u = 0; u < 4; ++u)
        a[(ar + 32 * ps) * LDK + ac + u] = combine_a<T, AOP>(g, k0 + ac + u, ra[ps][u], P.sgn);
#pragma unroll
    for (int v = 0; v < BV; ++v)
#pragma unroll
      for (int u = 0; u < 4; ++u)
        bt[bn * LDK + bk + 4 * BKG * v + u] =
            combine_b<T, BOP>(g, k0 + bk + 4 * BKG * v + u, n0 + bn, rb[v][u], P.sgn);
  };

  constexpr int NACC = (MT * NT == 1) ? 4 : 1;
  acc_t acc[MT][NT][NACC];
#pragma unroll
  for (int a_ = 0; a_ < MT; ++a_)
#pragma unroll
    for (int b_ = 0; b_ < NT; ++b_)
#pragma unroll
      for (int c_ = 0; c_ < NACC; ++c_)
#pragma unroll
        for (int r = 0; r < 4; ++r) acc[a_][b_][c_][r] = T(0);
  auto multiply = [&](int b) {
    const T* a = As(b);
    const T* bt = Bs(b);
#pragma unroll
    for (int g2 = 0; g2 < 2; ++g2) {
      T av[MT][4], bv[NT][4];
#pragma unroll
      for (int a_ = 0; a_ < MT; ++a_) {
        const T* src = a + (wm * (BM / 2) + a_ * 16 + r16) * LDK + 16 * g2 + 4 * q4;
#pragma unroll
        for (int s = 0; s < 4; ++s) av[a_][s] = src[s];
      }
#pragma unroll
      for (int b_ = 0; b_ < NT; ++b_) {
        const T* src = bt + (wn * (BN / 2) + b_ * 16 + r16) * LDK + 16 * g2 + 4 * q4;
#pragma unroll
        for (int s = 0; s < 4; ++s) bv[b_][s] = src[s];
      }
#pragma unroll
      for (int s = 0; s < 4; ++s)
#pragma unroll
        for (int a_ = 0; a_ < MT; ++a_)
#pragma unroll
          for (int b_ = 0; b_ < NT; ++b_) {
            const int c_ = NACC == 4 ? s : 0;
            acc[a_][b_][c_] = Mfma<T>::run(av[a_][s], bv[b_][s], acc[a_][b_][c_]);
          }
    }
  };

  // group grp takes chunks grp, grp + KS, ...; all groups run the same count
  const int nchunk = g.K / (kTile * KS);
  auto kof = [&](int it) { return (it * KS + grp) * kTile; };
  fetch(kof(0));
  store(0, kof(0));
  if (nchunk > 1) fetch(kof(1));
  __syncthreads();
  for (int it = 0; it < nchunk; ++it) {
    multiply(it & 1);
    if (it + 1 < nchunk) store((it + 1) & 1, kof(it + 1));
    if (it + 2 < nchunk) fetch(kof(it + 2));
    __syncthreads();
  }

  T val[MT][NT][4];
#pragma unroll
  for (int a_ = 0; a_ < MT; ++a_)
#pragma unroll
    for (int b_ = 0; b_ < NT; ++b_)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        T v = acc[a_][b_][0][r];
        if (NACC == 4) v = (acc[a_][b_][0][r] + acc[a_][b_][NACC > 1 ? 1 : 0][r]) +
                           (acc[a_][b_][NACC > 2 ? 2 : 0][r] + acc[a_][b_][NACC > 3 ? 3 : 0][r]);
        val[a_][b_][r] = v;
      }
  if (KS > 1) {
    // the loop ended on a barrier, so the staging area is free for the partials
    constexpr int PER = MT * NT * 4;
    if (grp > 0) {
      T* park = sm + (size_t)(grp - 1) * 256 * PER;
#pragma unroll
      for (int a_ = 0; a_ < MT; ++a_)
#pragma unroll
        for (int b_ = 0; b_ < NT; ++b_)
#pragma unroll
          for (int r = 0; r < 4; ++r) park[((a_ * NT + b_) * 4 + r) * 256 + gt] = val[a_][b_][r];
    }
    __syncthreads();
    if (grp != 0) return;
#pragma unroll
    for (int qg = 1; qg < KS; ++qg) {
      const T* src = sm + (size_t)(qg - 1) * 256 * PER;
#pragma unroll
      for (int a_ = 0; a_ < MT; ++a_)
#pragma unroll
        for (int b_ = 0; b_ < NT; ++b_)
#pragma unroll
          for (int r = 0; r < 4; ++r) val[a_][b_][r] += src[((a_ * NT + b_) * 4 + r) * 256 + gt];
    }
  }
  // epilogue inputs of every element first, so that their loads are in flight together
  EpiIn<T> in[MT][NT][4];
#pragma unroll
  for (int a_ = 0; a_ < MT; ++a_)
#pragma unroll
    for (int b_ = 0; b_ < NT; ++b_)
#pragma unroll
      for (int r = 0; r < 4; ++r)
        in[a_][b_][r] = epi_input<T, KIND, MAP>(g, P, m0 + wm * (BM / 2) + a_ * 16 + Mfma<T>::out_row(lane, r),
                                                n0 + wn * (BN / 2) + b_ * 16 + r16);
#pragma unroll
  for (int a_ = 0; a_ < MT; ++a_)
#pragma unroll
    for (int b_ = 0; b_ < NT; ++b_)
#pragma unroll
      for (int r = 0; r < 4; ++r)
        epi_store<T, KIND, MAP>(g, P, m0 + wm * (BM / 2) + a_ * 16 + Mfma<T>::out_row(lane, r),
                                n0 + wn * (BN / 2) + b_ * 16 + r16, val[a_][b_][r], in[a_][b_][r]);
#pragma unroll
  for (int a_ = 0; a_ < MT; ++a_)
#pragma unroll
    for (int b_ = 0; b_ < NT; ++b_)
#pragma unroll
      for (int r = 0; r < 4; ++r)
        epi_ring<T, KIND, MAP>(g, P, m0 + wm * (BM / 2) + a_ * 16 + Mfma<T>::out_row(lane, r),
                               n0 + wn * (BN / 2) + b_ * 16 + r16);
}

template <typename T, int AOP, int BOP, int KIND, int MAP>
static void gemm(GemmArgs<T> a, int M, int N, int nz, int cfg, hipStream_t s) {
  // nz = members x parities.  Fall back to configuration 0 when the shape
  // does not tile with the requested one.
  if (cfg < 0 || cfg >= kNumTileCfg || !tile_cfg_fits(cfg, M, N, a.K)) cfg = 0;
  int ks = kTileCfgs[cfg].ks;
  if (cfg == 0) {
    const int nchunk = a.K / kTile;
    ks = nchunk % 4 == 0 ? 4 : (nchunk % 2 == 0 ? 2 : 1);
  }
  const int bm = kTileCfgs[cfg].bm, bn = kTileCfgs[cfg].bn;
  a.nz = nz;
  a.tm = M / bm;
  a.tn = N / bn;
  const dim3 grid(nz * a.tm * a.tn), block(256 * ks);
#define VD_G(...) hipLaunchKernelGGL((__VA_ARGS__), grid, block, 0, s, a)
  switch (cfg) {
    case 0:
      if (ks == 4) VD_G(k_gemm<T, AOP, BOP, KIND, MAP, 32, 32, 4>);
      else if (ks == 2) VD_G(k_gemm<T, AOP, BOP, KIND, MAP, 32, 32, 2>);
      else VD_G(k_gemm<T, AOP, BOP, KIND, MAP, 32, 32, 1>);
      break;
    case 1: VD_G(k_gemm<T, AOP, BOP, KIND, MAP, 64, 64, 1>); break;
    case 2: VD_G(k_gemm<T, AOP, BOP, KIND, MAP, 64, 64, 2>); break;
    case 3: VD_G(k_gemm<T, AOP, BOP, KIND, MAP, 128, 64, 1>); break;
    case 4: VD_G(k_gemm<T, AOP, BOP, KIND, MAP, 128, 64, 2>); break;
    case 5: VD_G(k_gemm2<T, AOP, BOP, KIND, MAP, 32, 32, 4>); break;
    case 6: VD_G(k_gemm2<T, AOP, BOP, KIND, MAP, 64, 64, 1>); break;
    case 7: VD_G(k_gemm2<T, AOP, BOP, KIND, MAP, 64, 64, 2>); break;
    case 8: VD_G(k_gemm2<T, AOP, BOP, KIND, MAP, 128, 64, 1>); break;
    case 9: VD_G(k_gemm2<T, AOP, BOP, KIND, MAP, 64, 32, 2>); break;
    default: VD_G(k_gemm<T, AOP, BOP, KIND, MAP, 128, 128, 1>); break;
  }
#undef VD_G
}

template <typename T>
struct UniformLaunch {
  GemmArgs<T> a;
  int M, N, nz, cfg;
  hipStream_t s;
  template <int AOP, int BOP, int KIND, int MAP>
  void go() { gemm<T, AOP, BOP, KIND, MAP>(a, M, N, nz, cfg, s); }
};

template <typename T>
void launch_gemm_stage(const Plan& p, int stage, const T* F, const T* Fsub, const T* hb, T* out,
                       int mode, int cfg, hipStream_t s) {
  if constexpr (sizeof(T) == 8) {
    if (p.split_ns > 0) {
      launch_split_stage(p, stage, F, Fsub, hb, out, mode, cfg, s);
      return;
    }
  }
  UniformLaunch<T> l;
  l.a = stage_args<T>(p, stage);
  if (stage == 0) { l.a.F = F; l.a.Fsub = Fsub; }
  if (stage == 3) { l.a.out = out; l.a.hb = hb; }
  int K;
  gemm_stage_shape(p, stage, &l.M, &l.N, &K, &l.nz);
  l.cfg = cfg;
  l.s = s;
  dispatch_stage(p.fold != 0, stage, Fsub != nullptr, mode, l);
}

template <typename T>
void launch_inversion(const Plan& p, const T* F, const T* Fsub, const T* hb, T* out, int mode,
                      hipStream_t s) {
  for (int st = 0; st < 4; ++st) launch_gemm_stage<T>(p, st, F, Fsub, hb, out, mode, p.gemm_cfg[st], s);
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

// View of a field for the stencil helpers: the value at offset (di, dj) from
// the view's point of a row-major array (a global field or an LDS tile).
template <typename T>
struct PView {
  const T* p;
  int pitch;
  __device__ __forceinline__ T operator()(int di, int dj) const { return p[di * pitch + dj]; }
};
// Upwind difference sums of adv (jqgm.py:387-394), without the 1/(6 d) factor.
template <typename T>
struct Diffs {
  T xp, xm, yp, ym;
};
template <typename T, typename V>
__device__ __forceinline__ Diffs<T> diffs_v(const V& q) {
  const T c = q(0, 0);
  const T e1 = q(0, 1), e2 = q(0, 2);
  const T w1 = q(0, -1), w2 = q(0, -2);
  const T n1 = q(1, 0), n2 = q(2, 0);
  const T s1 = q(-1, 0), s2 = q(-2, 0);
  Diffs<T> d;
  d.xp = 2 * e1 + 3 * c - 6 * w1 + w2;
  d.xm = e2 - 6 * e1 + 3 * c + 2 * w1;
  d.yp = 2 * n1 + 3 * c - 6 * s1 + s2;
  d.ym = n2 - 6 * n1 + 3 * c + 2 * s1;
  return d;
}
template <typename T>
__device__ __forceinline__ Diffs<T> diffs(const T* q, int nx, int i, int j) {
  return diffs_v<T>(PView<T>{q + i * nx + j, nx});
}

template <typename T>
struct Vel {
  T uT, vT;
};
// u at column offset dj of the view's row, v at row offset di of its column
// (u_at / v_at of the view's field)
template <typename T, typename V>
__device__ __forceinline__ T u_v(const V& h, int dj, T cu) {
  return nz(cu * (h(1, dj - 1) + h(1, dj) - h(-1, dj) - h(-1, dj - 1)));
}
template <typename T, typename V>
__device__ __forceinline__ T v_v(const V& h, int di, T cv) {
  return nz(cv * (h(di, 1) + h(di - 1, 1) - h(di - 1, -1) - h(di, -1)));
}
template <typename T, typename V>
__device__ __forceinline__ Vel<T> vel_v(const V& h, const Phys<T>& P) {
  Vel<T> r;
  r.uT = T(0.5) * (u_v<T>(h, 0, P.cu) + u_v<T>(h, 1, P.cu));
  r.vT = T(0.5) * (v_v<T>(h, 0, P.cv) + v_v<T>(h, 1, P.cv));
  return r;
}
template <typename T>
__device__ __forceinline__ Vel<T> vel_T(const T* h, int nx, int i, int j, const Phys<T>& P) {
  return vel_v<T>(PView<T>{h + i * nx + j, nx}, P);
}
template <typename T, typename V>
__device__ __forceinline__ bool u_isnan_v(const V& h) {
  return isnan(h(1, -1) + h(1, 0) + h(-1, 0) + h(-1, -1));
}
template <typename T, typename V>
__device__ __forceinline__ bool v_isnan_v(const V& h) {
  return isnan(h(0, 1) + h(-1, 1) + h(-1, -1) + h(0, -1));
}

// rhs (jqgm.py:321-375) at an active point (mask == 3, so 2 <= i,j <= n-3),
// with h and q read through views of row pitch `pitch` at (li, lj) -- the
// fields themselves (pitch nx, (li, lj) = (i, j)) or an LDS tile -- and the
// optional beta and f*bathymetry fields at (i, j).
template <typename T>
__device__ __forceinline__ T rhs_view(const T* h, const T* q, int pitch, int li, int lj, int i,
                                      int j, const Phys<T>& P) {
  const int nx = P.nx;
  const Vel<T> V = vel_T(h, pitch, li, lj, P);
  const T up = V.uT < 0 ? T(0) : V.uT, um = V.uT > 0 ? T(0) : V.uT;
  const T vp = V.vT < 0 ? T(0) : V.vT, vm = V.vT > 0 ? T(0) : V.vT;
  const Diffs<T> D = diffs(q, pitch, li, lj);
  T k = (um * D.xm - up * D.xp) * P.inv6dx + (vm * D.ym - vp * D.yp) * P.inv6dy;
  if (P.fbathy) {
    const Diffs<T> B = diffs(P.fbathy, nx, i, j);
    k += (um * B.xm - up * B.xp) * P.inv6dx + (vm * B.ym - vp * B.yp) * P.inv6dy;
  }
  if (P.beta) k -= P.beta[i * nx + j] * V.vT;
  if (P.has_kdiff) {
    const T c = q[li * pitch + lj];
    k += P.kx * (q[li * pitch + lj + 1] + q[li * pitch + lj - 1] - 2 * c) +
         P.ky * (q[(li + 1) * pitch + lj] + q[(li - 1) * pitch + lj] - 2 * c);
  }
  return nz(k);
}
template <typename T>
__device__ __forceinline__ T rhs_at(const T* h, const T* q, int i, int j, const Phys<T>& P) {
  return rhs_view(h, q, P.nx, i, j, i, j, P);
}

// One thread per point; blockIdx.z is the batch member and zo the offset of
// its (ny, nx) field inside every batched array.
#define VD_POINT_2D                                    \
  const int j = blockIdx.x * blockDim.x + threadIdx.x; \
  const int i = blockIdx.y * blockDim.y + threadIdx.y; \
  if (i >= P.ny || j >= P.nx) return;                  \
  const int nx = P.nx;                                 \
  const size_t o = (size_t)i * nx + j;                 \
  const size_t zo = (size_t)blockIdx.z * P.nf;         \
  (void)zo;

template <typename Q>
__device__ __forceinline__ Q* bofs(Q* p, size_t zo) { return p ? p + zo : p; }
#define VD_B(p) p = bofs(p, zo)

static inline dim3 grid2d(int ny, int nx, int nb) { return dim3((nx + 63) / 64, (ny + 3) / 4, nb); }
static const dim3 kBlock2d(64, 4);

// ---------------------------------------------------------------------------
// Forward kernels
// ---------------------------------------------------------------------------
// h2pv (jqgm.py:264-293) at point (i, j) of field h with ring values from hb.
template <typename T>
__device__ __forceinline__ T h2pv_at(const T* h, const T* hb, int i, int j, uint8_t m,
                                     const Phys<T>& P) {
  const int nx = P.nx;
  if (m == 0 || m == 5) return T(0);
  if (m != 3 && m != 4) return -P.gfc2_b[blockIdx.z] * hb[i * nx + j];
  const T c = h[i * nx + j];
  const T lap = P.gf_dy2 * (h[(i + 1) * nx + j] + h[(i - 1) * nx + j] - 2 * c) +
                P.gf_dx2 * (h[i * nx + j + 1] + h[i * nx + j - 1] - 2 * c);
  return nz(lap - P.gfc2_b[blockIdx.z] * c);
}

// Point-wise bodies of the segment start/end kernels, shared by the plan
// kernels (grid over columns, rows and batch members) and the ragged-group
// kernels (tile table over members).  Field pointers are already offset to the
// member's field; hbm and qb of the segment start may be NULL (not wanted).
template <typename T>
__device__ __forceinline__ void prep_point(const Phys<T>& P, int i, int j, const T* h0,
                                           const T* hb, const T* mdt, T* hcur, T* hbm, T* q0,
                                           T* qb) {
  const int nx = P.nx;
  const size_t o = (size_t)i * nx + j;
  const uint8_t m = P.mask[o];
  if (mdt) {
    // q needs h + mdt at neighbours: evaluate the stencil on the fly.
    auto H = [&](int ii, int jj) { return h0[ii * nx + jj] + mdt[ii * nx + jj]; };
    auto HB = [&](int ii, int jj) { return hb[ii * nx + jj] + mdt[ii * nx + jj]; };
    hcur[o] = H(i, j);
    if (hbm) hbm[o] = HB(i, j);
    T qv = 0, qbv = 0;
    if (m == 3 || m == 4) {
      const T c = H(i, j), cb = HB(i, j);
      qv = nz(P.gf_dy2 * (H(i + 1, j) + H(i - 1, j) - 2 * c) +
              P.gf_dx2 * (H(i, j + 1) + H(i, j - 1) - 2 * c) - P.gfc2_b[blockIdx.z] * c);
      qbv = nz(P.gf_dy2 * (HB(i + 1, j) + HB(i - 1, j) - 2 * cb) +
               P.gf_dx2 * (HB(i, j + 1) + HB(i, j - 1) - 2 * cb) - P.gfc2_b[blockIdx.z] * cb);
    } else if (m == 1 || m == 2) {
      qv = -P.gfc2_b[blockIdx.z] * HB(i, j);
      qbv = qv;
    }
    q0[o] = qv;
    if (qb) qb[o] = qbv;
  } else {
    hcur[o] = h0[o];
    if (hbm) hbm[o] = hb[o];
    q0[o] = h2pv_at(h0, hb, i, j, m, P);
    if (qb) qb[o] = h2pv_at(hb, hb, i, j, m, P);
  }
}

template <typename T>
__global__ void k_prep(Phys<T> P, const T* __restrict__ h0, const T* __restrict__ hb,
                       const T* __restrict__ mdt, T* __restrict__ hcur, T* __restrict__ hbm,
                       T* __restrict__ q0, T* __restrict__ qb) {
  VD_POINT_2D
  VD_B(h0); VD_B(hb); VD_B(hcur); VD_B(hbm); VD_B(q0); VD_B(qb);
  prep_point(P, i, j, h0, hb, mdt, hcur, hbm, q0, qb);
}

template <typename T>
__global__ void k_h2pv(Phys<T> P, const T* __restrict__ h, const T* __restrict__ hb,
                       T* __restrict__ q) {
  VD_POINT_2D
  VD_B(h); VD_B(hb); VD_B(q);
  q[o] = h2pv_at(h, hb, i, j, P.mask[o], P);
}

template <typename T>
__global__ void k_fill_land(Phys<T> P, T* __restrict__ f, T value) {
  VD_POINT_2D
  VD_B(f);
  if (P.mask[o] == 0) f[o] = value;
}

template <typename T>
__device__ __forceinline__ void finish_point(const Phys<T>& P, int i, int j, const T* h,
                                             const T* mdt, T* h1) {
  const size_t o = (size_t)i * P.nx + j;
  h1[o] = P.mask[o] == 0 ? T(NAN) : (mdt ? h[o] - mdt[o] : h[o]);
}

template <typename T>
__global__ void k_finish(Phys<T> P, const T* __restrict__ h, const T* __restrict__ mdt,
                         T* __restrict__ h1) {
  VD_POINT_2D
  VD_B(h); VD_B(h1);
  finish_point(P, i, j, h, mdt, h1);
}

// Fused RK stage: qout = qbase + c0 k0 + c1 k1 + c2 rhs(h, q) (rhs stored in
// kstore when given); the mask selects rhs (active points) or zero.
// LDS-tiled: a 64 x kSY tile of points per workgroup (64 x 4 threads, kSY / 4
// rows each; 8 rows where the fields fit in cache, 16 rows on HBM-bound
// batches, where the taller tile halves the halo rows read per point: each
// measured faster than the other there).  h and q of the tile and its 2-point halo are loaded once into
// LDS (rows and columns clamped into the field, so always in bounds); every
// stencil read then comes from LDS, which leaves HBM as the bound (one global
// read per point of h and q instead of about twenty).  Points outside the
// active band (every active point lies in rows/columns 2..n-3) read halo
// values that the mask then discards; the optional beta and f*bathymetry
// fields are read at the point clamped into that band.
// kSY rows for launches whose fields fit in the Infinity Cache (and the
// ragged-group tables), kSYB for banded, HBM-bound launches (below).
constexpr int kSX = 64, kSY = 8, kSYB = 16, kSP = kSX + 4;
constexpr int kSCX = kSX / 64;  // columns per thread (64 threads across)
template <typename T, int kSY>
__device__ __forceinline__ void stage_tile(const Phys<T>& P, int i0, int j0, const T* h,
                                           const T* q, const T* qbase, const T* k0,
                                           const T* k1, T c0, T c1, T c2, T* qout, T* kstore) {
  __shared__ T hs[(kSY + 4) * kSP];
  __shared__ T qs[(kSY + 4) * kSP];
  const int nx = P.nx, ny = P.ny;
  const int tid = threadIdx.x + 64 * threadIdx.y;
  // this thread's point-wise inputs, issued before the tile loads so that
  // they are in flight together (clamped: always in bounds).  q_base and the
  // output are streamed once per launch: non-temporal load and store.
  T base[kSCX][kSY / 4];
  uint8_t msk[kSCX][kSY / 4];
#pragma unroll
  for (int cx = 0; cx < kSCX; ++cx) {
    const int j = j0 + threadIdx.x + 64 * cx;
#pragma unroll
    for (int rr = 0; rr < kSY / 4; ++rr) {
      const size_t o = (size_t)min(i0 + (int)threadIdx.y + 4 * rr, ny - 1) * nx + min(j, nx - 1);
      base[cx][rr] = __builtin_nontemporal_load(&qbase[o]);
      msk[cx][rr] = P.mask[o];
    }
  }
  for (int e = tid; e < (kSY + 4) * kSP; e += 256) {
    const int r = e / kSP, c = e % kSP;
    const size_t g = (size_t)min(max(i0 - 2 + r, 0), ny - 1) * nx + min(max(j0 - 2 + c, 0), nx - 1);
    hs[e] = h[g];
    qs[e] = q[g];
  }
  __syncthreads();
#pragma unroll
  for (int cx = 0; cx < kSCX; ++cx) {
    const int lj = threadIdx.x + 64 * cx, j = j0 + lj;
    if (j >= nx) break;
#pragma unroll
    for (int rr = 0; rr < kSY / 4; ++rr) {
      const int li = threadIdx.y + 4 * rr, i = i0 + li;
      if (i >= ny) break;
      const size_t o = (size_t)i * nx + j;
      const int ic = min(max(i, 2), ny - 3), jc = min(max(j, 2), nx - 3);  // global optional fields
      const T kk = rhs_view(hs, qs, kSP, li + 2, lj + 2, ic, jc, P);
      const T k = msk[cx][rr] == 3 ? kk : T(0);
      if (kstore) kstore[o] = k;
      T v = base[cx][rr];
      if (k0) v += c0 * k0[o];
      if (k1) v += c1 * k1[o];
      __builtin_nontemporal_store(v + c2 * k, &qout[o]);
    }
  }
}

// Banded XCD order of the stencil launches.  The GPU deals the workgroups of
// a 1-D grid round-robin to its 8 XCDs (block b on XCD b % 8), each with its
// own L2.  In row-major tile order the left and right neighbours of a tile
// run on other XCDs, so the cache lines of their shared halo columns are
// fetched from memory twice (measured: 1.34 x the algorithmic reads at 64 x
// 512^2).  Here the 8 XCDs work on one problem at a time, XCD x taking the
// x-th contiguous eighth of its row-major tiles: a band of rows in which
// every left/right and most upper/lower neighbours share its L2 (measured:
// 1.04 x), while all XCDs still stream through the same problem together.
// The last band of a problem may hold fewer tiles; its spare blocks exit.
// The order pays where the fields stream from HBM; when they fit in the
// 256 MB Infinity Cache, the plain (columns, rows, members) grid measured
// faster (8 x 512^2: adjoint stage 53.4 against 55.7 us), so the launcher
// picks it by working set (band = 0: the tile is the 3-D block index).
template <bool BAND>
__device__ __forceinline__ bool band_tile(int gx, int gy, int* bx, int* by, int* bz) {
  if constexpr (!BAND) {
    *bx = blockIdx.x;
    *by = blockIdx.y;
    *bz = blockIdx.z;
    return true;
  } else {
    const int T = gx * gy, G8 = (T + 7) >> 3;
    const int b = blockIdx.x, x = b & 7, k = b >> 3;
    const int z = k / G8, t = x * G8 + (k - z * G8);
    if (t >= T) return false;
    *bx = t % gx;
    *by = t / gx;
    *bz = z;
    return true;
  }
}
static inline dim3 band_grid(int band, int gx, int gy, int nb) {
  return band ? dim3(8 * ((gx * gy + 7) / 8) * nb) : dim3(gx, gy, nb);
}
// Banded order when the fields a launch touches exceed the Infinity Cache.
static inline int band_for(const Plan& p, int fields, size_t elem) {
  return (double)p.nb * (double)p.nf * (double)elem * fields > 256.0 * (1 << 20) ? 1 : 0;
}

template <typename T, int KSY, bool BAND>
__global__ __launch_bounds__(256) void k_stage(Phys<T> P, const T* __restrict__ h,
                                               const T* __restrict__ q,
                                               const T* __restrict__ qbase,
                                               const T* __restrict__ k0,
                                               const T* __restrict__ k1, T c0, T c1, T c2,
                                               T* __restrict__ qout, T* __restrict__ kstore) {
  int bx, by, bz;
  if (!band_tile<BAND>((P.nx + kSX - 1) / kSX, (P.ny + KSY - 1) / KSY, &bx, &by, &bz)) return;
  const size_t zo = (size_t)bz * P.nf;
  VD_B(h); VD_B(q); VD_B(qbase); VD_B(k0); VD_B(k1); VD_B(qout); VD_B(kstore);
  stage_tile<T, KSY>(P, by * KSY, bx * kSX, h, q, qbase, k0, k1, c0, c1, c2, qout, kstore);
}

// Ragged groups: block b takes tile (member, tile row, tile column) of the
// table; the member's constants and field offset come from its descriptor.
template <typename T>
__global__ __launch_bounds__(256) void k_stage_rg(const RgMember<T>* __restrict__ mem,
                                                  const int4* __restrict__ tiles,
                                                  const T* __restrict__ h,
                                                  const T* __restrict__ q,
                                                  const T* __restrict__ qbase,
                                                  const T* __restrict__ k0,
                                                  const T* __restrict__ k1, T c0, T c1, T c2,
                                                  T* __restrict__ qout, T* __restrict__ kstore) {
  const int4 t = tiles[blockIdx.x];
  if (t.x < 0) return;  // padding entry of a banded member
  const Phys<T> P = mem[t.x].P;
  const size_t zo = mem[t.x].foff;
  VD_B(h); VD_B(q); VD_B(qbase); VD_B(k0); VD_B(k1); VD_B(qout); VD_B(kstore);
  stage_tile<T, kSY>(P, t.y * kSY, t.z * kSX, h, q, qbase, k0, k1, c0, c1, c2, qout, kstore);
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
  VD_B(h); VD_B(q); VD_B(dh); VD_B(dq); VD_B(dqbase); VD_B(dk0); VD_B(dk1); VD_B(dqout); VD_B(dkstore);
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
  VD_B(dh0); VD_B(hcur); VD_B(dhcur); VD_B(dq0);
  dhcur[o] = dh0[o];
  T r = T(0);
  if ((P.mask[o] == 3 || P.mask[o] == 4) && !q_isnan(hcur, nx, i, j)) {
    const T c = dh0[o];
    r = P.gf_dy2 * (dh0[o + nx] + dh0[o - nx] - 2 * c) +
        P.gf_dx2 * (dh0[o + 1] + dh0[o - 1] - 2 * c) - P.gfc2_b[blockIdx.z] * c;
  }
  dq0[o] = r;
}

template <typename T>
__global__ void k_finish_tl(Phys<T> P, const T* __restrict__ dh, T* __restrict__ dh1) {
  VD_POINT_2D
  VD_B(dh); VD_B(dh1);
  dh1[o] = P.mask[o] == 0 ? T(0) : dh[o];
}

// ---------------------------------------------------------------------------
// Adjoint kernels (deterministic gathers, no atomics).
// ---------------------------------------------------------------------------
// Adjoint terms at an active T point t = (i, j) (mask == 3), for the
// cotangent ak = a0 s0 + a1 s1 + a2 s2 of rhs at t:
//   w[0] = lambda_uT, w[1] = lambda_vT (with the beta term),
//   w[2] = uT ak/(6dx), w[3] = vT ak/(6dy), w[4] = ak,
// and the upwind side of each velocity (pu: uT > 0, pv: vT > 0).  The
// transposed upwind stencils weight q by up ak/(6dx) = (pu ? w[2] : 0) and
// um ak/(6dx) = (pu ? 0 : w[2]) (both are zero where uT == 0), so one value
// and a flag carry both.  h and q are read through views of row pitch
// `pitch` at (hi, hj) (an LDS tile); s0..s2, beta and f*bathymetry at the
// field point (i, j).
template <typename T, typename HV, typename QV>
__device__ __forceinline__ void adj_terms(const HV& h, const QV& q, const T* s0, const T* s1,
                                          const T* s2, T a0, T a1, T a2, int i, int j,
                                          const Phys<T>& P, T w[5], bool* pu, bool* pv) {
  const int nx = P.nx;
  const size_t o = (size_t)i * nx + j;
  T ak = a0 * s0[o];
  if (s1) ak += a1 * s1[o];
  if (s2) ak += a2 * s2[o];
  const Vel<T> V = vel_v<T>(h, P);
  const Diffs<T> D = diffs_v<T>(q);
  T dxp = D.xp, dxm = D.xm, dyp = D.yp, dym = D.ym;
  if (P.fbathy) {
    const Diffs<T> B = diffs(P.fbathy, nx, i, j);
    dxp += B.xp; dxm += B.xm; dyp += B.yp; dym += B.ym;
  }
  const T aux = ak * P.inv6dx, auy = ak * P.inv6dy;
  w[0] = (V.uT < 0 ? T(0) : -aux * dxp) + (V.uT > 0 ? T(0) : aux * dxm);
  w[1] = (V.vT < 0 ? T(0) : -auy * dyp) + (V.vT > 0 ? T(0) : auy * dym);
  if (P.beta) w[1] -= P.beta[o] * ak;
  w[2] = V.uT * aux;
  w[3] = V.vT * auy;
  w[4] = ak;
  *pu = V.uT > 0;
  *pv = V.vT > 0;
}

// Fused adjoint stage (one launch, deterministic, no atomics).  A block owns
// an AX x AY tile of output points.  Phase 0 loads h and q of the tile plus a
// halo of 4 into LDS (rows and columns clamped into the field: in bounds, and
// only read where the true neighbours exist).  Phase 1 evaluates adj_terms on
// the tile plus a halo of 2 into LDS (zero where the point is inactive or
// outside the grid), together with the validity of u and v at every tile
// point (inside their domains and not NaN, jqgm.py:238-262), reading h and q
// from LDS.  Phase 2 gathers, for each output point p,
//   lambda_q(p) from the T points whose upwind / diffusion stencils read q(p),
//   lambda_h(p) from lambda_u, lambda_v at the velocity points that read h(p).
// The tile is AX x AY points (AX x AY threads): 64 x 16 in general, 32 x 32
// for ragged groups whose members are all at most 32 points wide (the tiles
// of QgmWithTiles), where a 64-point-wide tile would leave half its threads
// idle.
constexpr int kAX = 64, kAY = 16, kAH = 2;
constexpr int kAXN = 32, kAYN = 32;  // narrow members

// Phase 2 of an adjoint stage at output point p: lambda_q(p) (added to lq)
// from the T points whose upwind / diffusion stencils read q(p), lambda_h(p)
// from lambda_u, lambda_v at the velocity points that read h(p).  W(k, di,
// dj) is adjoint term k at offset (di, dj) from p (0: lambda_uT, 1: lambda_vT,
// 2: uT ak/(6dx), 3: vT ak/(6dy), 4: ak) and F(di, dj) its flags (bit 0 u
// valid, 1 v valid, 2 uT > 0, 3 vT > 0).
template <typename T, bool KD, typename WV, typename FV>
__device__ __forceinline__ void adj_gather(const Phys<T>& P, const WV& W, const FV& F, T lq,
                                           T* lq_out, T* lh_out) {
  // the up / um (vp / vm) weights at offset (a, b)
  auto wp = [&](int k, int bit, int a, int b) -> T { return (F(a, b) & bit) ? W(k, a, b) : T(0); };
  auto wm = [&](int k, int bit, int a, int b) -> T { return (F(a, b) & bit) ? T(0) : W(k, a, b); };
  // transpose of the x / y upwind stencils: T point t = p + d reads q(p)
  lq += (-2) * wp(2, 4, 0, -1) + (-6) * wm(2, 4, 0, -1) + (-3) * wp(2, 4, 0, 0) +
        3 * wm(2, 4, 0, 0) + 6 * wp(2, 4, 0, 1) + 2 * wm(2, 4, 0, 1) - wp(2, 4, 0, 2) +
        wm(2, 4, 0, -2);
  lq += (-2) * wp(3, 8, -1, 0) + (-6) * wm(3, 8, -1, 0) + (-3) * wp(3, 8, 0, 0) +
        3 * wm(3, 8, 0, 0) + 6 * wp(3, 8, 1, 0) + 2 * wm(3, 8, 1, 0) - wp(3, 8, 2, 0) +
        wm(3, 8, -2, 0);
  if constexpr (KD) {
    if (P.has_kdiff) {
      const T c = W(4, 0, 0);
      lq += P.kx * (W(4, 0, 1) + W(4, 0, -1) - 2 * c) + P.ky * (W(4, 1, 0) + W(4, -1, 0) - 2 * c);
    }
  }
  *lq_out = lq;
  // lambda_u at u point (a, b) = 0.5 (W0(a, b) + W0(a, b-1)) where u is valid
  auto lu = [&](int a, int b) -> T {
    return (F(a, b) & 1) ? T(0.5) * (W(0, a, b) + W(0, a, b - 1)) : T(0);
  };
  auto lv = [&](int a, int b) -> T {
    return (F(a, b) & 2) ? T(0.5) * (W(1, a, b) + W(1, a - 1, b)) : T(0);
  };
  *lh_out = P.cu * (lu(-1, 1) + lu(-1, 0) - lu(1, 0) - lu(1, 1)) +
            P.cv * (lv(0, -1) + lv(1, -1) - lv(1, 1) - lv(0, 1));
}

template <typename T, int AX = kAX, int AY = kAY, bool KD = true>
__device__ __forceinline__ void adj_tile(const Phys<T>& P, int i0, int j0, const T* h, const T* q,
                                         const T* s0, const T* s1, const T* s2, T a0, T a1, T a2,
                                         const T* acc0, const T* acc1, const T* acc2, T* lq_out,
                                         T* lh_out) {
  // acc* may alias lq_out (in-place accumulation): each thread reads and
  // writes only its own point of those fields.  s0..s2 are read at the halo
  // points of the tile, which other blocks write, so they must not alias
  // lq_out.  KD: the plan has PV diffusion (the ak array is kept for its
  // transposed Laplacian).
  constexpr int kRW = AX + 2 * kAH, kRH = AY + 2 * kAH;
  constexpr int kHW = AX + 4 * kAH, kHH = AY + 4 * kAH;  // h, q tile: halo 4
  constexpr int NWA = KD ? 5 : 4;
  __shared__ T W[NWA][kRH][kRW];
  __shared__ T Hs[kHH * kHW], Qs[kHH * kHW];
  // bit 0: u valid, 1: v valid, 2: uT > 0, 3: vT > 0 at each tile point
  __shared__ uint8_t fl[kRH][kRW];
  const int ny = P.ny, nx = P.nx;
  const int tid = threadIdx.y * AX + threadIdx.x;
  for (int e = tid; e < kHH * kHW; e += AX * AY) {
    const int r = e / kHW, c = e - r * kHW;
    const size_t g = (size_t)min(max(i0 - 2 * kAH + r, 0), ny - 1) * nx +
                     min(max(j0 - 2 * kAH + c, 0), nx - 1);
    Hs[e] = h[g];
    Qs[e] = q[g];
  }
  __syncthreads();
  for (int r = tid; r < kRH * kRW; r += AX * AY) {
    const int li = r / kRW, lj = r - li * kRW;
    const int i = i0 - kAH + li, j = j0 - kAH + lj;
    const int hi = li + kAH, hj = lj + kAH;  // the same point in the h, q tile
    const PView<T> hv{Hs + hi * kHW + hj, kHW}, qv{Qs + hi * kHW + hj, kHW};
    T w[5] = {0, 0, 0, 0, 0};
    bool pu = false, pv = false;
    if (i >= 0 && i < ny && j >= 0 && j < nx && P.mask[(size_t)i * nx + j] == 3)
      adj_terms(hv, qv, s0, s1, s2, a0, a1, a2, i, j, P, w, &pu, &pv);
#pragma unroll
    for (int k = 0; k < NWA; ++k) W[k][li][lj] = w[k];
    const bool oku = i >= 1 && i <= ny - 2 && j >= 1 && j <= nx - 1 && !u_isnan_v<T>(hv);
    const bool okv = i >= 1 && i <= ny - 1 && j >= 1 && j <= nx - 2 && !v_isnan_v<T>(hv);
    fl[li][lj] = (oku ? 1 : 0) | (okv ? 2 : 0) | (pu ? 4 : 0) | (pv ? 8 : 0);
  }
  __syncthreads();
  const int i = i0 + threadIdx.y, j = j0 + threadIdx.x;
  if (i >= ny || j >= nx) return;
  const size_t o = (size_t)i * nx + j;
  const int li = threadIdx.y + kAH, lj = threadIdx.x + kAH;
  T lq = 0;
  if (acc0) lq += acc0[o];
  if (acc1) lq += acc1[o];
  if (acc2) lq += acc2[o];
  auto Wv = [&](int k, int a, int b) -> T { return W[k < NWA ? k : 0][li + a][lj + b]; };
  auto Fv = [&](int a, int b) -> int { return fl[li + a][lj + b]; };
  adj_gather<T, KD>(P, Wv, Fv, lq, lq_out + o, lh_out + o);
}

// Without diffusion the FP64 tile takes 77 KB of LDS, so two workgroups share
// a CU when each thread keeps to 64 VGPRs (eight waves per SIMD).
template <typename T, bool BAND, bool KD>
__global__ __launch_bounds__(kAX * kAY) __attribute__((amdgpu_waves_per_eu(KD ? 4 : 8)))
void k_adj_fused(
    Phys<T> P, const T* __restrict__ h, const T* __restrict__ q, const T* __restrict__ s0,
    const T* __restrict__ s1, const T* __restrict__ s2, T a0, T a1, T a2, const T* acc0,
    const T* acc1, const T* acc2, T* lq_out, T* __restrict__ lh_out) {
  int bx, by, bz;
  if (!band_tile<BAND>((P.nx + kAX - 1) / kAX, (P.ny + kAY - 1) / kAY, &bx, &by, &bz)) return;
  const size_t zo = (size_t)bz * P.nf;
  VD_B(h); VD_B(q); VD_B(s0); VD_B(s1); VD_B(s2); VD_B(acc0); VD_B(acc1); VD_B(acc2);
  VD_B(lq_out); VD_B(lh_out);
  adj_tile<T, kAX, kAY, KD>(P, by * kAY, bx * kAX, h, q, s0, s1, s2, a0, a1, a2, acc0, acc1, acc2,
                            lq_out, lh_out);
}

template <typename T, int AX, int AY>
__global__ __launch_bounds__(AX * AY) void k_adj_fused_rg(
    const RgMember<T>* __restrict__ mem, const int4* __restrict__ tiles, const T* __restrict__ h,
    const T* __restrict__ q, const T* __restrict__ s0, const T* __restrict__ s1,
    const T* __restrict__ s2, T a0, T a1, T a2, const T* acc0, const T* acc1, const T* acc2,
    T* lq_out, T* __restrict__ lh_out) {
  const int4 t = tiles[blockIdx.x];
  if (t.x < 0) return;  // padding entry of a banded member
  const Phys<T> P = mem[t.x].P;
  const size_t zo = mem[t.x].foff;
  VD_B(h); VD_B(q); VD_B(s0); VD_B(s1); VD_B(s2); VD_B(acc0); VD_B(acc1); VD_B(acc2);
  VD_B(lq_out); VD_B(lh_out);
  adj_tile<T, AX, AY>(P, t.y * AY, t.z * AX, h, q, s0, s1, s2, a0, a1, a2, acc0, acc1, acc2,
                      lq_out, lh_out);
}


template <typename T>
__device__ __forceinline__ void adj_init_point(const Phys<T>& P, int i, int j, const T* adh1,
                                               T* lh, T* lq) {
  const size_t o = (size_t)i * P.nx + j;
  lh[o] = P.mask[o] == 0 ? T(0) : adh1[o];
  lq[o] = T(0);
}

template <typename T>
__global__ void k_adj_init(Phys<T> P, const T* __restrict__ adh1, T* __restrict__ lh,
                           T* __restrict__ lq) {
  VD_POINT_2D
  VD_B(adh1); VD_B(lh); VD_B(lq);
  adj_init_point(P, i, j, adh1, lh, lq);
}

template <typename T>
__device__ __forceinline__ void adj_final_point(const Phys<T>& P, int i, int j, const T* hcur,
                                                const T* lh, const T* lq, T* adh0) {
  const int nx = P.nx, ny = P.ny;
  const size_t o = (size_t)i * nx + j;
  // a(t) = lq(t) where h2pv used its Laplacian formula (mask 3 or 4, q not NaN)
  auto A = [&](int a, int b) -> T {
    if (a < 1 || a > ny - 2 || b < 1 || b > nx - 2) return T(0);
    const size_t t = (size_t)a * nx + b;
    if ((P.mask[t] != 3 && P.mask[t] != 4) || q_isnan(hcur, nx, a, b)) return T(0);
    return lq[t];
  };
  const T r = P.gf_dy2 * (A(i + 1, j) + A(i - 1, j)) + P.gf_dx2 * (A(i, j + 1) + A(i, j - 1)) +
              (-2 * P.gf_dy2 - 2 * P.gf_dx2 - P.gfc2_b[blockIdx.z]) * A(i, j);
  adh0[o] = lh[o] + r;
}

template <typename T>
__global__ void k_adj_final(Phys<T> P, const T* __restrict__ hcur, const T* __restrict__ lh,
                            const T* __restrict__ lq, T* __restrict__ adh0) {
  VD_POINT_2D
  VD_B(hcur); VD_B(lh); VD_B(lq); VD_B(adh0);
  adj_final_point(P, i, j, hcur, lh, lq, adh0);
}

// Ragged-group versions of the segment start/end kernels: block b takes the
// kSX x kSY point tile stage_tiles[b] of its member (64 x 4 threads) with the member's constants, MDT and field offset.  Every point runs
// the same body as in the member's own plan; the member's gfc2_b has one
// entry and blockIdx.z is 0 here, so the bodies read the member's value.
#define VD_POINT_RG                                              \
  const int4 t = tiles[blockIdx.x];                              \
  if (t.x < 0) return;                                           \
  const Phys<T> P = mem[t.x].P;                                  \
  const T* mdt = mem[t.x].mdt;                                   \
  (void)mdt;                                                     \
  const size_t zo = mem[t.x].foff;
#define VD_POINT_RG_LOOP(BODY)                                   \
  for (int cx = 0; cx < kSCX; ++cx) {                            \
    const int j = t.z * kSX + threadIdx.x + 64 * cx;             \
    _Pragma("unroll") for (int rr = 0; rr < kSY / 4; ++rr) {     \
      const int i = t.y * kSY + threadIdx.y + 4 * rr;            \
      if (i < P.ny && j < P.nx) BODY;                            \
    }                                                            \
  }

template <typename T>
__global__ __launch_bounds__(256) void k_prep_rg(const RgMember<T>* __restrict__ mem,
                                                 const int4* __restrict__ tiles, const T* h0,
                                                 const T* hb, T* hcur, T* hbm, T* q0, T* qb) {
  VD_POINT_RG
  VD_B(h0); VD_B(hb); VD_B(hcur); VD_B(hbm); VD_B(q0); VD_B(qb);
  VD_POINT_RG_LOOP(prep_point(P, i, j, h0, hb, mdt, hcur, hbm, q0, qb))
}

template <typename T>
__global__ __launch_bounds__(256) void k_finish_rg(const RgMember<T>* __restrict__ mem,
                                                   const int4* __restrict__ tiles, const T* h,
                                                   T* h1) {
  VD_POINT_RG
  VD_B(h); VD_B(h1);
  VD_POINT_RG_LOOP(finish_point(P, i, j, h, mdt, h1))
}

template <typename T>
__global__ __launch_bounds__(256) void k_adj_init_rg(const RgMember<T>* __restrict__ mem,
                                                     const int4* __restrict__ tiles,
                                                     const T* adh1, T* lh, T* lq) {
  VD_POINT_RG
  VD_B(adh1); VD_B(lh); VD_B(lq);
  VD_POINT_RG_LOOP(adj_init_point(P, i, j, adh1, lh, lq))
}

template <typename T>
__global__ __launch_bounds__(256) void k_adj_final_rg(const RgMember<T>* __restrict__ mem,
                                                      const int4* __restrict__ tiles,
                                                      const T* hcur, const T* lh, const T* lq,
                                                      T* adh0) {
  VD_POINT_RG
  VD_B(hcur); VD_B(lh); VD_B(lq); VD_B(adh0);
  VD_POINT_RG_LOOP(adj_final_point(P, i, j, hcur, lh, lq, adh0))
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
  P.gfc2_b = static_cast<const T*>(p.gfc2_b);
  P.nf = p.nf;
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
  hipLaunchKernelGGL(k_prep<T>, grid2d(p.ny, p.nx, p.nb), kBlock2d, 0, s, P, h0, hb, mdt, hcur, hbm,
                     q0, qb);
}

template <typename T>
void launch_h2pv(const Plan& p, const T* h, const T* hb, T* q, hipStream_t s) {
  Phys<T> P = make_phys<T>(p);
  hipLaunchKernelGGL(k_h2pv<T>, grid2d(p.ny, p.nx, p.nb), kBlock2d, 0, s, P, h, hb, q);
}

template <typename T>
void launch_fill_land(const Plan& p, T* f, T value, hipStream_t s) {
  Phys<T> P = make_phys<T>(p);
  hipLaunchKernelGGL(k_fill_land<T>, grid2d(p.ny, p.nx, p.nb), kBlock2d, 0, s, P, f, value);
}

template <typename T>
void launch_finish(const Plan& p, const T* h, T* h1, hipStream_t s) {
  Phys<T> P = make_phys<T>(p);
  const T* mdt = p.has_mdt ? static_cast<const T*>(p.mdt) : nullptr;
  hipLaunchKernelGGL(k_finish<T>, grid2d(p.ny, p.nx, p.nb), kBlock2d, 0, s, P, h, mdt, h1);
}

template <typename T>
void launch_stage(const Plan& p, const T* h, const T* q, const T* qbase, const T* kprev0,
                  const T* kprev1, T c0, T c1, T c2, T* qout, T* kstore, hipStream_t s) {
  Phys<T> P = make_phys<T>(p);
  const int band = band_for(p, 4, sizeof(T));  // h, q, q_base in; q_out
  const int gx = (p.nx + kSX - 1) / kSX;
  if (band)
    hipLaunchKernelGGL((k_stage<T, kSYB, true>), band_grid(1, gx, (p.ny + kSYB - 1) / kSYB, p.nb),
                       kBlock2d, 0, s, P, h, q, qbase, kprev0, kprev1, c0, c1, c2, qout, kstore);
  else
    hipLaunchKernelGGL((k_stage<T, kSY, false>), band_grid(0, gx, (p.ny + kSY - 1) / kSY, p.nb),
                       kBlock2d, 0, s, P, h, q, qbase, kprev0, kprev1, c0, c1, c2, qout, kstore);
}

template <typename T>
void launch_stage_tl(const Plan& p, const T* h, const T* q, const T* dh, const T* dq,
                     const T* dqbase, const T* dkprev0, const T* dkprev1, T c0, T c1, T c2,
                     T* dqout, T* dkstore, hipStream_t s) {
  Phys<T> P = make_phys<T>(p);
  hipLaunchKernelGGL(k_stage_tl<T>, grid2d(p.ny, p.nx, p.nb), kBlock2d, 0, s, P, h, q, dh, dq, dqbase,
                     dkprev0, dkprev1, c0, c1, c2, dqout, dkstore);
}

template <typename T>
void launch_prep_tl(const Plan& p, const T* dh0, const T* hcur, T* dhcur, T* dq0, hipStream_t s) {
  Phys<T> P = make_phys<T>(p);
  hipLaunchKernelGGL(k_prep_tl<T>, grid2d(p.ny, p.nx, p.nb), kBlock2d, 0, s, P, dh0, hcur, dhcur, dq0);
}

template <typename T>
void launch_finish_tl(const Plan& p, const T* dh, T* dh1, hipStream_t s) {
  Phys<T> P = make_phys<T>(p);
  hipLaunchKernelGGL(k_finish_tl<T>, grid2d(p.ny, p.nx, p.nb), kBlock2d, 0, s, P, dh, dh1);
}

template <typename T>
void launch_stage_adj(const Plan& p, const T* h, const T* q, const T* s0, const T* s1,
                      const T* s2, T a0, T a1, T a2, const T* acc0, const T* acc1,
                      const T* acc2, T* lq_out, T* lh_out, hipStream_t s) {
  Phys<T> P = make_phys<T>(p);
  const int band = band_for(p, 6, sizeof(T));  // h, q, cotangent, accumulator in; lq, lh out
  const dim3 grid = band_grid(band, (p.nx + kAX - 1) / kAX, (p.ny + kAY - 1) / kAY, p.nb);
#define VD_ADJ(BAND, KD)                                                                        \
  hipLaunchKernelGGL((k_adj_fused<T, BAND, KD>), grid, dim3(kAX, kAY), 0, s, P, h, q, s0, s1, s2, \
                     a0, a1, a2, acc0, acc1, acc2, lq_out, lh_out)
  if (band) {
    if (p.has_kdiff) VD_ADJ(true, true);
    else VD_ADJ(true, false);
  } else {
    if (p.has_kdiff) VD_ADJ(false, true);
    else VD_ADJ(false, false);
  }
#undef VD_ADJ
}

template <typename T>
void launch_adj_init(const Plan& p, const T* adh1, T* lh, T* lq, hipStream_t s) {
  Phys<T> P = make_phys<T>(p);
  hipLaunchKernelGGL(k_adj_init<T>, grid2d(p.ny, p.nx, p.nb), kBlock2d, 0, s, P, adh1, lh, lq);
}

template <typename T>
void launch_adj_final(const Plan& p, const T* hcur, const T* lh, const T* lq, T* adh0,
                      hipStream_t s) {
  Phys<T> P = make_phys<T>(p);
  hipLaunchKernelGGL(k_adj_final<T>, grid2d(p.ny, p.nx, p.nb), kBlock2d, 0, s, P, hcur, lh, lq, adh0);
}

// ---------------------------------------------------------------------------
// Ragged groups
// ---------------------------------------------------------------------------
template <typename T>
struct RaggedLaunch {
  RgCall<T> call;
  const RgGemm<T>* rg;
  const int4* tiles;
  int ntiles, ks;
  hipStream_t s;
  template <int AOP, int BOP, int KIND, int MAP>
  void go() {
    const dim3 grid(ntiles), block(256 * ks);
    if (ks == 4)
      hipLaunchKernelGGL((k_gemm_rg<T, AOP, BOP, KIND, MAP, 32, 32, 4>), grid, block, 0, s, call, rg, tiles);
    else if (ks == 2)
      hipLaunchKernelGGL((k_gemm_rg<T, AOP, BOP, KIND, MAP, 32, 32, 2>), grid, block, 0, s, call, rg, tiles);
    else
      hipLaunchKernelGGL((k_gemm_rg<T, AOP, BOP, KIND, MAP, 32, 32, 1>), grid, block, 0, s, call, rg, tiles);
  }
};

template <typename T>
void launch_stage_rg(const Group& g, int nact, const T* h, const T* q, const T* qbase,
                     const T* kprev0, const T* kprev1, T c0, T c1, T c2, T* qout, T* kstore,
                     hipStream_t s) {
  const int nb = g.stage_prefix[nact];
  if (nb == 0) return;
  hipLaunchKernelGGL(k_stage_rg<T>, dim3(nb), kBlock2d, 0, s, static_cast<const RgMember<T>*>(g.mem),
                     g.stage_tiles, h, q, qbase, kprev0, kprev1, c0, c1, c2, qout, kstore);
}

#define VD_RG_LAUNCH(KERNEL, ...)                                                      \
  const int nb = g.stage_prefix[nact];                                                 \
  if (nb == 0) return;                                                                 \
  hipLaunchKernelGGL(KERNEL<T>, dim3(nb), kBlock2d, 0, s,                              \
                     static_cast<const RgMember<T>*>(g.mem), g.stage_tiles, __VA_ARGS__)

template <typename T>
void launch_prep_rg(const Group& g, int nact, const T* h0, const T* hb, T* hcur, T* hbm, T* q0,
                    T* qb, hipStream_t s) {
  VD_RG_LAUNCH(k_prep_rg, h0, hb, hcur, hbm, q0, qb);
}
template <typename T>
void launch_finish_rg(const Group& g, int nact, const T* h, T* h1, hipStream_t s) {
  VD_RG_LAUNCH(k_finish_rg, h, h1);
}
template <typename T>
void launch_adj_init_rg(const Group& g, int nact, const T* adh1, T* lh, T* lq, hipStream_t s) {
  VD_RG_LAUNCH(k_adj_init_rg, adh1, lh, lq);
}
template <typename T>
void launch_adj_final_rg(const Group& g, int nact, const T* hcur, const T* lh, const T* lq,
                         T* adh0, hipStream_t s) {
  VD_RG_LAUNCH(k_adj_final_rg, hcur, lh, lq, adh0);
}

template <typename T>
void launch_stage_adj_rg(const Group& g, int nact, const T* h, const T* q, const T* s0,
                         const T* s1, const T* s2, T a0, T a1, T a2, const T* acc0,
                         const T* acc1, const T* acc2, T* lq_out, T* lh_out, hipStream_t s) {
  const int sh = g.adj_narrow ? 1 : 0;
  const int nb = g.adj_prefix[sh][nact];
  if (nb == 0) return;
  if (g.adj_narrow)
    hipLaunchKernelGGL((k_adj_fused_rg<T, kAXN, kAYN>), dim3(nb), dim3(kAXN, kAYN), 0, s,
                       static_cast<const RgMember<T>*>(g.mem), g.adj_tiles[1], h, q, s0, s1, s2, a0, a1,
                       a2, acc0, acc1, acc2, lq_out, lh_out);
  else
    hipLaunchKernelGGL((k_adj_fused_rg<T, kAX, kAY>), dim3(nb), dim3(kAX, kAY), 0, s,
                       static_cast<const RgMember<T>*>(g.mem), g.adj_tiles[0], h, q, s0, s1, s2, a0, a1,
                       a2, acc0, acc1, acc2, lq_out, lh_out);
}

// Helmholtz inversion of a small member in one workgroup (ragged groups of
// equal-size members with interiors of at most kSmallM points, the tiles of
// QgmWithTiles): the four sine products and the division by Lambda run on
// the member's interior held in LDS, instead of four GEMM launches whose
// 32 x 32 tiles would be mostly padding.  Per member:
//   A = F - Fsub, B = S_y A, A = (B S_x) / Lambda, B = S_y A,
//   out = hb + B S_x (interior; ring = hb) or out += B S_x (accumulate).
// Lambda = gof (lcx[j] + lcy[i]) - gfc2 as in the GEMM epilogue.
constexpr int kSmallM = 45;  // 2 m_y m_x + m_y^2 + m_x^2 doubles fit 64 KB of LDS
struct RgSmall {
  size_t foff;
  const double* lcx;
  const double* lcy;
  double gof, gfc2;
};

// C[i][j] = sum_k L[i][k] R[k][j] (M x N, row-major operands in LDS), each
// thread four rows of one column: the L values are the same for every lane
// of a row block (broadcast reads) and each R value is read once for four
// multiply-adds.  Every output sums over k in order with FMA.
template <typename F>
__device__ __forceinline__ void small_mm4(const double* L, int ldl, const double* R, int ldr,
                                          int K, int M, int N, F&& store) {
  const int nb = (M + 3) / 4;
  for (int e = threadIdx.x; e < nb * N; e += 256) {
    const int i0 = 4 * (e / N), j = e % N;
    double acc[4] = {0, 0, 0, 0};
    for (int k = 0; k < K; ++k) {
      const double r = R[k * ldr + j];
#pragma unroll
      for (int u = 0; u < 4; ++u)
        if (i0 + u < M) acc[u] = fma(L[(i0 + u) * ldl + k], r, acc[u]);
    }
#pragma unroll
    for (int u = 0; u < 4; ++u)
      if (i0 + u < M) store(i0 + u, j, acc[u]);
  }
}

template <int MODE>
__global__ __launch_bounds__(256) void k_inv_small(const RgSmall* __restrict__ mem,
                                                   const double* __restrict__ Sy, int ldy,
                                                   const double* __restrict__ Sx, int ldx, int my,
                                                   int mx, const double* F, const double* Fsub,
                                                   const double* hb, double* out) {
  extern __shared__ double sm[];
  double* A = sm;
  double* B = A + my * mx;
  double* SY = B + my * mx;
  double* SX = SY + my * my;
  const RgSmall m = mem[blockIdx.x];
  const int nx = mx + 2, ny = my + 2, tid = threadIdx.x;
  F += m.foff;
  if (Fsub) Fsub += m.foff;
  if (hb) hb += m.foff;
  out += m.foff;
  for (int e = tid; e < my * my; e += 256) SY[e] = Sy[(size_t)(e / my) * ldy + e % my];
  for (int e = tid; e < mx * mx; e += 256) SX[e] = Sx[(size_t)(e / mx) * ldx + e % mx];
  for (int e = tid; e < my * mx; e += 256) {
    const size_t o = (size_t)(e / mx + 1) * nx + e % mx + 1;
    A[e] = Fsub ? F[o] - Fsub[o] : F[o];
  }
  __syncthreads();
  // B = S_y A
  small_mm4(SY, my, A, mx, my, my, mx, [&](int i, int j, double v) { B[i * mx + j] = v; });
  __syncthreads();
  // A = (B S_x) / Lambda
  small_mm4(B, mx, SX, mx, mx, my, mx, [&](int i, int j, double v) {
    double lam;
    {
#pragma clang fp contract(off)
      lam = m.gof * (m.lcx[j] + m.lcy[i]) - m.gfc2;
    }
    A[i * mx + j] = v * (1.0 / lam);
  });
  __syncthreads();
  // B = S_y A
  small_mm4(SY, my, A, mx, my, my, mx, [&](int i, int j, double v) { B[i * mx + j] = v; });
  __syncthreads();
  // out = hb + B S_x, or out += B S_x
  small_mm4(B, mx, SX, mx, mx, my, mx, [&](int i, int j, double v) {
    const size_t o = (size_t)(i + 1) * nx + j + 1;
    if (MODE == INV_ACC)
      out[o] += v;
    else
      out[o] = hb ? hb[o] + v : v;
  });
  if (MODE != INV_ACC)  // the ring from hb
    for (int e = tid; e < 2 * nx + 2 * my; e += 256) {
      const size_t o = e < nx ? (size_t)e
                     : e < 2 * nx ? (size_t)(ny - 1) * nx + (e - nx)
                     : e < 2 * nx + my ? (size_t)(e - 2 * nx + 1) * nx
                                       : (size_t)(e - 2 * nx - my + 1) * nx + nx - 1;
      out[o] = hb ? hb[o] : 0.0;
    }
}

template <typename T>
void launch_inversion_rg(const Group& g, int nact, const T* F, const T* Fsub, const T* hb, T* out,
                         int mode, hipStream_t s) {
  if constexpr (sizeof(T) == 8) {
    if (g.split_ns > 0) {
      launch_inversion_split_rg(g, nact, F, Fsub, hb, out, mode, s);
      return;
    }
    if (g.use_small && mode != INV_TL) {
      if (nact == 0) return;
      const Plan& p0 = *g.plans[0];
      const size_t lds = (size_t)(2 * p0.my * p0.mx + p0.my * p0.my + p0.mx * p0.mx) * 8;
      if (mode == INV_ACC)
        hipLaunchKernelGGL(k_inv_small<INV_ACC>, dim3(nact), dim3(256), lds, s,
                           static_cast<const RgSmall*>(g.small_mem), static_cast<const double*>(p0.Sy),
                           p0.MY, static_cast<const double*>(p0.Sx), p0.MX, p0.my, p0.mx, F, Fsub, hb,
                           out);
      else
        hipLaunchKernelGGL(k_inv_small<INV_FWD>, dim3(nact), dim3(256), lds, s,
                           static_cast<const RgSmall*>(g.small_mem), static_cast<const double*>(p0.Sy),
                           p0.MY, static_cast<const double*>(p0.Sx), p0.MX, p0.my, p0.mx, F, Fsub, hb,
                           out);
      return;
    }
  }
  static const int kKs[3] = {4, 2, 1};
  for (int st = 0; st < 4; ++st) {
    for (int c = 0; c < 3; ++c) {
      RaggedLaunch<T> l;
      l.call = RgCall<T>{st == 0 ? F : nullptr, st == 0 ? Fsub : nullptr, st == 3 ? hb : nullptr,
                         st == 3 ? out : nullptr};
      l.rg = static_cast<const RgGemm<T>*>(g.gargs[st]);
      l.tiles = g.gtiles[st][c];
      l.ntiles = g.gprefix[st][c][nact];
      l.ks = kKs[c];
      l.s = s;
      if (l.ntiles > 0) dispatch_stage(g.fold != 0, st, Fsub != nullptr, mode, l);
    }
  }
}

template <typename V>
static int upload_vec(const std::vector<V>& v, void** d) {
  VD_HIP_TRY(hipMalloc(d, std::max<size_t>(1, v.size()) * sizeof(V)));
  if (!v.empty()) VD_HIP_TRY(hipMemcpy(*d, v.data(), v.size() * sizeof(V), hipMemcpyHostToDevice));
  return VD_OK;
}

// Stencil tile table of member b with gx x gy tiles.  With band set, members
// of at least 64 tiles are laid out in the banded XCD order of the
// single-problem launches (table entry e runs on XCD e % 8): padding entries
// (member -1) first align the member to a multiple of 8, then the 8 bands
// interleave.  Otherwise, and for smaller members, row-major order.
static void rg_tiles(int b, int gx, int gy, bool band, std::vector<int4>& out) {
  const int T = gx * gy;
  if (!band || T < 64) {
    for (int ty = 0; ty < gy; ++ty)
      for (int tx = 0; tx < gx; ++tx) out.push_back(make_int4(b, ty, tx, 0));
    return;
  }
  while (out.size() % 8) out.push_back(make_int4(-1, 0, 0, 0));
  const int G8 = (T + 7) / 8;
  for (int k = 0; k < G8; ++k)
    for (int x = 0; x < 8; ++x) {
      const int t = x * G8 + k;
      out.push_back(t < T ? make_int4(b, t / gx, t % gx, 0) : make_int4(-1, 0, 0, 0));
    }
}

template <typename T>
int group_build(Group& g) {
  const int n = g.n;
  g.foff.assign(n + 1, 0);
  std::vector<RgMember<T>> mem(n);
  std::vector<int4> st, at;
  // banded stencil order when the members' fields (four per stage launch)
  // exceed the 256 MB Infinity Cache, as for single-problem launches
  double total = 0;
  for (int b = 0; b < n; ++b) total += (double)g.plans[b]->nf * sizeof(T) * 4;
  const bool band = total > 256.0 * (1 << 20);
  // adjoint-stage tiles: 32 x 32 by default when every member is at most 32
  // points wide (a 64-wide tile would leave half its threads idle)
  g.adj_narrow = true;
  for (int b = 0; b < n; ++b) g.adj_narrow = g.adj_narrow && g.plans[b]->nx <= kAXN;
  std::vector<int4> at2;
  g.stage_prefix.assign(n + 1, 0);
  g.adj_prefix[0].assign(n + 1, 0);
  g.adj_prefix[1].assign(n + 1, 0);
  for (int b = 0; b < n; ++b) {
    const Plan& p = *g.plans[b];
    g.foff[b + 1] = g.foff[b] + p.nf;
    mem[b].P = make_phys<T>(p);
    mem[b].foff = g.foff[b];
    mem[b].mdt = p.has_mdt ? static_cast<const T*>(p.mdt) : nullptr;
    rg_tiles(b, (p.nx + kSX - 1) / kSX, (p.ny + kSY - 1) / kSY, band, st);
    rg_tiles(b, (p.nx + kAX - 1) / kAX, (p.ny + kAY - 1) / kAY, band, at);
    rg_tiles(b, (p.nx + kAXN - 1) / kAXN, (p.ny + kAYN - 1) / kAYN, band, at2);
    g.stage_prefix[b + 1] = (int)st.size();
    g.adj_prefix[0][b + 1] = (int)at.size();
    g.adj_prefix[1][b + 1] = (int)at2.size();
  }
  int rc;
  if ((rc = upload_vec(mem, &g.mem)) != VD_OK) return rc;
  if ((rc = upload_vec(st, (void**)&g.stage_tiles)) != VD_OK) return rc;
  if ((rc = upload_vec(at, (void**)&g.adj_tiles[0])) != VD_OK) return rc;
  if ((rc = upload_vec(at2, (void**)&g.adj_tiles[1])) != VD_OK) return rc;
  if (g.split_ns > 0) return split_group_build(g);
  // small equal-size FP64 members: fused inversion in one workgroup each
  if constexpr (sizeof(T) == 8) {
    const Plan& p0 = *g.plans[0];
    bool small = p0.my <= kSmallM && p0.mx <= kSmallM && p0.nb == 1;
    for (int b = 0; b < n && small; ++b)
      small = g.plans[b]->ny == p0.ny && g.plans[b]->nx == p0.nx && g.plans[b]->nb == 1;
    if (small) {
      std::vector<RgSmall> sm(n);
      for (int b = 0; b < n; ++b) {
        const Plan& p = *g.plans[b];
        double gfc2 = 0;
        VD_HIP_TRY(hipMemcpy(&gfc2, p.gfc2_d, sizeof(double), hipMemcpyDeviceToHost));
        sm[b] = RgSmall{g.foff[b], p.lam_cx, p.lam_cy, p.g / p.f0, gfc2};
      }
      if ((rc = upload_vec(sm, &g.small_mem)) != VD_OK) return rc;
      g.small = g.use_small = true;
    }
  }
  for (int stg = 0; stg < 4; ++stg) {
    // 32 x 32 tiles with configuration 0's K-group count for each member (4, 2
    // or 1 from its chunk count), so that every member computes exactly what
    // its own plan computes; one tile table per K-split class
    std::vector<RgGemm<T>> args(n);
    std::vector<int4> tiles[3];
    for (int c = 0; c < 3; ++c) g.gprefix[stg][c].assign(n + 1, 0);
    for (int b = 0; b < n; ++b) {
      const Plan& p = *g.plans[b];
      int M, N, K, nz;
      gemm_stage_shape(p, stg, &M, &N, &K, &nz);
      const int nchunk = K / kTile;
      const int c = nchunk % 4 == 0 ? 0 : (nchunk % 2 == 0 ? 1 : 2);
      args[b].g = stage_args<T>(p, stg);
      args[b].g.nz = nz;
      args[b].g.tm = M / kTile;
      args[b].g.tn = N / kTile;
      args[b].foff = g.foff[b];
      for (int z = 0; z < nz; ++z)
        for (int t = 0; t < (M / kTile) * (N / kTile); ++t) tiles[c].push_back(make_int4(b, z, t, 0));
      for (int cc = 0; cc < 3; ++cc) g.gprefix[stg][cc][b + 1] = (int)tiles[cc].size();
    }
    if ((rc = upload_vec(args, &g.gargs[stg])) != VD_OK) return rc;
    for (int c = 0; c < 3; ++c)
      if ((rc = upload_vec(tiles[c], (void**)&g.gtiles[stg][c])) != VD_OK) return rc;
  }
  return VD_OK;
}

void group_release(Group& g) {
  hipFree(g.mem);
  hipFree(g.stage_tiles);
  for (auto& t : g.adj_tiles) {
    hipFree(t);
    t = nullptr;
  }
  for (int st = 0; st < 4; ++st) {
    for (int sh = 0; sh < 2; ++sh) {
      hipFree(g.sargs[sh][st]);
      hipFree(g.stiles[sh][st]);
      g.sargs[sh][st] = nullptr;
      g.stiles[sh][st] = nullptr;
    }
    hipFree(g.gargs[st]);
    g.gargs[st] = nullptr;
    for (int c = 0; c < 3; ++c) {
      hipFree(g.gtiles[st][c]);
      g.gtiles[st][c] = nullptr;
    }
  }
  hipFree(g.small_mem);
  g.small_mem = nullptr;
  g.small = g.use_small = false;
  g.mem = nullptr;
  g.stage_tiles = nullptr;
}

#define VD_INSTANTIATE(T)                                                                     \
  template Phys<T> make_phys<T>(const Plan&);                                                \
  template void launch_inversion<T>(const Plan&, const T*, const T*, const T*, T*, int,       \
                                    hipStream_t);                                             \
  template void launch_gemm_stage<T>(const Plan&, int, const T*, const T*, const T*, T*, int,  \
                                     int, hipStream_t);                                       \
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
  template int group_build<T>(Group&);                                                       \
  template void launch_stage_rg<T>(const Group&, int, const T*, const T*, const T*, const T*,  \
                                   const T*, T, T, T, T*, T*, hipStream_t);                  \
  template void launch_inversion_rg<T>(const Group&, int, const T*, const T*, const T*, T*,    \
                                       int, hipStream_t);                                    \
  template void launch_stage_adj_rg<T>(const Group&, int, const T*, const T*, const T*,       \
                                       const T*, const T*, T, T, T, const T*, const T*,       \
                                       const T*, T*, T*, hipStream_t);                       \
  template void launch_stage_adj<T>(const Plan&, const T*, const T*, const T*, const T*,      \
                                    const T*, T, T, T, const T*, const T*, const T*, T*, T*,  \
                                    hipStream_t);                                             \
  template void launch_adj_init<T>(const Plan&, const T*, T*, T*, hipStream_t);               \
  template void launch_prep_rg<T>(const Group&, int, const T*, const T*, T*, T*, T*, T*,       \
                                  hipStream_t);                                               \
  template void launch_finish_rg<T>(const Group&, int, const T*, T*, hipStream_t);             \
  template void launch_adj_init_rg<T>(const Group&, int, const T*, T*, T*, hipStream_t);       \
  template void launch_adj_final_rg<T>(const Group&, int, const T*, const T*, const T*, T*,    \
                                       hipStream_t);                                          \
  template void launch_adj_final<T>(const Plan&, const T*, const T*, const T*, T*,            \
                                    hipStream_t);

VD_INSTANTIATE(float)
VD_INSTANTIATE(double)

}  // namespace vd
