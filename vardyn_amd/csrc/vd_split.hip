// Sine-matrix GEMMs of the Helmholtz inversion on int8 MFMA with FP64-class
// accuracy: the split-product (Ozaki) scheme.
//
// Method (Ozaki, Ogita, Oishi & Rump, "Error-free transformations of matrix
// multiplication by using fast routines of matrix multiplication and its
// applications", Numer. Algorithms 59, 2012; integer form: Ootomo, Ozaki &
// Yokota, "DGEMM on integer matrix multiplication unit", IJHPCA 2024).
// Every line (a row of the left operand, a column of the right one) of an
// FP64 operand is scaled by a power of two 2^-e so that its entries lie in
// (-1/4, 1/4), rounded to an integer Y = rint(x 2^(8 NS - e)) and written as
// NS balanced base-256 digits d_1 .. d_NS in [-128, 127]:
//   Y = sum_p d_p 256^(NS-p),  the bytes of (Y + C) ^ C  with  C = 0x80...80.
// A product of two lines is then
//   sum_k x_k y_k ~= 2^(eA + eB) sum_{p+q <= NS+1} 2^(-8 (p+q)) sum_k d_p(x_k) d_q(y_k)
// where every inner sum runs exactly in int32 on the int8 MFMA (|d| <= 128,
// so a level of at most NS pairs stays below 2^31 for K < 2^14); the levels
// are combined in FP64 at the end.  Dropped pairs (p+q > NS+1) and the
// rounding of Y bound the error by about 2^(-8 NS + 5) of the largest
// |x| |y| line products; NS = 6 gives 4e-13 relative on the 510 x 510 DST.
//
// The sine operand is split once per plan (digit planes S8, exponents eS).
// The data operand (a field or an intermediate, with the fold applied) is
// split inside the kernel, once per workgroup, over the whole K range (up to
// 512): fetched and folded into registers, reduced to line exponents, written
// as digit planes to LDS.  Sine digit fragments are read from global memory
// (L2-resident: 6 x 256 x 256 bytes per parity) straight into registers.
//
// MFMA v_mfma_i32_32x32x32_i8 (gfx950): lane l supplies 16 consecutive k
// (16 (l/32) .. +15) of line l%32 on both operands; result register r of
// lane l holds (S line 8 (r/4) + 4 (l/32) + r%4, data line l%32).
#include "vd_gemm.h"

#include <climits>
#include <cmath>

namespace vd {

typedef int i32x4 __attribute__((ext_vector_type(4)));
typedef int i32x16 __attribute__((ext_vector_type(16)));

// Data side of a stage: 0 = right operand (left-multiply stages 0 and 2:
// data lines are columns), 1 = left operand (stages 1 and 3: data lines are rows).
template <int AOP, int BOP, int KIND, int MAP>
constexpr int split_data_side() {
  return (MAP == MAP_COLFOLD || (MAP == MAP_PLAIN && KIND != EK_STORE)) ? 1 : 0;
}

struct SplitArgs {
  GemmArgs<double> g;  // tm = tiles along sine lines, tn = tiles along data lines
  const int8_t* S8;    // [par][NS] digit planes of the sine operand, fragment-major (lines x K each)
  const int* eS;       // [par][lines] line exponents
  size_t pstride;      // bytes between parity blocks of S8
  int lines;           // sine lines per parity in the digit planes (split_line_alloc)
  int nsl;             // sine lines of the product: output past them is not stored
  int ns;
};

// Data element fetch/fold of line `line`, position k.
template <int AOP, int BOP, int DS>
struct DataOp {
  static constexpr int NW = DS == 0 ? b_words<BOP>() : a_words<AOP>();
  static __device__ __forceinline__ void fetch(const GemmArgs<double>& g, const GemmPtrs<double>& P,
                                               int line, int k, double* r) {
    if (DS == 0) fetch_b<double, BOP>(g, P.B, P.Fld, P.Fsub, k, line, r);
    else fetch_a<double, AOP>(g, P.A, line, k, r);
  }
  static __device__ __forceinline__ double combine(const GemmArgs<double>& g, const GemmPtrs<double>& P,
                                                   int line, int k, const double* r) {
    if (DS == 0) return combine_b<double, BOP>(g, k, line, r, P.sgn);
    return combine_a<double, AOP>(g, k, r, P.sgn);
  }
  // Raw words of k0 .. k0+3 (k0 a multiple of 4).  Rows of a left data
  // operand are contiguous in k: 16-byte loads of the four words and, for a
  // column fold, of the four mirrored ones (columns mx-4-k0 .. mx-1-k0; the
  // element-wise path, with its clamping, covers quads that reach past the
  // fold).  Other operands: element-wise.
  static __device__ __forceinline__ void fetch_quad(const GemmArgs<double>& g, const GemmPtrs<double>& P,
                                                    int line, int k0, double (*r)[NW]) {
    if constexpr (DS == 1 && AOP == AOP_DENSE) {
      const double* row = P.A + (size_t)line * g.lda + k0;
      const double2 a = *reinterpret_cast<const double2*>(row);
      const double2 b = *reinterpret_cast<const double2*>(row + 2);
      r[0][0] = a.x; r[1][0] = a.y; r[2][0] = b.x; r[3][0] = b.y;
    } else if constexpr (DS == 1 && AOP == AOP_FOLDCOL) {
      if (k0 + 3 < g.hx) {
        const double* row = P.A + (size_t)line * g.lda;
        const int m0 = g.mx - 4 - k0;
        const double2 a = *reinterpret_cast<const double2*>(row + k0);
        const double2 b = *reinterpret_cast<const double2*>(row + k0 + 2);
        const double2 c = *reinterpret_cast<const double2*>(row + m0);
        const double2 d = *reinterpret_cast<const double2*>(row + m0 + 2);
        r[0][0] = a.x; r[1][0] = a.y; r[2][0] = b.x; r[3][0] = b.y;
        r[0][1] = d.y; r[1][1] = d.x; r[2][1] = c.y; r[3][1] = c.x;
      } else {
#pragma unroll
        for (int u = 0; u < 4; ++u) fetch(g, P, line, k0 + u, r[u]);
      }
    } else {
#pragma unroll
      for (int u = 0; u < 4; ++u) fetch(g, P, line, k0 + u, r[u]);
    }
  }
};

// Z = (Y + C) ^ C of Y = rint(v * sc): the NS digit bytes, most significant
// digit in byte NS-1.  |v * sc| < 2^(8 NS - 2).  For NS <= 6 that is below
// 2^51, and fma(v, sc, 1.5 * 2^52) rounds to the integer while leaving
// 0x4338000000000000 + Y in the bit pattern; for NS = 7 the value is split
// into two 32-bit words in FP64 (every step exact).
template <int NS>
__device__ __forceinline__ unsigned long long split_digits(double v, double sc) {
  constexpr unsigned long long C = 0x8080808080808080ull >> (8 * (8 - NS));
  if constexpr (NS <= 6) {
    const unsigned long long bits = (unsigned long long)__double_as_longlong(fma(v, sc, 0x1.8p52));
    return (bits + (C - 0x4338000000000000ull)) ^ C;
  } else {
    const double y = rint(v * sc);
    const double yh = floor(y * 0x1p-32);
    const double yl = fma(yh, -0x1p32, y);  // in [0, 2^32)
    const unsigned long long Y = ((unsigned long long)(unsigned)(int)yh << 32) | (unsigned)yl;
    return (Y + C) ^ C;
  }
}

// Byte transpose of four elements' digit words: plane word j holds byte j of
// w0..w3 in bytes 0..3.
__device__ __forceinline__ void transpose4(unsigned w0, unsigned w1, unsigned w2, unsigned w3,
                                           unsigned* p0, unsigned* p1, unsigned* p2, unsigned* p3) {
  const unsigned a = __builtin_amdgcn_perm(w1, w0, 0x05010400u);
  const unsigned b = __builtin_amdgcn_perm(w1, w0, 0x07030602u);
  const unsigned c = __builtin_amdgcn_perm(w3, w2, 0x05010400u);
  const unsigned d = __builtin_amdgcn_perm(w3, w2, 0x07030602u);
  *p0 = __builtin_amdgcn_perm(c, a, 0x05040100u);
  *p1 = __builtin_amdgcn_perm(c, a, 0x07060302u);
  *p2 = __builtin_amdgcn_perm(d, b, 0x05040100u);
  *p3 = __builtin_amdgcn_perm(d, b, 0x07060302u);
}

// One workgroup computes a BSL x BDL tile (sine lines x data lines) of one
// product over the whole K range (K <= 32 NSTEP):
//   1. every thread fetches and folds its quads (4 consecutive k of one data
//      line) for all K steps at once -- the loads of all steps are in flight
//      together -- and keeps them in registers;
//   2. per-line maxima reduce through LDS to the line exponents;
//   3. the values are split into NS digit planes in LDS (whole K range);
//   4. the MFMA loop runs over the K steps with no barrier: data digits from
//      LDS, sine digits from global memory one step ahead.
// The four waves tile the output WS x WD ways and split the K steps WK =
// 4 / (WS WD) ways (step st goes to wave group st % WK); K-split partial sums
// are added in int32 through LDS, so every configuration gives the same bits.
// With the data in LDS before the first MFMA, a second workgroup on the CU
// overlaps its fetch/split phases with this one's MFMAs.
template <int AOP, int BOP, int KIND, int MAP, int BSL, int BDL, int WS, int WD, int NS, int NSTEP>
__device__ __forceinline__ void split_tile(const SplitArgs& sa, int z, int tile) {
  constexpr int DS = split_data_side<AOP, BOP, KIND, MAP>();
  typedef DataOp<AOP, BOP, DS> D;
  constexpr int NW = D::NW;
  constexpr int NTHR = 256;
  constexpr int WK = 4 / (WS * WD);
  constexpr int WSL = BSL / WS, WDL = BDL / WD;  // lines per wave
  constexpr int MT = WSL / 32, NT = WDL / 32;
  constexpr int NQ = 8 * BDL / NTHR;             // quads per thread and K step
  constexpr int KP = 32 * NSTEP + 16;            // LDS pitch of a data line (bytes): conflict-free b128 reads
  static_assert(MT >= 1 && NT >= 1 && NQ >= 1 && WS * WD * WK == 4, "tile shape");
  static_assert(WK == 1 || (WK - 1) * WS * WD * MT * NT * 16 * 64 * 4 <= NS * BDL * KP,
                "K-split partials fit LDS");
  __shared__ __align__(16) unsigned char Dl[NS][BDL][KP];
  __shared__ double red[8 * BDL];
  __shared__ double dscale[BDL];
  __shared__ int dexp[BDL];
  __shared__ int sexp[BSL];

  const GemmArgs<double>& g = sa.g;
  const GemmPtrs<double> P = gemm_ptrs<double, AOP, BOP, MAP>(g, z);
  const int s0 = (tile / g.tn) * BSL, d0 = (tile % g.tn) * BDL;
  // the wave index is the same for every lane of a wave: in a scalar register
  const int tid = threadIdx.x, lane = tid & 63, wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int ws = wave % WS, wd = (wave / WS) % WD, wk = wave / (WS * WD);
  const int nstep = g.K / 32;  // 1 .. NSTEP

  // quad j of this thread: data line and first k within a step
  auto qline = [&](int j) {
    const int Q = tid + NTHR * j;
    return DS == 0 ? Q % BDL : Q / 8;
  };
  auto qk = [&](int j) {
    const int Q = tid + NTHR * j;
    return DS == 0 ? 4 * (Q / BDL) : 4 * (Q % 8);
  };

  // sine line exponents of the tile (read in the epilogue)
  for (int i = tid; i < BSL; i += NTHR) sexp[i] = sa.eS[(size_t)P.par * sa.lines + s0 + i];

  // 1. fetch and fold everything (steps past nstep re-read the last one and are zeroed)
  double v[NSTEP][NQ][4];
  {
    double r[NSTEP][NQ][4][NW];
#pragma unroll
    for (int st = 0; st < NSTEP; ++st) {
      const int sc = min(st, nstep - 1);
#pragma unroll
      for (int j = 0; j < NQ; ++j) D::fetch_quad(g, P, d0 + qline(j), sc * 32 + qk(j), r[st][j]);
    }
#pragma unroll
    for (int st = 0; st < NSTEP; ++st)
#pragma unroll
      for (int j = 0; j < NQ; ++j)
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          const double x = D::combine(g, P, d0 + qline(j), st * 32 + qk(j) + u, r[st][j][u]);
          v[st][j][u] = st < nstep ? x : 0.0;
        }
  }

  // 2. line exponents: |x| 2^-e < 1/4 on the line; NaN/Inf lines give NaN outputs
#pragma unroll
  for (int j = 0; j < NQ; ++j) {
    double m = 0.0;
    bool bad = false;
#pragma unroll
    for (int st = 0; st < NSTEP; ++st)
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        bad |= !isfinite(v[st][j][u]);
        m = fmax(m, fabs(v[st][j][u]));
      }
    red[tid + NTHR * j] = bad ? NAN : m;
  }
  __syncthreads();
  if (tid < BDL) {
    double m = 0.0;
    bool nanl = false;
    for (int q = 0; q < 8; ++q) {
      const double x = red[DS == 0 ? tid + BDL * q : 8 * tid + q];
      nanl |= isnan(x);
      m = fmax(m, x);
    }
    int e = 0;
    double sc = 0.0;  // zero line (or below 2^-900): all digits zero
    if (!nanl && m >= 0x1p-900) {
      e = ilogb(m) + 3;
      sc = ldexp(1.0, 8 * NS - e);
    }
    dexp[tid] = nanl ? INT_MIN : e;
    dscale[tid] = sc;
  }
  __syncthreads();

  // 3. digit planes of the whole K range
#pragma unroll
  for (int j = 0; j < NQ; ++j) {
    const int ln = qline(j), k = qk(j);
    const double sc = dscale[ln];
#pragma unroll
    for (int st = 0; st < NSTEP; ++st) {
      if (st >= nstep) break;
      unsigned lo[4], hi[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const unsigned long long Z = split_digits<NS>(v[st][j][u], sc);
        lo[u] = (unsigned)Z;
        hi[u] = (unsigned)(Z >> 32);
      }
      unsigned pl[8];
      transpose4(lo[0], lo[1], lo[2], lo[3], &pl[0], &pl[1], &pl[2], &pl[3]);
      if (NS > 4) transpose4(hi[0], hi[1], hi[2], hi[3], &pl[4], &pl[5], &pl[6], &pl[7]);
      // byte b of Z is digit NS - b (digit 1 most significant)
#pragma unroll
      for (int b = 0; b < NS; ++b)
        *reinterpret_cast<unsigned*>(&Dl[NS - 1 - b][ln][st * 32 + k]) = pl[b];
    }
  }
  __syncthreads();

  // 4. MFMA loop (this wave: steps wk, wk + WK, ...)
  const int8_t* S8 = sa.S8 + (size_t)P.par * sa.pstride;
  // fragment-major planes: the 16 bytes of lane l for 32-line block lb and K
  // step st at ((lb * K/32 + st) * 64 + l) * 16, so one load is 1 KB contiguous
  const size_t plane = (size_t)sa.lines * g.K;
  const int8_t* srow = S8 + (size_t)((s0 + ws * WSL) / 32) * 32 * g.K + 16 * lane;
  typedef i32x4 SineFrag[NS][MT];
  auto load_sine = [&](int st, SineFrag& f) {
#pragma unroll
    for (int p = 0; p < NS; ++p)
#pragma unroll
      for (int a = 0; a < MT; ++a)
        f[p][a] = *reinterpret_cast<const i32x4*>(srow + p * plane + (size_t)a * 32 * g.K + st * 1024);
  };
  i32x16 acc[NS][MT][NT];
#pragma unroll
  for (int l = 0; l < NS; ++l)
#pragma unroll
    for (int a = 0; a < MT; ++a)
#pragma unroll
      for (int b = 0; b < NT; ++b)
#pragma unroll
        for (int r = 0; r < 16; ++r) acc[l][a][b][r] = 0;
  // The loop is unrolled over the K steps of this wave (st = wk + i WK) so that
  // the fragment buffers are registers: sine fragments are loaded two steps
  // ahead (three buffers) and data fragments one step ahead (two buffers), which
  // keeps the L2 and LDS latency of the loads off the MFMA issue.
  typedef i32x4 DataFrag[NS][NT];
  auto load_data = [&](int st, DataFrag& df) {
#pragma unroll
    for (int q = 0; q < NS; ++q)
#pragma unroll
      for (int b = 0; b < NT; ++b)
        df[q][b] = *reinterpret_cast<const i32x4*>(
            &Dl[q][wd * WDL + b * 32 + (lane & 31)][st * 32 + 16 * (lane >> 5)]);
  };
  SineFrag sf[3];
  DataFrag dfr[2];
  auto mfma_step = [&](const SineFrag& cur, const DataFrag& df) {
#pragma unroll
    for (int p = 0; p < NS; ++p)
#pragma unroll
      for (int q = 0; q < NS - p; ++q)
#pragma unroll
        for (int a = 0; a < MT; ++a)
#pragma unroll
          for (int b = 0; b < NT; ++b)
            acc[p + q][a][b] =
                __builtin_amdgcn_mfma_i32_32x32x32_i8(cur[p][a], df[q][b], acc[p + q][a][b], 0, 0, 0);
  };
  if (nstep == NSTEP) {
    // The whole K range of the template: every load below is unconditional
    // (its condition is a compile-time function of the unrolled iteration),
    // so the compiler's wait counts track each prefetch batch exactly and an
    // MFMA waits only for its own step's fragments.  (Under run-time
    // conditions the waits had to assume the worst at every join and drained
    // all outstanding loads once per step.)
    constexpr int NI = NSTEP / WK;
    static_assert(NI * WK == NSTEP, "K steps divide among the wave groups");
    load_sine(wk, sf[0]);
    load_data(wk, dfr[0]);
    if constexpr (NI > 1) load_sine(wk + WK, sf[1]);
#pragma unroll
    for (int i = 0; i < NI; ++i) {
      const int st = wk + i * WK;
      if (i + 2 < NI) load_sine(st + 2 * WK, sf[(i + 2) % 3]);
      if (i + 1 < NI) load_data(st + WK, dfr[(i + 1) % 2]);
      // keep the prefetches ahead of this step's MFMAs: the scheduler would
      // otherwise sink each load next to its first use to shorten live ranges
      __builtin_amdgcn_sched_barrier(0);
      mfma_step(sf[i % 3], dfr[i % 2]);
      __builtin_amdgcn_sched_barrier(0);
    }
  } else {
    if (wk < nstep) {
      load_sine(wk, sf[0]);
      load_data(wk, dfr[0]);
    }
    if (wk + WK < nstep) load_sine(wk + WK, sf[1]);
#pragma unroll
    for (int i = 0; i < NSTEP; ++i) {
      const int st = wk + i * WK;
      if (st >= nstep) break;
      if (st + 2 * WK < nstep) load_sine(st + 2 * WK, sf[(i + 2) % 3]);
      if (st + WK < nstep) load_data(st + WK, dfr[(i + 1) % 2]);
      mfma_step(sf[i % 3], dfr[i % 2]);
    }
  }
  if constexpr (WK > 1) {
    // K-split partials: groups 1..WK-1 park each level in LDS (over the data
    // planes), group 0 adds them in group order
    int* park = reinterpret_cast<int*>(&Dl[0][0][0]);
    constexpr int PER = MT * NT * 16;
    const int slot = ws + WS * wd;  // output block of this wave
#pragma unroll
    for (int l = 0; l < NS; ++l) {
      __syncthreads();
      if (wk > 0) {
#pragma unroll
        for (int a = 0; a < MT; ++a)
#pragma unroll
          for (int b = 0; b < NT; ++b)
#pragma unroll
            for (int r = 0; r < 16; ++r)
              park[(((wk - 1) * (WS * WD) + slot) * PER + (a * NT + b) * 16 + r) * 64 + lane] = acc[l][a][b][r];
      }
      __syncthreads();
      if (wk == 0) {
#pragma unroll
        for (int q = 1; q < WK; ++q)
#pragma unroll
          for (int a = 0; a < MT; ++a)
#pragma unroll
            for (int b = 0; b < NT; ++b)
#pragma unroll
              for (int r = 0; r < 16; ++r)
                acc[l][a][b][r] += park[(((q - 1) * (WS * WD) + slot) * PER + (a * NT + b) * 16 + r) * 64 + lane];
      }
    }
  }

  // 5. combine levels in FP64 (Horner in 2^-8, exact conversions), scale by
  // the line exponents, epilogue.  Left operands' products (DS = 1) have data
  // lines as output rows: their tile goes through LDS so that consecutive
  // threads store (and read Lambda / hb at) consecutive columns of a row.
  auto value = [&](int a, int b, int r) {
    const int dl = wd * WDL + b * 32 + (lane & 31);
    const int sl = ws * WSL + a * 32 + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
    double x = (double)acc[NS - 1][a][b][r];
#pragma unroll
    for (int l = NS - 2; l >= 0; --l) x = fma(x, 0x1p-8, (double)acc[l][a][b][r]);
    const int ed = dexp[dl];
    return ed == INT_MIN ? (double)NAN : ldexp(x, sexp[sl] + ed - 16);
  };
  if constexpr (DS == 0) {
    if (wk == 0) {
#pragma unroll
      for (int a = 0; a < MT; ++a)
#pragma unroll
        for (int b = 0; b < NT; ++b)
#pragma unroll
          for (int r = 0; r < 16; ++r) {
            const int sl = s0 + ws * WSL + a * 32 + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
            if (sl < sa.nsl) store_out<double, KIND, MAP>(g, P, sl, d0 + wd * WDL + b * 32 + (lane & 31), value(a, b, r));
          }
    }
  } else {
    constexpr int TP = BSL + 1;  // row pitch (doubles) of the transposed tile
    constexpr int NE = BDL * BSL / NTHR;  // output elements per thread
    static_assert(BDL * TP * 8 <= NS * BDL * KP, "output tile fits the data planes");
    static_assert(NE * NTHR == BDL * BSL, "the threads cover the output tile");
    // the epilogue's global inputs (hb, the Lambda tables) of this thread's
    // elements, all in flight while the tile goes through LDS
    EpiIn<double> in[NE];
#pragma unroll
    for (int it = 0; it < NE; ++it) {
      const int e = tid + NTHR * it;
      in[it] = epi_input<double, KIND, MAP>(g, P, d0 + e / BSL, s0 + e % BSL);
    }
    double* T = reinterpret_cast<double*>(&Dl[0][0][0]);
    __syncthreads();  // every wave is done with the data planes
    if (wk == 0) {
#pragma unroll
      for (int a = 0; a < MT; ++a)
#pragma unroll
        for (int b = 0; b < NT; ++b)
#pragma unroll
          for (int r = 0; r < 16; ++r)
            T[(wd * WDL + b * 32 + (lane & 31)) * TP + ws * WSL + a * 32 + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3)] =
                value(a, b, r);
    }
    __syncthreads();
#pragma unroll
    for (int it = 0; it < NE; ++it) {
      const int e = tid + NTHR * it, dl = e / BSL, sl = e % BSL;
      if (s0 + sl < sa.nsl) epi_store<double, KIND, MAP>(g, P, d0 + dl, s0 + sl, T[dl * TP + sl], in[it]);
    }
#pragma unroll
    for (int it = 0; it < NE; ++it) {
      const int e = tid + NTHR * it;
      if (s0 + e % BSL < sa.nsl) epi_ring<double, KIND, MAP>(g, P, d0 + e / BSL, s0 + e % BSL);
    }
  }
}

template <int AOP, int BOP, int KIND, int MAP, int BSL, int BDL, int WS, int WD, int NS, int NSTEP>
__global__ __launch_bounds__(256) void k_gemm_split(SplitArgs sa) {
  int z, tile;
  work_item(sa.g, &z, &tile);
  split_tile<AOP, BOP, KIND, MAP, BSL, BDL, WS, WD, NS, NSTEP>(sa, z, tile);
}

// Ragged groups: workgroup b takes tile table entry b = (member, product,
// tile); the member's static operands come from its descriptor, the per-call
// fields from `call` at the member's offset.
struct RgSplit {
  SplitArgs sa;
  size_t foff;
};
template <int AOP, int BOP, int KIND, int MAP, int NS, int BSL>
__global__ __launch_bounds__(256) void k_gemm_split_rg(RgCall<double> call, const RgSplit* __restrict__ rg,
                                                       const int4* __restrict__ tiles) {
  const int4 t = tiles[blockIdx.x];
  SplitArgs sa = rg[t.x].sa;
  const size_t fo = rg[t.x].foff;
  sa.g.F = call.F ? call.F + fo : nullptr;
  sa.g.Fsub = call.Fsub ? call.Fsub + fo : nullptr;
  sa.g.hb = call.hb ? call.hb + fo : nullptr;
  sa.g.out = call.out ? call.out + fo : nullptr;
  // BSL sine lines (64 or 32) x 32 data lines: configuration 2 or 3
  split_tile<AOP, BOP, KIND, MAP, BSL, 32, BSL / 32, 1, NS, 8>(sa, t.y, t.z);
}

// Tile configurations of the split kernel: sine lines x data lines per
// workgroup, waves along the sine and the data lines (the rest split K);
// each wave owns one 32 x 32 output block.
struct SplitCfg {
  int bsl, bdl, ws, wd;
};
constexpr int kNumSplitCfg = 4;
constexpr SplitCfg kSplitCfgs[kNumSplitCfg] = {{128, 32, 4, 1}, {64, 64, 2, 2}, {64, 32, 2, 1}, {32, 32, 1, 1}};
constexpr int kSplitMaxK = 512;  // K range a workgroup keeps in registers and LDS

int split_num_cfgs() { return kNumSplitCfg; }
int split_max_k() { return kSplitMaxK; }

// (sine lines, data lines) of stage `stage`.
static void split_lines(const Plan& p, int stage, int* nsl, int* ndl, int* K, int* nz) {
  int M, N;
  gemm_stage_shape(p, stage, &M, &N, K, nz);
  const bool left = stage == 0 || stage == 2;
  *nsl = left ? M : N;
  *ndl = left ? N : M;
}

bool split_cfg_fits(const Plan& p, int stage, int c) {
  if (c < 0 || c >= kNumSplitCfg) return false;
  int nsl, ndl, K, nz;
  split_lines(p, stage, &nsl, &ndl, &K, &nz);
  // 64 data lines hold their whole K range in registers only up to K = 256
  return nsl % kSplitCfgs[c].bsl == 0 && ndl % kSplitCfgs[c].bdl == 0 && K <= kSplitMaxK &&
         (kSplitCfgs[c].bdl == 32 || K <= 256);
}

int split_default_cfg(const Plan& p, int stage) {
  int nsl, ndl, K, nz;
  split_lines(p, stage, &nsl, &ndl, &K, &nz);
  // the largest tile (fewest re-splits of a data line) with a workgroup per
  // CU; 32 x 32 fits every shape
  for (int c = 0; c < 3; ++c)
    if (split_cfg_fits(p, stage, c) &&
        (long)nz * (nsl / kSplitCfgs[c].bsl) * (ndl / kSplitCfgs[c].bdl) >= 256)
      return c;
  return 3;
}

template <int AOP, int BOP, int KIND, int MAP, int NS, int NSTEP>
static void split_launch_k(SplitArgs a, int cfg, hipStream_t s) {
  const dim3 grid(a.g.nz * a.g.tm * a.g.tn), block(256);
#define VD_S(BSL, BDL, WS, WD) \
  hipLaunchKernelGGL((k_gemm_split<AOP, BOP, KIND, MAP, BSL, BDL, WS, WD, NS, NSTEP>), grid, block, 0, s, a)
  switch (cfg) {
    case 0: VD_S(128, 32, 4, 1); break;
    case 1:
      if constexpr (NSTEP <= 8) VD_S(64, 64, 2, 2);
      break;
    case 2: VD_S(64, 32, 2, 1); break;
    default: VD_S(32, 32, 1, 1); break;
  }
#undef VD_S
}

template <int AOP, int BOP, int KIND, int MAP, int NS>
static void split_launch_ns(SplitArgs a, int cfg, hipStream_t s) {
  if (a.g.K <= 256) split_launch_k<AOP, BOP, KIND, MAP, NS, 8>(a, cfg, s);
  else split_launch_k<AOP, BOP, KIND, MAP, NS, 16>(a, cfg, s);
}

struct SplitLaunch {
  const Plan* p;
  SplitArgs a;
  int cfg;
  hipStream_t s;
  template <int AOP, int BOP, int KIND, int MAP>
  void go() {
    switch (a.ns) {
      case 5: split_launch_ns<AOP, BOP, KIND, MAP, 5>(a, cfg, s); break;
      case 6: split_launch_ns<AOP, BOP, KIND, MAP, 6>(a, cfg, s); break;
      default: split_launch_ns<AOP, BOP, KIND, MAP, 7>(a, cfg, s); break;
    }
  }
};

void launch_split_stage(const Plan& p, int stage, const double* F, const double* Fsub,
                        const double* hb, double* out, int mode, int cfg, hipStream_t s) {
  SplitLaunch l;
  l.p = &p;
  l.a.g = stage_args<double>(p, stage);
  if (stage == 0) { l.a.g.F = F; l.a.g.Fsub = Fsub; }
  if (stage == 3) { l.a.g.out = out; l.a.g.hb = hb; }
  int nsl, ndl, K, nz;
  split_lines(p, stage, &nsl, &ndl, &K, &nz);
  if (!split_cfg_fits(p, stage, cfg)) cfg = split_default_cfg(p, stage);
  const bool left = stage == 0 || stage == 2;
  l.a.S8 = static_cast<const int8_t*>(left ? p.S8y : p.S8x);
  l.a.eS = left ? p.eSy : p.eSx;
  l.a.lines = split_line_alloc(nsl);
  l.a.nsl = nsl;
  l.a.ns = p.split_ns;
  l.a.pstride = (size_t)p.split_ns * l.a.lines * K;
  l.a.g.nz = nz;
  l.a.g.tm = nsl / kSplitCfgs[cfg].bsl;
  l.a.g.tn = ndl / kSplitCfgs[cfg].bdl;
  l.cfg = cfg;
  l.s = s;
  dispatch_stage(p.fold != 0, stage, Fsub != nullptr, mode, l);
}

// Split-product GEMMs of a ragged group: one launch per inversion stage over
// the tiles of every active member, K up to 256, in one of two shapes:
//   0: 64 sine x 32 data lines (configuration 2).  A member's sine lines run
//      to split_line_alloc(lines), a multiple of 64; the tiles' rows past the
//      product's own multiply zero digit lines, and the epilogues discard
//      them.  Each data line is fetched and split once per 64 sine lines, and
//      two wave groups split K (6.7% more MFMA work on the 48-problem set's
//      shapes than shape 1).
//   1: 32 x 32 (configuration 3): no padding, each data line fetched and
//      split once per 32 sine lines, four wave groups split K.
// Every shape gives the same bits; group_build's default is shape 0 and
// vd_group_autotune times both per stage.
static SplitArgs split_stage_args(const Plan& p, int stage, int shape) {
  SplitArgs a;
  a.g = stage_args<double>(p, stage);
  int nsl, ndl, K, nz;
  split_lines(p, stage, &nsl, &ndl, &K, &nz);
  const bool left = stage == 0 || stage == 2;
  a.S8 = static_cast<const int8_t*>(left ? p.S8y : p.S8x);
  a.eS = left ? p.eSy : p.eSx;
  a.lines = split_line_alloc(nsl);
  a.nsl = nsl;
  a.ns = p.split_ns;
  a.pstride = (size_t)p.split_ns * a.lines * K;
  a.g.nz = nz;
  const int cfg = shape == 0 ? 2 : 3;
  a.g.tm = (shape == 0 ? a.lines : nsl) / kSplitCfgs[cfg].bsl;
  a.g.tn = ndl / kSplitCfgs[cfg].bdl;
  return a;
}

int split_group_build(Group& g) {
  for (int shape = 0; shape < 2; ++shape)
    for (int stg = 0; stg < 4; ++stg) {
      std::vector<RgSplit> args(g.n);
      std::vector<int4> tiles;
      g.sprefix[shape][stg].assign(g.n + 1, 0);
      for (int b = 0; b < g.n; ++b) {
        const Plan& p = *g.plans[b];
        int nsl, ndl, K, nz;
        split_lines(p, stg, &nsl, &ndl, &K, &nz);
        VD_REQUIRE(K <= 256, VD_ERR_UNSUPPORTED,
                   "split-product groups need sine products with K up to 256 (interior sizes up to 512 folded)");
        args[b].sa = split_stage_args(p, stg, shape);
        args[b].foff = g.foff[b];
        const int ntile = args[b].sa.g.tm * args[b].sa.g.tn;
        for (int z = 0; z < nz; ++z)
          for (int t = 0; t < ntile; ++t) tiles.push_back(make_int4(b, z, t, 0));
        g.sprefix[shape][stg][b + 1] = (int)tiles.size();
      }
      VD_HIP_TRY(hipMalloc(&g.sargs[shape][stg], args.size() * sizeof(RgSplit)));
      VD_HIP_TRY(hipMemcpy(g.sargs[shape][stg], args.data(), args.size() * sizeof(RgSplit), hipMemcpyHostToDevice));
      VD_HIP_TRY(hipMalloc((void**)&g.stiles[shape][stg], std::max<size_t>(1, tiles.size()) * sizeof(int4)));
      VD_HIP_TRY(hipMemcpy(g.stiles[shape][stg], tiles.data(), tiles.size() * sizeof(int4), hipMemcpyHostToDevice));
    }
  for (int stg = 0; stg < 4; ++stg) g.split_shape[stg] = 0;
  return VD_OK;
}

struct RaggedSplitLaunch {
  RgCall<double> call;
  const RgSplit* rg;
  const int4* tiles;
  int ntiles, ns, shape;
  hipStream_t s;
  template <int AOP, int BOP, int KIND, int MAP>
  void go() {
    if (shape == 0) go_bsl<AOP, BOP, KIND, MAP, 64>();
    else go_bsl<AOP, BOP, KIND, MAP, 32>();
  }
  template <int AOP, int BOP, int KIND, int MAP, int BSL>
  void go_bsl() {
    const dim3 grid(ntiles), block(256);
    if (ns == 5) hipLaunchKernelGGL((k_gemm_split_rg<AOP, BOP, KIND, MAP, 5, BSL>), grid, block, 0, s, call, rg, tiles);
    else if (ns == 6) hipLaunchKernelGGL((k_gemm_split_rg<AOP, BOP, KIND, MAP, 6, BSL>), grid, block, 0, s, call, rg, tiles);
    else hipLaunchKernelGGL((k_gemm_split_rg<AOP, BOP, KIND, MAP, 7, BSL>), grid, block, 0, s, call, rg, tiles);
  }
};

void launch_split_stage_rg(const Group& g, int st, int shape, int nact, const double* F, const double* Fsub,
                           const double* hb, double* out, int mode, hipStream_t s) {
  RaggedSplitLaunch l;
  l.call = RgCall<double>{st == 0 ? F : nullptr, st == 0 ? Fsub : nullptr, st == 3 ? hb : nullptr,
                          st == 3 ? out : nullptr};
  l.rg = static_cast<const RgSplit*>(g.sargs[shape][st]);
  l.tiles = g.stiles[shape][st];
  l.ntiles = g.sprefix[shape][st][nact];
  l.ns = g.split_ns;
  l.shape = shape;
  l.s = s;
  if (l.ntiles > 0) dispatch_stage(g.fold != 0, st, Fsub != nullptr, mode, l);
}

void launch_inversion_split_rg(const Group& g, int nact, const double* F, const double* Fsub,
                               const double* hb, double* out, int mode, hipStream_t s) {
  for (int st = 0; st < 4; ++st) launch_split_stage_rg(g, st, g.split_shape[st], nact, F, Fsub, hb, out, mode, s);
}

}  // namespace vd
