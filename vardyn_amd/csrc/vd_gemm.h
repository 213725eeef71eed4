// GEMM pieces shared by the inversion kernels of libvardyn_hip: operand
// descriptors, operand fetch/combine, work mapping, epilogues and the
// per-stage operand setup.
#pragma once

#include "vd_internal.h"

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
// through LDS with a two-chunk register prefetch).  The KS partial tiles are
// summed through LDS in a fixed order, so results are deterministic.  More
// groups give the CU more waves to hide L2 latency with; the product stays
// one launch.
//
// Operands: A dense (AOP_DENSE) or column-folded (AOP_FOLDCOL); B dense,
// the interior of a field, optionally minus a second field (BOP_FIELD,
// BOP_FIELD_SUB), or row-folded (BOP_FOLDROW_*).  All dense operands are zero-padded to
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
enum {
  BOP_DENSE = 0,
  BOP_FIELD = 1,
  BOP_FOLDROW_FIELD = 2,
  BOP_FOLDROW_DENSE = 3,
  BOP_FIELD_SUB = 4,
  BOP_FOLDROW_FIELD_SUB = 5
};
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
  // EK_SCALE factor 1/Lambda, Lambda = gof (lcx[col] + lcy[row]) - gfc2[member]
  // (separable: two 1-D tables and one constant per member), zero in the padding
  const double* lcx;
  const double* lcy;
  const double* gfc2;
  double gof;
  const T* hb;     // EK_FIELD_SET offset / ring field (nullptr = 0)
  T* out;          // EK_FIELD_*: (ny, nx) destination
  // batch strides (elements) of each operand; 0 for operands shared by all members
  size_t bsA, bsB, bsF, bsC, bsOut;
  // launch shape: nz products (parities x members), tm x tn tiles each, nb members
  int nz, tm, tn, nb;
};

// Ragged groups: per-call field pointers of a GEMM launch (member b's fields
// start at its offset in every flat buffer), and the static operands of one
// member for one inversion stage.
template <typename T>
struct RgCall {
  const T* F;
  const T* Fsub;
  const T* hb;
  T* out;
};
template <typename T>
struct RgGemm {
  GemmArgs<T> g;
  size_t foff;
};

// Operand elements.  A(i, l): dense, or column-folded (sum/difference of
// columns l and mx-1-l).  B(k, n): dense, the interior of field F (minus a
// second field Fsub for the *_SUB kinds), or row-folded.  Indices past the
// folded/interior extent read as zero.
//
// Each access is split into fetch (the raw words, from clamped -- always
// valid -- addresses) and combine (fold, subtraction and range select).  The
// kernel fetches chunks ahead and combines only when it writes them to LDS,
// so no arithmetic on a prefetched word forces an early wait for it and every
// load of a chunk is in flight while the previous chunk is multiplied.
template <int AOP>
constexpr int a_words() { return AOP == AOP_DENSE ? 1 : 2; }
template <int BOP>
constexpr int b_words() {
  return (BOP == BOP_DENSE || BOP == BOP_FIELD) ? 1 : (BOP == BOP_FOLDROW_FIELD_SUB ? 4 : 2);
}
template <typename T, int AOP>
__device__ __forceinline__ void fetch_a(const GemmArgs<T>& g, const T* A, int i, int l, T* r) {
  if (AOP == AOP_DENSE) {
    r[0] = A[(size_t)i * g.lda + l];
  } else {
    const int lc = min(l, g.hx - 1);
    r[0] = A[(size_t)i * g.lda + lc];
    r[1] = A[(size_t)i * g.lda + (g.mx - 1 - lc)];
  }
}
template <typename T, int AOP>
__device__ __forceinline__ T combine_a(const GemmArgs<T>& g, int l, const T* r, T sgn) {
  if (AOP == AOP_DENSE) return r[0];
  return l < g.hx ? r[0] + sgn * r[1] : T(0);
}
template <typename T, int BOP>
__device__ __forceinline__ void fetch_b(const GemmArgs<T>& g, const T* B, const T* Fld,
                                        const T* Fsub, int k, int n, T* r) {
  if (BOP == BOP_DENSE) {
    r[0] = B[(size_t)k * g.ldb + n];
  } else if (BOP == BOP_FOLDROW_DENSE) {
    const int kc = min(k, g.hy - 1);
    r[0] = B[(size_t)kc * g.ldb + n];
    r[1] = B[(size_t)(g.my - 1 - kc) * g.ldb + n];
  } else if (BOP == BOP_FIELD || BOP == BOP_FIELD_SUB) {
    const size_t idx = (size_t)(min(k, g.my - 1) + 1) * g.nx + (min(n, g.mx - 1) + 1);
    r[0] = Fld[idx];
    if (BOP == BOP_FIELD_SUB) r[1] = Fsub[idx];
  } else {
    const int kc = min(k, g.hy - 1), nc = min(n, g.mx - 1);
    const size_t i1 = (size_t)(kc + 1) * g.nx + (nc + 1);
    const size_t i2 = (size_t)(g.my - kc) * g.nx + (nc + 1);
    r[0] = Fld[i1];
    r[1] = Fld[i2];
    if (BOP == BOP_FOLDROW_FIELD_SUB) {
      r[2] = Fsub[i1];
      r[3] = Fsub[i2];
    }
  }
}
template <typename T, int BOP>
__device__ __forceinline__ T combine_b(const GemmArgs<T>& g, int k, int n, const T* r, T sgn) {
  if (BOP == BOP_DENSE) return r[0];
  if (BOP == BOP_FOLDROW_DENSE) return k < g.hy ? r[0] + sgn * r[1] : T(0);
  if (BOP == BOP_FIELD) return (k < g.my && n < g.mx) ? r[0] : T(0);
  if (BOP == BOP_FIELD_SUB) return (k < g.my && n < g.mx) ? r[0] - r[1] : T(0);
  if (BOP == BOP_FOLDROW_FIELD) return (k < g.hy && n < g.mx) ? r[0] + sgn * r[1] : T(0);
  return (k < g.hy && n < g.mx) ? (r[0] - r[2]) + sgn * (r[1] - r[3]) : T(0);
}

// Work item of a workgroup.  Workgroups are dispatched round-robin over the 8
// XCDs (linear id % 8), each with its own L2; the mapping gives every XCD
// whole products (or a contiguous run of row-panels of one product) so that
// its L2 holds one member's operands.  z = parity * nb + member for folded
// products, member otherwise.
template <typename T>
__device__ __forceinline__ void work_item(const GemmArgs<T>& g, int* z, int* tile) {
  const int L = blockIdx.x, T_ = g.tm * g.tn;
  const int x = L & 7, slot = L >> 3;
  if (g.nz % 8 == 0) {
    *z = x + 8 * (slot / T_);
    *tile = slot % T_;
  } else if (8 % g.nz == 0 && T_ % (8 / g.nz) == 0) {
    const int per = 8 / g.nz;
    *z = x % g.nz;
    *tile = (x / g.nz) * (T_ / per) + slot;
  } else {
    *z = L / T_;
    *tile = L % T_;
  }
}

// Operand and output pointers of one product.
template <typename T>
struct GemmPtrs {
  const T *A, *B, *Fld, *Fsub, *hbm;
  T *Cout, *fout;
  int par;
  T sgn;
  double gfc2;  // g f / c^2 of the member
};
template <typename T, int AOP, int BOP, int MAP>
__device__ __forceinline__ GemmPtrs<T> gemm_ptrs(const GemmArgs<T>& g, int z) {
  constexpr bool FOLD = MAP != MAP_PLAIN;
  GemmPtrs<T> P;
  P.par = FOLD ? z / g.nb : 0;
  const size_t mb = FOLD ? z % g.nb : z;
  P.sgn = P.par ? T(-1) : T(1);
  P.A = g.A + mb * g.bsA + (AOP == AOP_DENSE ? P.par * g.strideA : 0);
  P.B = g.B + mb * g.bsB + (BOP == BOP_DENSE ? P.par * g.strideB : 0);
  P.Fld = g.F + mb * g.bsF;
  P.Fsub = g.Fsub ? g.Fsub + mb * g.bsF : nullptr;
  P.Cout = g.C + mb * g.bsC;
  P.gfc2 = g.gfc2 ? g.gfc2[mb] : 0.0;
  P.fout = g.out + mb * g.bsOut;
  P.hbm = g.hb ? g.hb + mb * g.bsOut : nullptr;
  return P;
}

// Epilogue for output element (row, col) of a product, in three parts so that
// a kernel can issue the loads of all its elements before the first store and
// keep the few boundary elements' extra stores out of the store sequence (an
// output pointer may alias an input, so the compiler keeps every load behind
// the stores before it, and a load between stores waits for all of them):
//   epi_input  the loads: hb for EK_FIELD_SET, the field for EK_FIELD_ACC,
//              the two Lambda tables for EK_SCALE;
//   epi_store  the element's own store, given those values;
//   epi_ring   EK_FIELD_SET's ring stores (h = hb on the boundary rows and
//              columns) of an element next to the boundary.
template <typename T>
struct EpiIn {
  T x;             // hb or the field
  double lcx, lcy;  // Lambda tables at the element's column and row
};

template <typename T, int KIND, int MAP>
__device__ __forceinline__ EpiIn<T> epi_input(const GemmArgs<T>& g, const GemmPtrs<T>& P, int row, int col) {
  // Straight-line code: every load is unconditional, from an index clamped
  // into its array (an element outside the interior reads element 0);
  // epi_store discards what it does not use.  So all of a thread's loads are
  // in flight together.  EK_FIELD_SET without hb loads nothing (the branch is
  // on a kernel argument, the same for every thread).
  EpiIn<T> in{T(0), 0.0, 0.0};
  if (MAP == MAP_ROWFOLD || KIND == EK_STORE) return in;
  if (KIND == EK_FIELD_SET && !P.hbm) return in;
  const int oc = MAP == MAP_COLFOLD ? 2 * col + P.par : col;
  const bool ok = row < g.my && (MAP == MAP_COLFOLD ? col < g.hx : col < g.mx) && oc < g.mx;
  if (KIND == EK_SCALE) {
    in.lcx = g.lcx[ok ? oc : 0];
    in.lcy = g.lcy[ok ? row : 0];
    return in;
  }
  const T* src = KIND == EK_FIELD_SET ? P.hbm : P.fout;
  const size_t i = ok ? row + 1 : 1, j = ok ? oc + 1 : 1;
  in.x = src[i * g.nx + j];
  return in;
}

template <typename T, int KIND, int MAP>
__device__ __forceinline__ void epi_store(const GemmArgs<T>& g, const GemmPtrs<T>& P, int row, int col, T v,
                                          const EpiIn<T>& in) {
  if (MAP == MAP_ROWFOLD) {
    if (row < g.hy) P.Cout[(size_t)(2 * row + P.par) * g.ldc + col] = v;
    return;
  }
  const int oc = MAP == MAP_COLFOLD ? 2 * col + P.par : col;
  if (KIND == EK_STORE || KIND == EK_SCALE) {
    if (MAP == MAP_COLFOLD && col >= g.hx) return;
    const size_t o = (size_t)row * g.ldc + oc;
    if (KIND == EK_SCALE) {
      // the same operations, in the same order and unfused, as the host's
      // Lambda (so the factor is the correctly rounded 1/Lambda of before)
      double lam;
      {
#pragma clang fp contract(off)
        lam = g.gof * (in.lcx + in.lcy) - P.gfc2;
      }
      P.Cout[o] = (row < g.my && oc < g.mx) ? v * T(1.0 / lam) : T(0);
    } else {
      P.Cout[o] = v;
    }
    return;
  }
  const bool ok = row < g.my && (MAP == MAP_COLFOLD ? col < g.hx : col < g.mx);
  if (!ok) return;
  const size_t i = row + 1, j = oc + 1;
  P.fout[i * g.nx + j] = (KIND == EK_FIELD_ACC || P.hbm ? in.x : T(0)) + v;
}

template <typename T, int KIND, int MAP>
__device__ __forceinline__ void epi_ring(const GemmArgs<T>& g, const GemmPtrs<T>& P, int row, int col) {
  if (MAP == MAP_ROWFOLD || KIND != EK_FIELD_SET) return;
  const bool ok = row < g.my && (MAP == MAP_COLFOLD ? col < g.hx : col < g.mx);
  const int oc = MAP == MAP_COLFOLD ? 2 * col + P.par : col;
  const bool top = row == 0, bot = row == g.my - 1, lft = oc == 0, rgt = oc == g.mx - 1;
  if (!ok || !(top || bot || lft || rgt)) return;
  const int nx = g.nx, ny = g.ny;
  const size_t i = row + 1, j = oc + 1;
  T* fout = P.fout;
  const T* hb = P.hbm;
  // ring: h = hb on rows 0, ny-1 and columns 0, nx-1 (jqgm.py:316-317)
  auto ring = [&](size_t ii, size_t jj) { fout[ii * nx + jj] = hb ? hb[ii * nx + jj] : T(0); };
  if (top) ring(0, j);
  if (bot) ring(ny - 1, j);
  if (lft) ring(i, 0);
  if (rgt) ring(i, nx - 1);
  if (top && lft) ring(0, 0);
  if (top && rgt) ring(0, nx - 1);
  if (bot && lft) ring(ny - 1, 0);
  if (bot && rgt) ring(ny - 1, nx - 1);
}

template <typename T, int KIND, int MAP>
__device__ __forceinline__ void store_out(const GemmArgs<T>& g, const GemmPtrs<T>& P, int row, int col, T v) {
  epi_store<T, KIND, MAP>(g, P, row, col, v, epi_input<T, KIND, MAP>(g, P, row, col));
  epi_ring<T, KIND, MAP>(g, P, row, col);
}


// Static operands of inversion stage `stage` of plan p (the per-call field
// pointers F, Fsub, hb, out are left null):
//   0: T1 = S_y (F - Fsub)   1: T2 = (T1 S_x) ./ Lambda
//   2: T1 = S_y T2           3: out = epilogue(T1 S_x)
template <typename T>
GemmArgs<T> stage_args(const Plan& p, int stage) {
  T* T1 = static_cast<T*>(p.gT[0]);
  T* T2 = static_cast<T*>(p.gT[1]);
  GemmArgs<T> a{};
  a.ny = p.ny; a.nx = p.nx; a.my = p.my; a.mx = p.mx; a.hy = p.hy; a.hx = p.hx;
  a.lcx = p.lam_cx;
  a.lcy = p.lam_cy;
  a.gfc2 = p.gfc2_d;
  a.gof = p.g / p.f0;
  const size_t NF = p.nf, NG = p.ng;
  a.nb = p.nb;
  a.bsF = NF; a.bsOut = NF;
  const bool fold = p.fold;
  const T* Ay = static_cast<const T*>(fold ? p.SyF : p.Sy);
  const T* Bx = static_cast<const T*>(fold ? p.SxF : p.Sx);
  const int ldy = fold ? p.HY : p.MY, ldx = fold ? p.HX : p.MX;
  if (fold) {
    a.strideA = (size_t)p.HY * p.HY;
    a.strideB = (size_t)p.HX * p.HX;
  }
  if (stage == 0) {
    a.A = Ay; a.lda = ldy; a.K = ldy; a.C = T1; a.ldc = p.MX;
    a.bsA = 0; a.bsB = 0; a.bsC = NG;
  } else if (stage == 1) {
    a.A = T1; a.lda = p.MX; a.B = Bx; a.ldb = ldx; a.K = ldx; a.C = T2; a.ldc = p.MX;
    a.bsA = NG; a.bsB = 0; a.bsC = NG;
  } else if (stage == 2) {
    a.A = Ay; a.lda = ldy; a.B = T2; a.ldb = p.MX; a.K = ldy; a.C = T1; a.ldc = p.MX;
    a.bsA = 0; a.bsB = NG; a.bsC = NG;
  } else {
    a.A = T1; a.lda = p.MX; a.B = Bx; a.ldb = ldx; a.K = ldx;
    a.bsA = NG; a.bsB = 0; a.bsC = 0;
  }
  return a;
}

// Calls L::go<AOP, BOP, KIND, MAP>() with the operand and epilogue kinds of
// inversion stage `stage` (fold or dense, subtracted field or not,
// overwrite or accumulate).
template <typename L>
void dispatch_stage(bool fold, int stage, bool sub, int mode, L& l) {
  const bool acc = mode == INV_ACC;
  if (fold) {
    if (stage == 0) {
      if (sub) l.template go<AOP_DENSE, BOP_FOLDROW_FIELD_SUB, EK_STORE, MAP_ROWFOLD>();
      else l.template go<AOP_DENSE, BOP_FOLDROW_FIELD, EK_STORE, MAP_ROWFOLD>();
    } else if (stage == 1) {
      l.template go<AOP_FOLDCOL, BOP_DENSE, EK_SCALE, MAP_COLFOLD>();
    } else if (stage == 2) {
      l.template go<AOP_DENSE, BOP_FOLDROW_DENSE, EK_STORE, MAP_ROWFOLD>();
    } else if (acc) {
      l.template go<AOP_FOLDCOL, BOP_DENSE, EK_FIELD_ACC, MAP_COLFOLD>();
    } else {
      l.template go<AOP_FOLDCOL, BOP_DENSE, EK_FIELD_SET, MAP_COLFOLD>();
    }
    return;
  }
  if (stage == 0) {
    if (sub) l.template go<AOP_DENSE, BOP_FIELD_SUB, EK_STORE, MAP_PLAIN>();
    else l.template go<AOP_DENSE, BOP_FIELD, EK_STORE, MAP_PLAIN>();
  } else if (stage == 1) {
    l.template go<AOP_DENSE, BOP_DENSE, EK_SCALE, MAP_PLAIN>();
  } else if (stage == 2) {
    l.template go<AOP_DENSE, BOP_DENSE, EK_STORE, MAP_PLAIN>();
  } else if (acc) {
    l.template go<AOP_DENSE, BOP_DENSE, EK_FIELD_ACC, MAP_PLAIN>();
  } else {
    l.template go<AOP_DENSE, BOP_DENSE, EK_FIELD_SET, MAP_PLAIN>();
  }
}

}  // namespace vd
