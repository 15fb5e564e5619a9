// K2: batched complex GEMM on CDNA4 matrix cores, plus small elementwise kernels.
//
// C[z] = alpha * op(A[z]) op(B[z]) + beta * C0[z] + shift[z] * I
//
// Design (gfx950, wave64):
//  * workgroup = 256 threads = 4 waves in a 2 x 2 arrangement; workgroup tile
//    64 x 64 complex, wave tile 32 x 32 = 2 x 2 MFMA tiles of 16 x 16;
//  * K advances in chunks of 16 complex staged through two LDS buffers: the
//    next chunk's global loads are issued before the current chunk's MFMAs and
//    land in the other buffer afterwards (one barrier per chunk);
//  * a complex product is 3 real MFMAs on v_mfma_f64_16x16x4_f64 (3M:
//    re*re, im*im, (re+im)*(re+im)) in FP64, and 4 on v_mfma_f32_16x16x4_f32
//    in complex64 mode (operands rounded to float when staged into LDS, FP32
//    accumulation);
//  * op(A) and op(B) tiles are written to LDS k-major, so each MFMA operand
//    read is 16 consecutive complex per quarter-wave;
//  * the linear workgroup id is remapped so that consecutive tiles (same GEMM,
//    neighbouring batch entries) land on the same XCD and share its L2;
//  * smaller configurations (32 x 32 down to 16 x 16, some splitting k among
//    their waves) serve products too small or too few to fill 256 CUs with
//    64 x 64 tiles.
//
// MFMA 16x16x4 operand layout (lane l of 64): A[i][k] with i = l % 16,
// k = l / 16; B[k][j] with j = l % 16, k = l / 16.  Accumulator item v holds
// D[l/16 + 4*v][l % 16] for the FP64 instruction and D[4*(l/16) + v][l % 16]
// for the FP32 one (measured on gfx950 with lane-coded operands).
#include <cstdarg>
#include <cstdio>

#include "sinv_core.hpp"

namespace sinv {

namespace {

typedef double d4v __attribute__((ext_vector_type(4)));
typedef float f4v __attribute__((ext_vector_type(4)));

template <typename R>
struct CR {
  R re, im;
};

template <typename R>
struct Acc;
template <>
struct Acc<double> {
  using V = d4v;
};
template <>
struct Acc<float> {
  using V = f4v;
};

__device__ __forceinline__ d4v mfma(double a, double b, d4v c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ f4v mfma(float a, float b, f4v c) {
  return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

// row of accumulator item v in lane l, within a 16 x 16 tile
template <typename R>
__device__ __forceinline__ int acc_row(int lane, int v);
template <>
__device__ __forceinline__ int acc_row<double>(int lane, int v) {
  return (lane >> 4) + 4 * v;
}
template <>
__device__ __forceinline__ int acc_row<float>(int lane, int v) {
  return (lane >> 4) * 4 + v;
}

// complex product form: 3 real MFMAs (3M) for FP64, 4 for FP32 (3M's extra
// cancellation is not worth it at single precision)
template <typename R>
constexpr bool kThreeM = sizeof(R) == 8;

constexpr int BK = 16;
constexpr int PAD = 1;

constexpr int kGroupM = 8;  // tile rows per band of the grouped tile order

struct TileMap {
  int z_in, z_out, m0, n0;
};

// XCD-aware remap: hardware places workgroup `bid` on XCD bid % 8.  Give each
// XCD a contiguous range of logical tiles.
template <int BMt, int BNt>
__device__ __forceinline__ TileMap map_tile(int tiles_m, int tiles_n, int inner) {
  const int total = gridDim.x;
  int bid = blockIdx.x;
  if ((total & 7) == 0) {
    const int per = total >> 3;
    bid = (bid & 7) * per + (bid >> 3);
  }
  const int tiles = tiles_m * tiles_n;
  const int z = bid / tiles;
  const int t = bid - z * tiles;
  TileMap m;
  m.z_out = z / inner;
  m.z_in = z - m.z_out * inner;
  if (tiles_m >= kGroupM && tiles_n >= kGroupM) {
    // large tile grids (unbatched products): walk the tiles in bands of
    // kGroupM tile rows, column by column inside a band, so an XCD's
    // contiguous range covers a compact block of C and reads few panels of
    // op(A) and op(B) (32 tiles of 32 x 32 over K = 512: 12 panels, 3 MB,
    // held in the XCD's 4 MB L2, against 18 panels, 4.5 MB, row by row)
    const int band = t / (kGroupM * tiles_n);
    const int first = band * kGroupM;
    const int rows = min(tiles_m - first, kGroupM);
    const int r = t - band * (kGroupM * tiles_n);
    m.m0 = (first + r % rows) * BMt;
    m.n0 = (r / rows) * BNt;
  } else {
    m.m0 = (t / tiles_n) * BMt;
    m.n0 = (t % tiles_n) * BNt;
  }
  return m;
}

__device__ __forceinline__ cplx ld_c(const cplx* p) {
  const double2 v = *reinterpret_cast<const double2*>(p);
  return cplx{v.x, v.y};
}

// Tile configuration: WM x WN waves, each owning TM x TN MFMA tiles of 16 x 16;
// workgroup tile (16 WM TM) x (16 WN TN).  OCC is the occupancy the register
// budget is sized for (waves per SIMD).  WK > 1 splits K inside the
// workgroup: WK groups of WM x WN waves share the output tile, each takes its
// own BK-slice of every KC = WK BK chunk, and the groups' accumulators are
// summed through LDS, in group order, before the epilogue.  That multiplies
// the workgroups a small or unbatched product can occupy by the tile-area
// reduction while each wave still issues a full chunk's MFMAs per barrier.
template <int WM_, int WN_, int TM_, int TN_, int OCC_, int WK_ = 1>
struct Cfg {
  static constexpr int WM = WM_, WN = WN_, TM = TM_, TN = TN_, OCC = OCC_, WK = WK_;
  static constexpr int BM = 16 * WM * TM, BN = 16 * WN * TN;
  static constexpr int NT = 64 * WM * WN * WK;
  static constexpr int KC = BK * WK;  // k per LDS chunk
  // load groups: VA (VB) elements consecutive in memory, GA (GB) groups per thread
  static constexpr int VA = BM * KC / NT >= 4 ? 4 : BM * KC / NT;
  static constexpr int VB = BN * KC / NT >= 4 ? 4 : BN * KC / NT;
  static constexpr int GA = BM * KC / VA / NT;
  static constexpr int GB = BN * KC / VB / NT;
  static_assert(VA >= 1 && VB >= 1 && GA * VA * NT == BM * KC && GB * VB * NT == BN * KC,
                "tile must split evenly into load groups");
};

template <typename R, typename C>
__global__ __launch_bounds__(C::NT) __attribute__((amdgpu_waves_per_eu(C::OCC)))
void gemm_kernel(GemmArgs g, int tiles_m, int tiles_n) {
  constexpr int BMt = C::BM, BNt = C::BN, NT = C::NT, TM = C::TM, TN = C::TN;
  constexpr int GA = C::GA, GB = C::GB, VA = C::VA, VB = C::VB, KC = C::KC, WK = C::WK;
  // one allocation: the two chunk buffers of A and of B, reused after the K
  // loop for the split-K partial sums (WK > 1)
  constexpr int kAElems = 2 * KC * (BMt + PAD), kBElems = 2 * KC * (BNt + PAD);
  constexpr int kRedVals = (WK - 1) * C::WM * C::WN * 64 * 3 * TM * TN * 4;
  static_assert(kRedVals * sizeof(R) <= (kAElems + kBElems) * sizeof(CR<R>),
                "split-K partial sums must fit in the chunk buffers");
  __shared__ CR<R> lds_raw[kAElems + kBElems];
  auto& As2 = *reinterpret_cast<CR<R>(*)[2][KC][BMt + PAD]>(lds_raw);
  auto& Bs2 = *reinterpret_cast<CR<R>(*)[2][KC][BNt + PAD]>(lds_raw + kAElems);

  const TileMap tm = map_tile<BMt, BNt>(tiles_m, tiles_n, g.inner);
  // converged instances of an iteration leave at once (uniform per workgroup)
  if (g.skip != nullptr && g.skip[tm.z_out * g.inner + tm.z_in] < g.skip_iter) return;
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int wk = wave / (C::WM * C::WN);         // k-slice group (0 when WK == 1)
  const int wmn = wave - wk * (C::WM * C::WN);
  const int wm = wmn / C::WN, wn = wmn % C::WN;

  const cplx* A = g.A + tm.z_in * g.sA_in + tm.z_out * g.sA_out;
  const cplx* B = g.B + tm.z_in * g.sB_in + tm.z_out * g.sB_out;

  // op(A) tile is BM x BK, op(B) tile is BK x BN, staged k-major in LDS.  Each
  // thread moves EA (EB) elements, 16 B each.  An operand stored contiguous
  // along k (A as M x K, B as N x K) is read in runs of VA (VB) consecutive k
  // per thread: a run lands in VA different LDS rows, and the padded row
  // stride spreads the 16 lanes of a quarter-wave over all banks.  An operand
  // contiguous along m or n is read lane-consecutively along that dimension:
  // each wave reads whole rows (coalesced) and writes them to one LDS row
  // (bank-conflict free); runs along n would put 16 lanes on 4 bank groups.
  constexpr int EA = GA * VA, EB = GB * VB;
  const bool a_kcontig = (g.opA == OP_N);
  const bool b_ncontig = (g.opB == OP_N);
  const double a_conj = (g.opA == OP_C) ? -1.0 : 1.0;
  const double b_conj = (g.opB == OP_C) ? -1.0 : 1.0;

  int am[EA], ak[EA];  // tile coordinates (m, k) of each element this thread moves
#pragma unroll
  for (int j = 0; j < EA; ++j) {
    if (a_kcontig) {
      const int L = tid + (j / VA) * NT;
      am[j] = L / (KC / VA);
      ak[j] = (L % (KC / VA)) * VA + j % VA;
    } else {
      const int E = tid + j * NT;
      ak[j] = E / BMt;
      am[j] = E % BMt;
    }
  }
  int bk[EB], bn[EB];  // (k, n)
#pragma unroll
  for (int j = 0; j < EB; ++j) {
    if (b_ncontig) {
      const int E = tid + j * NT;
      bk[j] = E / BNt;
      bn[j] = E % BNt;
    } else {
      const int L = tid + (j / VB) * NT;
      bn[j] = L / (KC / VB);
      bk[j] = (L % (KC / VB)) * VB + j % VB;
    }
  }

  struct Regs {
    cplx ra[EA], rb[EB];
    bool oka[EA], okb[EB];
    bool fullA, fullB;  // whole chunk in range: no masks (wave-uniform)
  };
  Regs R0;
  // per-element pointers at k = 0, advanced by k0 (or k0 * ld) per chunk
  const cplx* pa[EA];
  const cplx* pb[EB];
  const bool fullM = tm.m0 + BMt <= g.M, fullN = tm.n0 + BNt <= g.N;
#pragma unroll
  for (int j = 0; j < EA; ++j)
    pa[j] = a_kcontig ? A + (long long)(tm.m0 + am[j]) * g.lda + ak[j]
                      : A + (long long)ak[j] * g.lda + (tm.m0 + am[j]);
#pragma unroll
  for (int j = 0; j < EB; ++j)
    pb[j] = b_ncontig ? B + (long long)bk[j] * g.ldb + (tm.n0 + bn[j])
                      : B + (long long)(tm.n0 + bn[j]) * g.ldb + bk[j];
  const long long a_kstep = a_kcontig ? 1 : g.lda;  // elements per unit of k
  const long long b_kstep = b_ncontig ? g.ldb : 1;

  // Branch-free tile loads: an out-of-range element is read from (0, 0), which
  // always exists, and zeroed when staged into LDS.  All loads of a chunk are
  // then in flight together instead of each behind its own branch.
  auto load_tiles = [&](Regs& X, int k0) {
    X.fullA = fullM && k0 + KC <= g.K;
    X.fullB = fullN && k0 + KC <= g.K;
    if (X.fullA) {
#pragma unroll
      for (int j = 0; j < EA; ++j) X.ra[j] = ld_c(pa[j] + k0 * a_kstep);
    } else {
#pragma unroll
    for (int j = 0; j < EA; ++j) {
      const int m = tm.m0 + am[j], k = k0 + ak[j];
      X.oka[j] = (m < g.M) && (k < g.K);
      const long long mm = X.oka[j] ? m : 0, kk = X.oka[j] ? k : 0;
      X.ra[j] = a_kcontig ? ld_c(A + mm * g.lda + kk) : ld_c(A + kk * g.lda + mm);
    }
    }
    if (X.fullB) {
#pragma unroll
      for (int j = 0; j < EB; ++j) X.rb[j] = ld_c(pb[j] + k0 * b_kstep);
    } else {
#pragma unroll
    for (int j = 0; j < EB; ++j) {
      const int k = k0 + bk[j], n = tm.n0 + bn[j];
      X.okb[j] = (k < g.K) && (n < g.N);
      const long long kk = X.okb[j] ? k : 0, nn = X.okb[j] ? n : 0;
      X.rb[j] = b_ncontig ? ld_c(B + kk * g.ldb + nn) : ld_c(B + nn * g.ldb + kk);
    }
    }
  };
  auto store_tiles = [&](const Regs& X, int buf) {
    auto& As = As2[buf];
    auto& Bs = Bs2[buf];
    if (X.fullA && g.opA != OP_C) {
#pragma unroll
      for (int j = 0; j < EA; ++j) As[ak[j]][am[j]] = CR<R>{(R)X.ra[j].re, (R)X.ra[j].im};
    } else if (X.fullA) {
#pragma unroll
      for (int j = 0; j < EA; ++j) As[ak[j]][am[j]] = CR<R>{(R)X.ra[j].re, (R)(-X.ra[j].im)};
    } else {
#pragma unroll
    for (int j = 0; j < EA; ++j) {
      const double re = X.oka[j] ? X.ra[j].re : 0.0;
      const double im = X.oka[j] ? X.ra[j].im * a_conj : 0.0;
      As[ak[j]][am[j]] = CR<R>{(R)re, (R)im};
    }
    }
    if (X.fullB && g.opB != OP_C) {
#pragma unroll
      for (int j = 0; j < EB; ++j) Bs[bk[j]][bn[j]] = CR<R>{(R)X.rb[j].re, (R)X.rb[j].im};
    } else if (X.fullB) {
#pragma unroll
      for (int j = 0; j < EB; ++j) Bs[bk[j]][bn[j]] = CR<R>{(R)X.rb[j].re, (R)(-X.rb[j].im)};
    } else {
#pragma unroll
    for (int j = 0; j < EB; ++j) {
      const double re = X.okb[j] ? X.rb[j].re : 0.0;
      const double im = X.okb[j] ? X.rb[j].im * b_conj : 0.0;
      Bs[bk[j]][bn[j]] = CR<R>{(R)re, (R)im};
    }
    }
  };

  using V = typename Acc<R>::V;
  // 3M (Karatsuba) complex product: T1 = Ar Br, T2 = Ai Bi, T3 = (Ar + Ai)(Br + Bi);
  // Re = T1 - T2, Im = T3 - T1 - T2.  Three real MFMAs per complex product
  // instead of four, normwise error of the same order as the 4-product form.
  V t1[TM][TN], t2[TM][TN], t3[TM][TN];
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j) {
      t1[i][j] = V{0, 0, 0, 0};
      t2[i][j] = V{0, 0, 0, 0};
      t3[i][j] = V{0, 0, 0, 0};
    }

  // one LDS buffer is read by the MFMAs while the next chunk is staged into
  // the other: a single barrier per chunk
  auto compute = [&](int buf) {
    auto& As = As2[buf];
    auto& Bs = Bs2[buf];
#pragma unroll
    for (int kk = 0; kk < BK; kk += 4) {
      const int kr = wk * BK + kk + (lane >> 4);
      CR<R> a[TM], b[TN];
#pragma unroll
      for (int q = 0; q < TM; ++q) a[q] = As[kr][wm * TM * 16 + q * 16 + (lane & 15)];
#pragma unroll
      for (int q = 0; q < TN; ++q) b[q] = Bs[kr][wn * TN * 16 + q * 16 + (lane & 15)];
      if constexpr (kThreeM<R>) {
        R as[TM], bs[TN];
#pragma unroll
        for (int q = 0; q < TM; ++q) as[q] = a[q].re + a[q].im;
#pragma unroll
        for (int q = 0; q < TN; ++q) bs[q] = b[q].re + b[q].im;
#pragma unroll
        for (int i = 0; i < TM; ++i)
#pragma unroll
          for (int j = 0; j < TN; ++j) t1[i][j] = mfma(a[i].re, b[j].re, t1[i][j]);
#pragma unroll
        for (int i = 0; i < TM; ++i)
#pragma unroll
          for (int j = 0; j < TN; ++j) t2[i][j] = mfma(a[i].im, b[j].im, t2[i][j]);
#pragma unroll
        for (int i = 0; i < TM; ++i)
#pragma unroll
          for (int j = 0; j < TN; ++j) t3[i][j] = mfma(as[i], bs[j], t3[i][j]);
      } else {
        // 4 real products: t1 = Re, t3 = Im (t2 unused)
#pragma unroll
        for (int i = 0; i < TM; ++i)
#pragma unroll
          for (int j = 0; j < TN; ++j) t1[i][j] = mfma(a[i].re, b[j].re, t1[i][j]);
#pragma unroll
        for (int i = 0; i < TM; ++i)
#pragma unroll
          for (int j = 0; j < TN; ++j) t3[i][j] = mfma(a[i].re, b[j].im, t3[i][j]);
#pragma unroll
        for (int i = 0; i < TM; ++i)
#pragma unroll
          for (int j = 0; j < TN; ++j) t1[i][j] = mfma(-a[i].im, b[j].im, t1[i][j]);
#pragma unroll
        for (int i = 0; i < TM; ++i)
#pragma unroll
          for (int j = 0; j < TN; ++j) t3[i][j] = mfma(a[i].im, b[j].re, t3[i][j]);
      }
    }
  };

  const int ktiles = (g.K + KC - 1) / KC;
  load_tiles(R0, 0);
  store_tiles(R0, 0);
  __syncthreads();
  for (int kt = 0; kt < ktiles; ++kt) {
    const bool more = kt + 1 < ktiles;
    if (more) load_tiles(R0, (kt + 1) * KC);
    compute(kt & 1);
    if (more) store_tiles(R0, (kt + 1) & 1);
    __syncthreads();
  }

  if constexpr (WK > 1) {
    // split-K: groups 1 .. WK-1 park their accumulators in LDS (the last
    // barrier of the K loop has retired every chunk read); group 0 adds them
    // in group order, so the sum is the same on every run, and writes C.
    constexpr int NV = 3 * TM * TN * 4;  // accumulator values per lane
    R* red = reinterpret_cast<R*>(lds_raw);
    if (wk > 0) {
      R* dst = red + ((wk - 1) * (C::WM * C::WN) + wmn) * NV * 64 + lane;
#pragma unroll
      for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int j = 0; j < TN; ++j)
#pragma unroll
          for (int v = 0; v < 4; ++v) {
            const int o = ((i * TN + j) * 4 + v) * 3;
            dst[(o + 0) * 64] = t1[i][j][v];
            dst[(o + 1) * 64] = t2[i][j][v];
            dst[(o + 2) * 64] = t3[i][j][v];
          }
    }
    __syncthreads();
    if (wk > 0) return;
#pragma unroll
    for (int s = 1; s < WK; ++s) {
      const R* src = red + ((s - 1) * (C::WM * C::WN) + wmn) * NV * 64 + lane;
#pragma unroll
      for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int j = 0; j < TN; ++j)
#pragma unroll
          for (int v = 0; v < 4; ++v) {
            const int o = ((i * TN + j) * 4 + v) * 3;
            t1[i][j][v] += src[(o + 0) * 64];
            t2[i][j][v] += src[(o + 1) * 64];
            t3[i][j][v] += src[(o + 2) * 64];
          }
    }
  }

  // epilogue
  cplx* Cp = g.C + tm.z_in * g.sC_in + tm.z_out * g.sC_out;
  const bool use_c0 = g.C0 != nullptr && (g.beta.re != 0.0 || g.beta.im != 0.0);
  const cplx* C0 = use_c0 ? g.C0 + tm.z_in * g.sC0_in + tm.z_out * g.sC0_out : nullptr;
  cplx sh{0.0, 0.0};
  const bool use_shift = g.shift != nullptr;
  if (use_shift) sh = g.shift[tm.z_out];
  // complex result of accumulator item (i, j, v)
  auto result = [&](int i, int j, int v, double& xr, double& xi) {
    if constexpr (kThreeM<R>) {
      const double p1 = (double)t1[i][j][v], p2 = (double)t2[i][j][v];
      xr = p1 - p2;
      xi = (double)t3[i][j][v] - p1 - p2;
    } else {
      xr = (double)t1[i][j][v];
      xi = (double)t3[i][j][v];
    }
  };
  if (fullM && fullN) {
    // interior tile: no bounds checks, all C0 loads issued before any use
    cplx c0v[TM][TN][4];
    if (use_c0) {
#pragma unroll
      for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int j = 0; j < TN; ++j)
#pragma unroll
          for (int v = 0; v < 4; ++v) {
            const int row = tm.m0 + wm * TM * 16 + i * 16 + acc_row<R>(lane, v);
            const int col = tm.n0 + wn * TN * 16 + j * 16 + (lane & 15);
            c0v[i][j][v] = ld_c(C0 + (long long)row * g.ldc0 + col);
          }
    }
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
      for (int j = 0; j < TN; ++j)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int row = tm.m0 + wm * TM * 16 + i * 16 + acc_row<R>(lane, v);
          const int col = tm.n0 + wn * TN * 16 + j * 16 + (lane & 15);
          double xr, xi;
          result(i, j, v, xr, xi);
          double outr = g.alpha.re * xr - g.alpha.im * xi;
          double outi = g.alpha.re * xi + g.alpha.im * xr;
          if (use_c0) {
            const cplx c0 = c0v[i][j][v];
            outr += g.beta.re * c0.re - g.beta.im * c0.im;
            outi += g.beta.re * c0.im + g.beta.im * c0.re;
          }
          if (use_shift && row == col) {
            outr += sh.re;
            outi += sh.im;
          }
          *reinterpret_cast<double2*>(Cp + (long long)row * g.ldc + col) = double2{outr, outi};
        }
    return;
  }
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int row = tm.m0 + wm * TM * 16 + i * 16 + acc_row<R>(lane, v);
        const int col = tm.n0 + wn * TN * 16 + j * 16 + (lane & 15);
        if (row < g.M && col < g.N) {
          double xr, xi;
          result(i, j, v, xr, xi);
          double outr = g.alpha.re * xr - g.alpha.im * xi;
          double outi = g.alpha.re * xi + g.alpha.im * xr;
          if (use_c0) {
            const cplx c0 = ld_c(C0 + (long long)row * g.ldc0 + col);
            outr += g.beta.re * c0.re - g.beta.im * c0.im;
            outi += g.beta.re * c0.im + g.beta.im * c0.re;
          }
          if (use_shift && row == col) {
            outr += sh.re;
            outi += sh.im;
          }
          *reinterpret_cast<double2*>(Cp + (long long)row * g.ldc + col) = double2{outr, outi};
        }
      }
}

// The tile configurations the launcher can choose from (index = config id).
// Every wave of the larger tiles keeps the 32 x 32 (2 x 2 MFMA tiles) shape of
// the default one; larger tiles add waves, not registers.
using Cfg64 = Cfg<2, 2, 2, 2, 2>;     // 0: 64 x 64, 4 waves
using Cfg32w4 = Cfg<2, 2, 1, 1, 4>;   // 1: 32 x 32, 4 waves of 16 x 16
using Cfg32w1 = Cfg<1, 1, 2, 2, 4>;   // 2: 32 x 32, 1 wave
using Cfg128x64 = Cfg<4, 2, 2, 2, 2>; // 3: 128 x 64, 8 waves
using Cfg64x128 = Cfg<2, 4, 2, 2, 2>; // 4: 64 x 128, 8 waves
using Cfg64x32 = Cfg<2, 2, 2, 1, 3>;  // 5: 64 x 32, 4 waves of 32 x 16 (3 workgroups per CU)
using Cfg32x64 = Cfg<2, 2, 1, 2, 3>;  // 6: 32 x 64, 4 waves of 16 x 32
// split-K inside the workgroup, for products too small or too few to fill the
// chip with output tiles alone (unbatched 64^3 .. 512^3)
using Cfg32k4 = Cfg<1, 1, 2, 2, 1, 4>;  // 7: 32 x 32, 4 k-slices of 32 x 32 waves
using Cfg16k4 = Cfg<1, 1, 1, 1, 2, 4>;  // 8: 16 x 16, 4 k-slices
using Cfg16k8 = Cfg<1, 1, 1, 1, 1, 8>;  // 9: 16 x 16, 8 k-slices
// the 32 x 32 tile of config 1 (4 waves of 16 x 16) with its k split 4 or 2
// ways: 16 (8) waves per workgroup, so one workgroup per CU still puts 4 (2)
// waves on each SIMD to cover the chunk loads of an unbatched 512^3 product
using Cfg32w4k4 = Cfg<2, 2, 1, 1, 4, 4>;  // 10: 32 x 32, 4 k-slices of 4 waves
using Cfg32w4k2 = Cfg<2, 2, 1, 1, 4, 2>;  // 11: 32 x 32, 2 k-slices of 4 waves
constexpr int kNumCfg = 12;
constexpr int kCfgBM[kNumCfg] = {Cfg64::BM,     Cfg32w4::BM,  Cfg32w1::BM,  Cfg128x64::BM,
                                 Cfg64x128::BM, Cfg64x32::BM, Cfg32x64::BM, Cfg32k4::BM,
                                 Cfg16k4::BM,   Cfg16k8::BM,  Cfg32w4k4::BM, Cfg32w4k2::BM};
constexpr int kCfgBN[kNumCfg] = {Cfg64::BN,     Cfg32w4::BN,  Cfg32w1::BN,  Cfg128x64::BN,
                                 Cfg64x128::BN, Cfg64x32::BN, Cfg32x64::BN, Cfg32k4::BN,
                                 Cfg16k4::BN,   Cfg16k8::BN,  Cfg32w4k4::BN, Cfg32w4k2::BN};
constexpr int kCfgWK[kNumCfg] = {1, 1, 1, 1, 1, 1, 1, Cfg32k4::WK, Cfg16k4::WK, Cfg16k8::WK,
                                 Cfg32w4k4::WK, Cfg32w4k2::WK};

template <typename R, typename C>
void launch_cfg(const GemmArgs& g, long long batch, hipStream_t s) {
  const int tiles_m = (g.M + C::BM - 1) / C::BM;
  const int tiles_n = (g.N + C::BN - 1) / C::BN;
  hipLaunchKernelGGL((gemm_kernel<R, C>), dim3((unsigned)((long long)tiles_m * tiles_n * batch)),
                     dim3(C::NT), 0, s, g, tiles_m, tiles_n);
}

template <typename R>
void launch_by_id(int id, const GemmArgs& g, long long batch, hipStream_t s) {
  switch (id) {
    case 1: launch_cfg<R, Cfg32w4>(g, batch, s); break;
    case 2: launch_cfg<R, Cfg32w1>(g, batch, s); break;
    case 3: launch_cfg<R, Cfg128x64>(g, batch, s); break;
    case 4: launch_cfg<R, Cfg64x128>(g, batch, s); break;
    case 5: launch_cfg<R, Cfg64x32>(g, batch, s); break;
    case 6: launch_cfg<R, Cfg32x64>(g, batch, s); break;
    case 7: launch_cfg<R, Cfg32k4>(g, batch, s); break;
    case 8: launch_cfg<R, Cfg16k4>(g, batch, s); break;
    case 9: launch_cfg<R, Cfg16k8>(g, batch, s); break;
    case 10: launch_cfg<R, Cfg32w4k4>(g, batch, s); break;
    case 11: launch_cfg<R, Cfg32w4k2>(g, batch, s); break;
    default: launch_cfg<R, Cfg64>(g, batch, s); break;
  }
}

// ---- tuned configurations: (shape key) -> config id, written by
// sinv_gemm_tune_set, read at every launch.  Fixed-size open table; not
// synchronised: set entries before launching work that uses them.
struct TuneEntry {
  int M, N, K, opA, opB, c0, prec, cfg;
  long long batch;
  bool used;
};
constexpr int kTuneSlots = 1024;
TuneEntry g_tune[kTuneSlots];
int g_tune_count = 0;

int tune_find(const GemmArgs& g, long long batch, int prec) {
  const int c0 = g.C0 != nullptr ? 1 : 0;
  for (int i = 0; i < g_tune_count; ++i) {
    const TuneEntry& e = g_tune[i];
    if (e.used && e.M == g.M && e.N == g.N && e.K == g.K && e.batch == batch && e.opA == g.opA &&
        e.opB == g.opB && e.c0 == c0 && e.prec == prec)
      return e.cfg;
  }
  return -1;
}

// Untuned shapes: the 64 x 64 tile, unless its grid leaves the chip under half
// full (under 2 workgroups per CU), where 32 x 32 tiles of 4 waves give 4x the
// workgroups.  Grids of 129 .. 640 such tiles (about one per CU) take the same
// tile with its k split 2 ways (config 11), or 4 ways up to 256 tiles with
// K >= 384 (config 10): two or four waves per SIMD then cover the chunk loads
// (graph-timed per launch: 512^3 21.8 / 27.0 us, 416^3 19.4 / 23.1 us,
// 384^3 x 2 29.1 / 31.8 us, 256^3 x 4 13.8 / 16.1 us against the plain 32 x 32
// tile).  Up to 128 such tiles (unbatched products of 256^3 or less), 16 x 16
// tiles with k split four ways (64^3 4.3 us, 128^3 5.4 us, 256^3 7.4 us
// against 5.8, 8.9 and 14.8 us with the 32 x 32 tile).
int default_cfg(const GemmArgs& g, long long batch) {
  const long long t64 = (long long)((g.M + 63) / 64) * ((g.N + 63) / 64) * batch;
  const long long t32 = (long long)((g.M + 31) / 32) * ((g.N + 31) / 32) * batch;
  if (t32 <= 128) return g.K >= 48 ? 8 : 1;
  if (t32 <= 640 && g.K >= 64) return (t32 <= 256 && g.K >= 384) ? 10 : 11;
  return t64 >= 512 ? 0 : 1;
}

int g_force_cfg = -1;  // >= 0: every launch uses this config (tuning runs)

// ---------------------------------------------------------------------------
// elementwise helpers (memory-bound; one thread per complex element)
__global__ void axpby_shift_kernel(int M, int N, cplx a, const cplx* D, int ldd, long long sD_in,
                                   long long sD_out, cplx* C, int ldc, long long sC_in,
                                   long long sC_out, const cplx* shift, int inner) {
  const int z = blockIdx.y;
  const int zo = z / inner, zi = z - zo * inner;
  const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (long long)M * N) return;
  const int r = (int)(e / N), c = (int)(e - (long long)r * N);
  const cplx d = ld_c(D + zi * sD_in + zo * sD_out + (long long)r * ldd + c);
  double outr = a.re * d.re - a.im * d.im;
  double outi = a.re * d.im + a.im * d.re;
  if (shift != nullptr && r == c) {
    outr += shift[zo].re;
    outi += shift[zo].im;
  }
  *reinterpret_cast<double2*>(C + zi * sC_in + zo * sC_out + (long long)r * ldc + c) =
      double2{outr, outi};
}

// C = alpha op(S) + shift I, op = T or C.  32 x 32 tiles through LDS so both
// the read and the write are coalesced.  shift (indexed by the outer batch
// index) is conjugated when conj_shift is set.
__global__ __launch_bounds__(256) void transpose_kernel(int M, int N, double conj, cplx alpha,
                                                        const cplx* S, int lds, long long sS_in,
                                                        long long sS_out, cplx* C, int ldc,
                                                        long long sC_in, long long sC_out,
                                                        const cplx* shift, int conj_shift,
                                                        int inner, int tiles_n) {
  __shared__ cplx t[32][33];
  const int z = blockIdx.y;
  const int zo = z / inner, zi = z - zo * inner;
  const int tile = blockIdx.x;
  const int r0 = (tile / tiles_n) * 32, c0 = (tile % tiles_n) * 32;  // output coords
  S += zi * sS_in + zo * sS_out;
  C += zi * sC_in + zo * sC_out;
  cplx sh{0.0, 0.0};
  if (shift != nullptr) {
    sh = shift[zo];
    if (conj_shift) sh.im = -sh.im;
  }
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 x 8
  // input is N x M: read S[c][r] for output (r, c)
  for (int y = ty; y < 32; y += 8) {
    const int sr = c0 + y, sc = r0 + tx;
    if (sr < N && sc < M) t[y][tx] = ld_c(S + (long long)sr * lds + sc);
  }
  __syncthreads();
  for (int y = ty; y < 32; y += 8) {
    const int r = r0 + y, c = c0 + tx;
    if (r < M && c < N) {
      const cplx v = t[tx][y];
      const double vr = v.re, vi = conj * v.im;
      double outr = alpha.re * vr - alpha.im * vi;
      double outi = alpha.re * vi + alpha.im * vr;
      if (r == c) {
        outr += sh.re;
        outi += sh.im;
      }
      *reinterpret_cast<double2*>(C + (long long)r * ldc + c) = double2{outr, outi};
    }
  }
}

__global__ void fill_kernel(cplx* p, long long n, cplx v) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) *reinterpret_cast<double2*>(p + i) = double2{v.re, v.im};
}

__global__ void fill_u64_kernel(unsigned long long* p, long long n, unsigned long long v) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

}  // namespace

hipError_t launch_gemm(const GemmArgs& g, Prec p, hipStream_t s) {
  if (g.M <= 0 || g.N <= 0 || g.inner <= 0 || g.outer <= 0) return hipSuccess;
  if (g.K <= 0) {
    // C = beta * C0 + shift: no product term
    if (g.C0 == nullptr) return hipErrorInvalidValue;
    return launch_axpby_shift(g.M, g.N, g.beta, g.C0, g.ldc0, g.sC0_in, g.sC0_out, g.C, g.ldc,
                              g.sC_in, g.sC_out, g.shift, g.inner, g.outer, s);
  }
  const long long batch = (long long)g.inner * g.outer;
  const int prec = p == Prec::F64 ? SINV_PREC_FP64 : SINV_PREC_MIXED;
  int id = g_force_cfg;
  if (id < 0) id = tune_find(g, batch, prec);
  if (id < 0) id = default_cfg(g, batch);
  const long long blocks = (long long)((g.M + kCfgBM[id] - 1) / kCfgBM[id]) *
                           ((g.N + kCfgBN[id] - 1) / kCfgBN[id]) * batch;
  if (blocks > 0x7fffffffLL) return hipErrorInvalidConfiguration;
  const int pid = prof_open(s);
  if (p == Prec::F64)
    launch_by_id<double>(id, g, batch, s);
  else
    launch_by_id<float>(id, g, batch, s);
  const double nb = (double)batch;
  prof_shape(pid, g.M, g.N, g.K, batch, g.opA | (g.opB << 2) | (g.C0 != nullptr ? 16 : 0));
  const double elems = (double)g.M * g.K + (double)g.K * g.N +
                       (double)g.M * g.N * (g.C0 != nullptr ? 2.0 : 1.0);
  if (g.prof_as > 0)
    prof_close(pid, s, g.prof_as - 1, 0.0, sizeof(cplx) * elems * nb);
  else
    prof_close(pid, s, p == Prec::F64 ? PROF_GEMM_F64 : PROF_GEMM_F32,
               8.0 * g.M * (double)g.N * g.K * nb, sizeof(cplx) * elems * nb);
  return hipGetLastError();
}

hipError_t launch_axpby_shift(int M, int N, cplx a, const cplx* D, int ldd, long long sD_in,
                              long long sD_out, cplx* C, int ldc, long long sC_in,
                              long long sC_out, const cplx* shift, int inner, int outer,
                              hipStream_t s) {
  const long long n = (long long)M * N;
  if (n <= 0 || inner * outer <= 0) return hipSuccess;
  dim3 grid((unsigned)((n + 255) / 256), (unsigned)(inner * outer));
  hipLaunchKernelGGL(axpby_shift_kernel, grid, dim3(256), 0, s, M, N, a, D, ldd, sD_in, sD_out, C,
                     ldc, sC_in, sC_out, shift, inner);
  return hipGetLastError();
}

hipError_t launch_transpose(int M, int N, int op, const cplx* S, int lds, long long sS_in,
                            long long sS_out, cplx* C, int ldc, long long sC_in,
                            long long sC_out, int inner, int outer, hipStream_t s, cplx alpha,
                            const cplx* shift, bool conj_shift) {
  if (M <= 0 || N <= 0 || inner * outer <= 0) return hipSuccess;
  const int tm = (M + 31) / 32, tn = (N + 31) / 32;
  dim3 grid((unsigned)(tm * tn), (unsigned)(inner * outer));
  hipLaunchKernelGGL(transpose_kernel, grid, dim3(256), 0, s, M, N, op == OP_C ? -1.0 : 1.0,
                     alpha, S, lds, sS_in, sS_out, C, ldc, sC_in, sC_out, shift,
                     conj_shift ? 1 : 0, inner, tn);
  return hipGetLastError();
}

hipError_t launch_fill(cplx* p, long long n, cplx v, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(fill_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, p, n, v);
  return hipGetLastError();
}

hipError_t launch_fill_u64(unsigned long long* p, long long n, unsigned long long v,
                           hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(fill_u64_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, p, n,
                     v);
  return hipGetLastError();
}

}  // namespace sinv

// ---------------------------------------------------------------------------
// C ABI
extern "C" int sinv_zgemm_batched(int M, int N, int K, int opA, int opB, sinv_c128 alpha,
                                  const void* A, int lda, long long strideA, const void* B,
                                  int ldb, long long strideB, sinv_c128 beta, const void* C0,
                                  int ldc0, long long strideC0, void* C, int ldc,
                                  long long strideC, const void* shift, int batch, int precision,
                                  void* stream) {
  using namespace sinv;
  if (M < 0 || N < 0 || K < 0 || batch < 0 || opA < 0 || opA > 2 || opB < 0 || opB > 2 ||
      (precision != SINV_PREC_FP64 && precision != SINV_PREC_MIXED) || C == nullptr ||
      (K > 0 && (A == nullptr || B == nullptr))) {
    set_error("sinv_zgemm_batched: invalid argument");
    return SINV_E_INVAL;
  }
  const int arows = opA == OP_N ? M : K, acols = opA == OP_N ? K : M;
  const int brows = opB == OP_N ? K : N, bcols = opB == OP_N ? N : K;
  (void)arows;
  (void)brows;
  if ((K > 0 && (lda < acols || ldb < bcols)) || ldc < N || (C0 != nullptr && ldc0 < N)) {
    set_error("sinv_zgemm_batched: leading dimension too small");
    return SINV_E_INVAL;
  }
  if (K == 0 && C0 == nullptr) {
    set_error("sinv_zgemm_batched: K == 0 needs C0");
    return SINV_E_INVAL;
  }
  GemmArgs g{};
  g.M = M; g.N = N; g.K = K; g.opA = opA; g.opB = opB;
  g.A = (const cplx*)A; g.lda = lda; g.sA_in = strideA; g.sA_out = 0;
  g.B = (const cplx*)B; g.ldb = ldb; g.sB_in = strideB; g.sB_out = 0;
  g.C0 = (const cplx*)C0; g.ldc0 = ldc0; g.sC0_in = strideC0; g.sC0_out = 0;
  g.C = (cplx*)C; g.ldc = ldc; g.sC_in = strideC; g.sC_out = 0;
  g.alpha = cplx{alpha.re, alpha.im};
  g.beta = cplx{beta.re, beta.im};
  // shift is indexed by the outer batch index: make every entry its own outer index
  g.shift = (const cplx*)shift;
  if (shift != nullptr) {
    g.inner = 1; g.outer = batch;
    g.sA_out = strideA; g.sA_in = 0; g.sB_out = strideB; g.sB_in = 0;
    g.sC0_out = strideC0; g.sC0_in = 0; g.sC_out = strideC; g.sC_in = 0;
  } else {
    g.inner = batch; g.outer = 1;
  }
  if (batch == 0) return SINV_OK;
  SINV_HIP_TRY(launch_gemm(g, precision == SINV_PREC_FP64 ? Prec::F64 : Prec::F32,
                           (hipStream_t)stream));
  return SINV_OK;
}

// ---------------------------------------------------------------------------
// GEMM tile configurations and the tuned-shape table (C ABI)
extern "C" int sinv_gemm_config_count(void) { return sinv::kNumCfg; }

extern "C" int sinv_gemm_config_tile(int cfg, int* bm, int* bn) {
  using namespace sinv;
  if (cfg < 0 || cfg >= kNumCfg || bm == nullptr || bn == nullptr) {
    set_error("sinv_gemm_config_tile: bad config id");
    return SINV_E_INVAL;
  }
  *bm = kCfgBM[cfg];
  *bn = kCfgBN[cfg];
  return SINV_OK;
}

extern "C" int sinv_gemm_config_ksplit(int cfg) {
  using namespace sinv;
  if (cfg < 0 || cfg >= kNumCfg) {
    set_error("sinv_gemm_config_ksplit: bad config id");
    return -SINV_E_INVAL;  // negative: never a k-split count
  }
  return kCfgWK[cfg];
}

extern "C" int sinv_gemm_force_config(int cfg) {
  using namespace sinv;
  if (cfg < -1 || cfg >= kNumCfg) {
    set_error("sinv_gemm_force_config: bad config id");
    return SINV_E_INVAL;
  }
  g_force_cfg = cfg;
  return SINV_OK;
}

extern "C" int sinv_gemm_tune_set(int M, int N, int K, long long batch, int opA, int opB,
                                  int has_c0, int precision, int cfg) {
  using namespace sinv;
  if (cfg < 0 || cfg >= kNumCfg || M <= 0 || N <= 0 || K <= 0 || batch <= 0) {
    set_error("sinv_gemm_tune_set: invalid argument");
    return SINV_E_INVAL;
  }
  const int c0 = has_c0 ? 1 : 0;
  for (int i = 0; i < g_tune_count; ++i) {
    TuneEntry& e = g_tune[i];
    if (e.M == M && e.N == N && e.K == K && e.batch == batch && e.opA == opA && e.opB == opB &&
        e.c0 == c0 && e.prec == precision) {
      e.cfg = cfg;
      e.used = true;
      return SINV_OK;
    }
  }
  if (g_tune_count >= kTuneSlots) {
    set_error("sinv_gemm_tune_set: table full");
    return SINV_E_INVAL;
  }
  g_tune[g_tune_count++] = TuneEntry{M, N, K, opA, opB, c0, precision, cfg, batch, true};
  return SINV_OK;
}

extern "C" int sinv_gemm_tune_clear(void) {
  sinv::g_tune_count = 0;
  return SINV_OK;
}

extern "C" int sinv_gemm_select(int M, int N, int K, long long batch, int opA, int opB,
                                int has_c0, int precision) {
  using namespace sinv;
  GemmArgs g{};
  g.M = M; g.N = N; g.K = K; g.opA = opA; g.opB = opB;
  g.C0 = has_c0 ? reinterpret_cast<const cplx*>(16) : nullptr;  // only tested against null
  if (g_force_cfg >= 0) return g_force_cfg;
  const int id = tune_find(g, batch, precision);
  return id >= 0 ? id : default_cfg(g, batch);
}
