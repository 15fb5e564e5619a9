// K1: batched in-place complex128 block inverse.
//
// Leaves (n <= 64): one workgroup per matrix, Gauss-Jordan elimination with
// partial pivoting on |re| + |im| (the modulus LAPACK's izamax uses inside
// ?getrf, which numpy.linalg.inv calls).  Default kernel: the block lives in
// registers, one row per lane, and the steps run panel by panel, one wave
// per panel (inv_leaf_kernel_panel; the pivot is the earliest row within
// 2^-15 of the column maximum).  The per-step register kernel
// (inv_leaf_kernel_rows, exact first maximum) and the older LDS-resident
// kernel (inv_leaf_kernel) are kept as selectable alternatives.
//
// Larger blocks: 2 x 2 Schur-complement recursion, all of whose arithmetic is
// batched GEMM on FP64 MFMA:
//   X11 = inv(A11); T21 = A21 X11; T12 = X11 A12; S = A22 - A21 T12
//   Y = inv(S); A12 <- -T12 Y; A21 <- -Y T21; A11 <- X11 - A12 T21
// Pivoting is therefore local to each leaf: the recursion needs nonsingular
// leading sub-blocks, which diagonally dominant blocks and blocks whose
// anti-Hermitian part is definite (the Schur complements of (E + i eta) I - H,
// eta > 0) always have.  Cost 8 n^3 real flops either way.
#include <cstdlib>
#include <cstring>

#include "sinv_core.hpp"

namespace sinv {

namespace {

constexpr int LEAF = kLeaf;

__device__ __forceinline__ cplx cmul(cplx a, cplx b) {
  return cplx{a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}

// Smith's complex reciprocal (no intermediate overflow for large moduli)
__device__ __forceinline__ cplx crecip(cplx p) {
  if (fabs(p.re) >= fabs(p.im)) {
    const double r = p.im / p.re;
    const double den = p.re + p.im * r;
    return cplx{1.0 / den, -r / den};
  } else {
    const double r = p.re / p.im;
    const double den = p.re * r + p.im;
    return cplx{r / den, -1.0 / den};
  }
}

__global__ __launch_bounds__(256) void inv_leaf_kernel(cplx* A, int n, int lda, long long s_in,
                                                       long long s_out, int inner,
                                                       unsigned long long* info, int tag) {
  __shared__ cplx a[LEAF][LEAF + 1];
  __shared__ cplx colk[LEAF];
  __shared__ int piv[LEAF];
  __shared__ int perm[LEAF];
  __shared__ int sp;
  __shared__ cplx spv;

  const int z = blockIdx.x;
  const int zo = z / inner, zi = z - zo * inner;
  A += zi * s_in + zo * s_out;
  const int tid = threadIdx.x;
  const int j = tid & 63;   // column owned in the elimination step
  const int i0 = tid >> 6;  // first row (rows i0, i0+4, ...)

  for (int i = i0; i < n; i += 4)
    if (j < n) {
      const double2 v = *reinterpret_cast<const double2*>(A + (long long)i * lda + j);
      a[i][j] = cplx{v.x, v.y};
    }
  __syncthreads();

  bool singular = false;
  for (int k = 0; k < n; ++k) {
    // --- pivot search on wave 0: first argmax of |re| + |im| over rows k..n-1
    if (tid < 64) {
      double v = -1.0;
      int idx = tid;
      if (tid >= k && tid < n) v = fabs(a[tid][k].re) + fabs(a[tid][k].im);
#pragma unroll
      for (int off = 32; off > 0; off >>= 1) {
        const double ov = __shfl_xor(v, off);
        const int oi = __shfl_xor(idx, off);
        if (ov > v || (ov == v && oi < idx)) {
          v = ov;
          idx = oi;
        }
      }
      // No comparable candidate (every entry NaN, or a lane >= n won the
      // reduction): keep row k, so the pivot is always a row in [k, n).
      if (idx < k || idx >= n) idx = k;
      if (tid == 0) {
        sp = idx;
        piv[k] = idx;
        spv = a[idx][k];
      }
      if (v == 0.0) singular = true;
    }
    __syncthreads();
    const int p = sp;
    const cplx d = crecip(spv);
    // --- swap rows k and p, scale the pivot row, save the multipliers of column k
    if (tid < n) {
      const cplx tk = a[k][tid], tp = a[p][tid];
      const cplx src = (tid == k) ? cplx{1.0, 0.0} : tp;
      a[k][tid] = cmul(src, d);
      if (p != k) a[p][tid] = tk;
      if (tid != k && tid != p) colk[tid] = a[tid][k];
      if (tid == k) colk[p] = (p != k) ? tk : cplx{0.0, 0.0};
    }
    __syncthreads();
    // --- eliminate column k from every other row (rank-1 update)
    if (j < n) {
      const cplx rk = a[k][j];
      for (int i = i0; i < n; i += 4) {
        if (i == k) continue;
        const cplx f = colk[i];
        const cplx cur = (j == k) ? cplx{0.0, 0.0} : a[i][j];
        a[i][j] = cplx{cur.re - (f.re * rk.re - f.im * rk.im), cur.im - (f.re * rk.im + f.im * rk.re)};
      }
    }
    __syncthreads();
  }

  // inv(A) = inv(P A) P: replay the row swaps as column swaps in reverse order
  if (tid == 0) {
    for (int c = 0; c < n; ++c) perm[c] = c;
    for (int k = n - 1; k >= 0; --k) {
      const int t = perm[k];
      perm[k] = perm[piv[k]];
      perm[piv[k]] = t;
    }
  }
  if (tid == 0 && singular && info != nullptr)
    atomicMin(info, ((unsigned long long)(unsigned)tag << 32) | (unsigned long long)(unsigned)z);
  __syncthreads();
  for (int i = i0; i < n; i += 4)
    if (j < n) {
      const cplx v = a[i][perm[j]];
      *reinterpret_cast<double2*>(A + (long long)i * lda + j) = double2{v.re, v.im};
    }
}

__device__ __forceinline__ double readlane_f64(double x, int lane) {
  const long long v = __double_as_longlong(x);
  const int lo = __builtin_amdgcn_readlane((int)v, lane);
  const int hi = __builtin_amdgcn_readlane((int)(v >> 32), lane);
  return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}

// DPP controls (gfx9 encoding) of the wave64 argmax below
constexpr int kDppXor1 = 0xB1;       // quad_perm [1,0,3,2]
constexpr int kDppXor2 = 0x4E;       // quad_perm [2,3,0,1]
constexpr int kDppHalfMirror = 0x141;
constexpr int kDppMirror = 0x140;
constexpr int kDppBcast15 = 0x142;
constexpr int kDppBcast31 = 0x143;

// One combine step of the pivot argmax: candidate (m, key) from the DPP
// source lane wins if its modulus is larger, or equal with a smaller key
// (key = position * 64 + row, so ties go to the first row in position order).
template <int CTRL, int ROWS>
__device__ __forceinline__ void amax_step(double& m, int& key) {
  const long long mb = __double_as_longlong(m);
  const int lo = __builtin_amdgcn_update_dpp((int)mb, (int)mb, CTRL, ROWS, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp((int)(mb >> 32), (int)(mb >> 32), CTRL, ROWS, 0xF,
                                             false);
  const int ok = __builtin_amdgcn_update_dpp(key, key, CTRL, ROWS, 0xF, false);
  const double om = __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
  // bitwise, not short-circuit: keeps the step free of exec-mask branches
  const bool take = (om > m) | ((om == m) & (ok < key));
  m = take ? om : m;
  key = take ? ok : key;
}

__device__ __forceinline__ void wave_fence() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Register-resident leaf, row layout: thread (w, lane i) keeps row i,
// columns w*R .. w*R+R-1 (R = 64/W), in registers for the whole elimination.
// Rows are never moved: each row carries its position in the row order that
// swapping Gauss-Jordan would have, the pivot search picks the first maximum
// of |re| + |im| in that order among unused rows (izamax's choice), and the
// positions are applied when the result is written.  Per step the wave that
// owns column k finds the pivot with a DPP argmax and publishes column k, the
// pivot row index/position and 1/pivot (one barrier, double-buffered); each
// wave then reads its segment of the pivot row from the pivot lane through a
// wave-private LDS slot and applies  a <- keep * a - g * P  to every element
// (g = a_ik / pivot, or -1/pivot on the pivot row, whose keep is 0; column
// k's entry of P is 1 and column k itself is cleared first).  Rows and
// columns past n are identity padding, untouched by the n steps.
template <int W>
__global__ __launch_bounds__(64 * W) void inv_leaf_kernel_rows(cplx* A, int n, int lda,
                                                               long long s_in, long long s_out,
                                                               int inner,
                                                               unsigned long long* info,
                                                               int tag) {
  constexpr int R = 64 / W;
  __shared__ cplx colk[2][64];
  __shared__ cplx dinv[2];
  __shared__ int pivrow[2];
  __shared__ int pivpos[64];
  __shared__ cplx prw[W][R];
  __shared__ int perm[64];
  __shared__ int invp[64];

  const int z = blockIdx.x;
  const int zo = z / inner, zi = z - zo * inner;
  A += zi * s_in + zo * s_out;
  const int lane = threadIdx.x & 63;
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int c0 = w * R;

  cplx a[R];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int c = c0 + r;
    if (lane < n && c < n) {
      const double2 v = *reinterpret_cast<const double2*>(A + (long long)lane * lda + c);
      a[r] = cplx{v.x, v.y};
    } else {
      a[r] = cplx{lane == c ? 1.0 : 0.0, 0.0};
    }
  }

  int pos = lane;  // position of this row in the swapped row order
  bool singular = false;
  for (int kb = 0; kb < W; ++kb) {
#pragma unroll
    for (int kr = 0; kr < R; ++kr) {
      const int k = kb * R + kr;
      if (k < n) {
        const int buf = k & 1;
        if (w == kb) {
          // ---- owner of column k: pivot search, publish column k and 1/pivot
          const cplx f = a[kr];
          double m = fabs(f.re) + fabs(f.im);
          if (!(m >= 0.0) || pos < k || lane >= n) m = -1.0;  // used rows, padding, NaN
          int key = pos * 64 + lane;
          amax_step<kDppXor1, 0xF>(m, key);
          amax_step<kDppXor2, 0xF>(m, key);
          amax_step<kDppHalfMirror, 0xF>(m, key);
          amax_step<kDppMirror, 0xF>(m, key);
          amax_step<kDppBcast15, 0xA>(m, key);
          amax_step<kDppBcast31, 0xC>(m, key);
          const long long mbits = __double_as_longlong(m);
          const double mw = __longlong_as_double(
              ((long long)__builtin_amdgcn_readlane((int)(mbits >> 32), 63) << 32) |
              (long long)(unsigned)__builtin_amdgcn_readlane((int)mbits, 63));
          int kw = __builtin_amdgcn_readlane(key, 63);
          int p, pp;
          if (mw >= 0.0) {
            p = kw & 63;
            pp = kw >> 6;
          } else {  // no comparable candidate: keep the row at position k
            const unsigned long long bal = __ballot(pos == k);
            p = (int)__builtin_ctzll(bal);
            pp = k;
          }
          if (mw == 0.0) singular = true;
          const cplx pv{readlane_f64(f.re, p), readlane_f64(f.im, p)};
          const cplx d = crecip(pv);
          colk[buf][lane] = f;
          if (lane == 0) {
            dinv[buf] = d;
            pivrow[buf] = p;
            pivpos[k] = pp;
          }
        }
        __syncthreads();
        // ---- every wave: eliminate column k with the published pivot
        const int p = __builtin_amdgcn_readfirstlane(pivrow[buf]);
        const int pp = pivpos[k];
        const cplx d = dinv[buf];
        const cplx f = colk[buf][lane];
        cplx g = cmul(f, d);
        double keep = 1.0;
        if (lane == p) {
          g = cplx{-d.re, -d.im};
          keep = 0.0;
        }
        if (lane == p) {
  #pragma unroll
          for (int r = 0; r < R; ++r) prw[w][r] = a[r];
        }
        wave_fence();
        if (w == kb) a[kr] = cplx{0.0, 0.0};
  #pragma unroll
        for (int r = 0; r < R; ++r) {
          const cplx P = (w == kb && r == kr) ? cplx{1.0, 0.0} : prw[w][r];
          a[r] = cplx{keep * a[r].re - (g.re * P.re - g.im * P.im),
                      keep * a[r].im - (g.re * P.im + g.im * P.re)};
        }
        wave_fence();
        // the row at position k and the pivot row exchange positions
        if (pos == k)
          pos = pp;
        else if (lane == p)
          pos = k;
      }
    }
  }

  // inv(A) = inv(P A) P, with P the row order built above: column c of the
  // eliminated matrix is column invp[c] of the result, row i goes to row pos
  if (threadIdx.x == 0) {
    for (int c = 0; c < n; ++c) perm[c] = c;
    for (int k = n - 1; k >= 0; --k) {
      const int t = perm[k];
      perm[k] = perm[pivpos[k]];
      perm[pivpos[k]] = t;
    }
    for (int c = 0; c < n; ++c) invp[perm[c]] = c;
  }
  if (lane == 0 && singular && info != nullptr)
    atomicMin(info, ((unsigned long long)(unsigned)tag << 32) | (unsigned long long)(unsigned)z);
  __syncthreads();
  if (lane < n) {
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int c = c0 + r;
      if (c < n)
        *reinterpret_cast<double2*>(A + (long long)pos * lda + invp[c]) =
            double2{a[r].re, a[r].im};
    }
  }
}

// Complex reciprocal of a wave-uniform value, without a division: the power
// of two sc = 2^(1023 - e), e the larger biased exponent of re and im (SALU
// bit arithmetic), brings the larger part into [1, 2) exactly; the squared
// modulus d in [1, 8) is inverted with v_rcp_f64 and two Newton steps; the
// result is sc (re' - i im') / d.  Error a few ulp, like Smith's method, at
// about a third of its latency.
__device__ __forceinline__ cplx crecip_uniform(cplx p) {
  const unsigned hre = (unsigned)((unsigned long long)__double_as_longlong(p.re) >> 32) & 0x7FFFFFFFu;
  const unsigned him = (unsigned)((unsigned long long)__double_as_longlong(p.im) >> 32) & 0x7FFFFFFFu;
  const unsigned e = min((hre > him ? hre : him) >> 20, 2045u);
  const double sc = __longlong_as_double((long long)(2046u - e) << 52);
  const double re = p.re * sc, im = p.im * sc;
  const double d = fma(re, re, im * im);
  double q = __builtin_amdgcn_rcp(d);
  q = fma(q, fma(-d, q, 1.0), q);
  q = fma(q, fma(-d, q, 1.0), q);
  q *= sc;
  return cplx{re * q, -im * q};
}

// Wave64 maximum of a 32-bit key, left in lane 63: six v_max_u32 with DPP
// sources (quad swaps, half-row and row mirrors, row broadcasts 15 and 31).
// Each DPP read follows the write of its source by the required two wait
// states; lanes outside a row mask keep their value.
__device__ __forceinline__ unsigned wave_umax(unsigned x) {
  asm volatile(
      "s_nop 1\n\t"
      "v_max_u32_dpp %0, %0, %0 quad_perm:[1,0,3,2] row_mask:0xf bank_mask:0xf\n\t"
      "s_nop 1\n\t"
      "v_max_u32_dpp %0, %0, %0 quad_perm:[2,3,0,1] row_mask:0xf bank_mask:0xf\n\t"
      "s_nop 1\n\t"
      "v_max_u32_dpp %0, %0, %0 row_half_mirror row_mask:0xf bank_mask:0xf\n\t"
      "s_nop 1\n\t"
      "v_max_u32_dpp %0, %0, %0 row_mirror row_mask:0xf bank_mask:0xf\n\t"
      "s_nop 1\n\t"
      "v_max_u32_dpp %0, %0, %0 row_bcast:15 row_mask:0xa bank_mask:0xf\n\t"
      "s_nop 1\n\t"
      "v_max_u32_dpp %0, %0, %0 row_bcast:31 row_mask:0xc bank_mask:0xf\n\t"
      "s_nop 1"
      : "+v"(x));
  return x;
}

// Pivot of one elimination step, as one wave-wide max of a 32-bit key: the
// high word of the modulus |re| + |im| (exponent and 20 mantissa bits;
// raised to at least the smallest normal double) with its 5 lowest bits
// replaced by 63 - position.  The winner is the earliest row, in position
// order, whose modulus is within 2^-15 relative of the column maximum
// (LAPACK's izamax takes the exact first maximum; the two choices coincide
// unless two candidates are that close, and either bounds the multipliers by
// 1 + 2^-15).  Rows already used, padding lanes and NaN moduli do not take
// part; with no candidate at all the row at position k is kept.  Returns the
// pivot lane; keys are distinct, so the lane is found by comparing keys.
__device__ __forceinline__ int pivot_lane(cplx f, int pos, int k, int lane, int n) {
  const double m = fabs(f.re) + fabs(f.im);
  const unsigned long long mbits = (unsigned long long)__double_as_longlong(m);
  const unsigned hi = (unsigned)(mbits >> 32);
  const bool cand = mbits <= 0x7FF0000000000000ull && pos >= k && lane < n;  // NaN excluded
  const unsigned hk = (max(hi, 0x00100000u) << 1) & ~63u;
  const unsigned mykey = cand ? (hk | (unsigned)(63 - pos)) : 0u;
  const unsigned top = (unsigned)__builtin_amdgcn_readlane((int)wave_umax(mykey), 63);
  const unsigned long long hit = __ballot(mykey == top);
  const unsigned long long keep_k = __ballot(pos == k);
  return (int)__builtin_ctzll(top != 0 ? hit : keep_k);
}

// Panel form of the register leaf: the same row layout and update formula as
// inv_leaf_kernel_rows, with the steps grouped by the wave that owns their
// columns (W waves of R = 64/W columns).  The owner of panel kb runs its R
// steps alone -- pivot search (pivot_lane), 1/pivot (crecip_uniform) and
// the update of its own columns, the next pivot column first -- as straight-
// line code, and publishes each step's multiplier vector g, pivot row and
// position; after ONE barrier per panel every other wave applies the R
// published rank-1 updates to its columns, reading the pivot row's entries
// from its own registers (lane p) as each step reaches them.  Every element
// receives the same updates in the same order as in a step-by-step
// elimination; the barriers drop from n to W.  Wave priorities follow the
// critical path: the factoring wave first, then the next panel's owner while
// it applies, then everyone else.
template <int R, bool FULL>
__device__ __forceinline__ void leaf_factor_panel(cplx (&a)[R], int kb, int n, int lane, int& pos,
                                                  bool& singular, cplx (*gb)[64], int* pb,
                                                  int* pivpos) {
  int p_of = 0, pp_of = 0;  // lane kr keeps step kr's pivot row and position
#pragma unroll
  for (int kr = 0; kr < R; ++kr) {
    const int k = kb * R + kr;
    if (FULL || k < n) {
      const cplx f = a[kr];
      const int p = pivot_lane(f, pos, k, lane, n);
      const int pp = __builtin_amdgcn_readlane(pos, p);
      const cplx piv{readlane_f64(f.re, p), readlane_f64(f.im, p)};
      singular |= piv.re == 0.0 && piv.im == 0.0;
      const cplx d = crecip_uniform(piv);
      const bool is_p = lane == p;
      const cplx gf = cmul(f, d);
      const cplx g{is_p ? -d.re : gf.re, is_p ? -d.im : gf.im};
      const double keep = is_p ? 0.0 : 1.0;
      gb[kr][lane] = g;
      p_of = lane == kr ? p : p_of;
      pp_of = lane == kr ? pp : pp_of;
#pragma unroll
      for (int j = 1; j <= R; ++j) {
        const int r = (kr + j) % R;  // column kr + 1 (the next pivot column) first
        cplx P{readlane_f64(a[r].re, p), readlane_f64(a[r].im, p)};
        cplx av = a[r];
        if (r == kr) {  // column k: cleared, unit entry in P
          P = cplx{1.0, 0.0};
          av = cplx{0.0, 0.0};
        }
        a[r] = cplx{keep * av.re - (g.re * P.re - g.im * P.im),
                    keep * av.im - (g.re * P.im + g.im * P.re)};
      }
      pos = pos == k ? pp : (is_p ? k : pos);
    }
  }
  if (lane < R && (FULL || kb * R + lane < n)) {
    pb[lane] = p_of;
    pivpos[kb * R + lane] = pp_of;
  }
}

template <int R, bool FULL>
__device__ __forceinline__ void leaf_apply_panel(cplx (&a)[R], int kb, int n, int lane, int& pos,
                                                 const cplx (*gb)[64], const int* pb,
                                                 const int* pivpos) {
#pragma unroll
  for (int kr = 0; kr < R; ++kr) {
    const int k = kb * R + kr;
    if (FULL || k < n) {
      const int p = __builtin_amdgcn_readfirstlane(pb[kr]);
      const int pp = __builtin_amdgcn_readfirstlane(pivpos[k]);
      const cplx g = gb[kr][lane];
      const double keep = lane == p ? 0.0 : 1.0;
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const cplx P{readlane_f64(a[r].re, p), readlane_f64(a[r].im, p)};
        a[r] = cplx{keep * a[r].re - (g.re * P.re - g.im * P.im),
                    keep * a[r].im - (g.re * P.im + g.im * P.re)};
      }
      pos = pos == k ? pp : (lane == p ? k : pos);
    }
  }
}

// A full panel's R steps applied column by column (each column takes the R
// steps in order, so every element sees the same operations as above): the
// first column is done after R updates, and the pivot search of the next
// panel, which needs only that column, can start while the other columns
// are still being updated.
template <int R>
__device__ __forceinline__ void leaf_apply_panel_cols(cplx (&a)[R], int kb, int lane, int& pos,
                                                      const cplx (*gb)[64], const int* pb,
                                                      const int* pivpos) {
  int p[R];
  cplx g[R];
  double keep[R];
#pragma unroll
  for (int kr = 0; kr < R; ++kr) {
    p[kr] = __builtin_amdgcn_readfirstlane(pb[kr]);
    g[kr] = gb[kr][lane];
    keep[kr] = lane == p[kr] ? 0.0 : 1.0;
  }
#pragma unroll
  for (int r = 0; r < R; ++r) {
#pragma unroll
    for (int kr = 0; kr < R; ++kr) {
      const cplx P{readlane_f64(a[r].re, p[kr]), readlane_f64(a[r].im, p[kr])};
      a[r] = cplx{keep[kr] * a[r].re - (g[kr].re * P.re - g[kr].im * P.im),
                  keep[kr] * a[r].im - (g[kr].re * P.im + g[kr].im * P.re)};
    }
  }
#pragma unroll
  for (int kr = 0; kr < R; ++kr) {
    const int k = kb * R + kr;
    const int pp = __builtin_amdgcn_readfirstlane(pivpos[k]);
    pos = pos == k ? pp : (lane == p[kr] ? k : pos);
  }
}

// Panel hand-off: with FLAGS every panel has its own LDS slot and a ready
// flag that its owner sets (release) once the slot is written; the other
// waves wait for that flag (acquire) instead of a workgroup barrier, so no
// wave waits for a slower one it does not depend on.  The wait is bounded
// (spin_limit polls, kLeafSpinLimit unless a test lowers it); a wave that
// runs out stops applying and the block is reported through info like a
// singular one.  Without FLAGS two slots alternate and one barrier per panel
// separates writers from readers.
constexpr int kLeafSpinLimit = 1 << 22;
int g_leaf_spin_limit = kLeafSpinLimit;

template <int W, bool FLAGS = false>
__global__ __launch_bounds__(64 * W) void inv_leaf_kernel_panel(cplx* A, int n, int lda,
                                                                long long s_in, long long s_out,
                                                                int inner,
                                                                unsigned long long* info,
                                                                int tag, int spin_limit) {
  constexpr int R = 64 / W;
  constexpr int NSLOT = FLAGS ? W : 2;
  __shared__ cplx gbuf[NSLOT][R][64];  // multiplier vector of each step of a panel
  __shared__ int pbuf[NSLOT][R];       // pivot row of each step
  __shared__ int ready[W];             // FLAGS: panel published
  __shared__ int pivpos[64];
  __shared__ int invp[64];

  const int z = blockIdx.x;
  const int zo = z / inner, zi = z - zo * inner;
  A += zi * s_in + zo * s_out;
  const int lane = threadIdx.x & 63;
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int c0 = w * R;

  cplx a[R];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int c = c0 + r;
    if (lane < n && c < n) {
      const double2 v = *reinterpret_cast<const double2*>(A + (long long)lane * lda + c);
      a[r] = cplx{v.x, v.y};
    } else {
      a[r] = cplx{lane == c ? 1.0 : 0.0, 0.0};
    }
  }

  int pos = lane;  // position of this row in the swapped row order
  bool singular = false;
  bool stuck = false;
  const int npanels = (n + R - 1) / R;
  if constexpr (FLAGS) {
    if (threadIdx.x < W) ready[threadIdx.x] = 0;
    __syncthreads();
  }
  bool prefactored = false;  // FLAGS: this wave's panel was factored (and published) early
  for (int kb = 0; kb < npanels; ++kb) {
    const int buf = FLAGS ? kb : (kb & 1);
    const bool full = kb * R + R <= n;
    if (w == kb && !prefactored) {
      __builtin_amdgcn_s_setprio(3);
      if (full)
        leaf_factor_panel<R, true>(a, kb, n, lane, pos, singular, gbuf[buf], pbuf[buf], pivpos);
      else
        leaf_factor_panel<R, false>(a, kb, n, lane, pos, singular, gbuf[buf], pbuf[buf], pivpos);
      if constexpr (FLAGS) {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        if (lane == 0) __hip_atomic_store(&ready[kb], 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
      }
    }
    if constexpr (!FLAGS) __syncthreads();
    if (w != kb) {
      if constexpr (FLAGS) {
        if (stuck) continue;
        int polls = 0;
        while (__hip_atomic_load(&ready[kb], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) == 0) {
          if (++polls > spin_limit) break;
          __builtin_amdgcn_s_sleep(1);
        }
        if (polls > spin_limit) {
          stuck = true;
          continue;
        }
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
        // the next panel's owner, both panels full: apply this one column by
        // column and factor its own right away, so that its first pivot
        // search overlaps the rest of the apply; publish it now
        if (w == kb + 1 && full && (kb + 1) * R + R <= n) {
          __builtin_amdgcn_s_setprio(3);
          leaf_apply_panel_cols<R>(a, kb, lane, pos, gbuf[buf], pbuf[buf], pivpos);
          leaf_factor_panel<R, true>(a, kb + 1, n, lane, pos, singular, gbuf[kb + 1],
                                     pbuf[kb + 1], pivpos);
          __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
          if (lane == 0)
            __hip_atomic_store(&ready[kb + 1], 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
          prefactored = true;
          continue;
        }
      }
      // the next panel's owner applies first: it is on the critical path
      if (w == kb + 1)
        __builtin_amdgcn_s_setprio(2);
      else
        __builtin_amdgcn_s_setprio(0);
      if (full)
        leaf_apply_panel<R, true>(a, kb, n, lane, pos, gbuf[buf], pbuf[buf], pivpos);
      else
        leaf_apply_panel<R, false>(a, kb, n, lane, pos, gbuf[buf], pbuf[buf], pivpos);
    }
  }

  // inv(A) = inv(P A) P: column c of the eliminated matrix is column invp[c]
  // of the result, where invp inverts the final row positions
  if (w == 0 && lane < n) invp[pos] = lane;
  if (lane == 0 && (singular || stuck) && info != nullptr)
    atomicMin(info, ((unsigned long long)(unsigned)tag << 32) | (unsigned long long)(unsigned)z);
  __syncthreads();
  if (lane < n) {
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int c = c0 + r;
      if (c < n)
        *reinterpret_cast<double2*>(A + (long long)pos * lda + invp[c]) =
            double2{a[r].re, a[r].im};
    }
  }
}

// Leaf kernel choice: the panel kernel with flag hand-off, 16 waves of 4
// columns when the launch leaves CUs idle (fewer than two matrices per CU:
// the per-step latency of the factoring wave is what counts) and 8 waves of
// 8 columns otherwise (fewer hand-offs; workgroups share CUs).
// sinv_leaf_force_variant, or initially SINV_LEAF_KERNEL ("panel16f",
// "panel8f", "panel16" / "panel8": one barrier per panel, "rows8": the
// register kernel with one barrier per step, "lds": the LDS-resident
// kernel), forces one.
int g_leaf_variant = -1;  // SINV_LEAF_* value; -1 until first read
int leaf_variant() {
  if (g_leaf_variant < 0) {
    const char* e = getenv("SINV_LEAF_KERNEL");
    int v = SINV_LEAF_AUTO;
    if (e != nullptr && strcmp(e, "lds") == 0) v = SINV_LEAF_LDS;
    if (e != nullptr && strcmp(e, "rows8") == 0) v = SINV_LEAF_ROWS8;
    if (e != nullptr && strcmp(e, "panel8") == 0) v = SINV_LEAF_PANEL8;
    if (e != nullptr && strcmp(e, "panel16") == 0) v = SINV_LEAF_PANEL16;
    if (e != nullptr && strcmp(e, "panel16f") == 0) v = SINV_LEAF_PANEL16_FLAGS;
    if (e != nullptr && strcmp(e, "panel8f") == 0) v = SINV_LEAF_PANEL8_FLAGS;
    g_leaf_variant = v;
  }
  return g_leaf_variant;
}
constexpr int kLeafPanel8MinGrid = 512;  // 2 x the 256 CUs of MI355X

// Split so that the recursion's GEMM dimensions stay multiples of the 64-wide
// GEMM tile where possible: the largest multiple of 64 not above n/2, else a
// multiple of 16 near n/2.
int split_point(int n) {
  int n1 = (n / 2) / 64 * 64;
  if (n1 == 0) n1 = ((n / 2 + 15) / 16) * 16;
  if (n1 >= n || n1 <= 0) n1 = n / 2;
  return n1;
}

}  // namespace

size_t inverse_workspace_elems(int n, int batch) {
  if (n <= LEAF) return 0;
  const int n1 = split_point(n), n2 = n - n1;
  const size_t own = 2ull * (size_t)n1 * (size_t)n2 * (size_t)batch;
  const size_t a = inverse_workspace_elems(n1, batch), b = inverse_workspace_elems(n2, batch);
  return own + (a > b ? a : b);
}

hipError_t launch_inverse(cplx* A, int n, int lda, long long s_in, long long s_out, int inner,
                          int outer, cplx* work, unsigned long long* info, int tag,
                          hipStream_t s) {
  if (n <= 0 || inner <= 0 || outer <= 0) return hipSuccess;
  if (n <= LEAF) {
    const int pid = prof_open(s);
    const dim3 grid((unsigned)(inner * outer));
    int v = leaf_variant();
    if (v == SINV_LEAF_AUTO)
      v = grid.x >= (unsigned)kLeafPanel8MinGrid ? SINV_LEAF_PANEL8_FLAGS
                                                 : SINV_LEAF_PANEL16_FLAGS;
    switch (v) {
      case SINV_LEAF_LDS:
        hipLaunchKernelGGL(inv_leaf_kernel, grid, dim3(256), 0, s, A, n, lda, s_in, s_out, inner,
                           info, tag);
        break;
      case SINV_LEAF_ROWS8:
        hipLaunchKernelGGL(inv_leaf_kernel_rows<8>, grid, dim3(512), 0, s, A, n, lda, s_in,
                           s_out, inner, info, tag);
        break;
      case SINV_LEAF_PANEL8:
        hipLaunchKernelGGL(inv_leaf_kernel_panel<8>, grid, dim3(512), 0, s, A, n, lda, s_in,
                           s_out, inner, info, tag, g_leaf_spin_limit);
        break;
      case SINV_LEAF_PANEL16_FLAGS:
        hipLaunchKernelGGL((inv_leaf_kernel_panel<16, true>), grid, dim3(1024), 0, s, A, n, lda,
                           s_in, s_out, inner, info, tag, g_leaf_spin_limit);
        break;
      case SINV_LEAF_PANEL8_FLAGS:
        hipLaunchKernelGGL((inv_leaf_kernel_panel<8, true>), grid, dim3(512), 0, s, A, n, lda,
                           s_in, s_out, inner, info, tag, g_leaf_spin_limit);
        break;
      default:
        hipLaunchKernelGGL(inv_leaf_kernel_panel<16>, grid, dim3(1024), 0, s, A, n, lda, s_in,
                           s_out, inner, info, tag, g_leaf_spin_limit);
    }
    prof_close(pid, s, PROF_INV_LEAF, 8.0 * n * (double)n * n * inner * (double)outer,
               2.0 * sizeof(cplx) * n * (double)n * inner * (double)outer);
    return hipGetLastError();
  }
  const int n1 = split_point(n), n2 = n - n1;
  cplx* A11 = A;
  cplx* A12 = A + n1;
  cplx* A21 = A + (long long)n1 * lda;
  cplx* A22 = A + (long long)n1 * lda + n1;
  const long long w = (long long)n1 * n2;
  const long long batch = (long long)inner * outer;
  cplx* W21 = work;              // n2 x n1, ld n1
  cplx* W12 = work + w * batch;  // n1 x n2, ld n2
  cplx* sub = work + 2 * w * batch;

  auto gemm = [&](int M, int N, int K, const cplx* Ap, int la, long long sa_in, long long sa_out,
                  const cplx* Bp, int lb, long long sb_in, long long sb_out, cplx* Cp, int lc,
                  long long sc_in, long long sc_out, cplx alpha, bool accumulate) {
    GemmArgs g{};
    g.M = M; g.N = N; g.K = K; g.opA = OP_N; g.opB = OP_N;
    g.A = Ap; g.lda = la; g.sA_in = sa_in; g.sA_out = sa_out;
    g.B = Bp; g.ldb = lb; g.sB_in = sb_in; g.sB_out = sb_out;
    g.C = Cp; g.ldc = lc; g.sC_in = sc_in; g.sC_out = sc_out;
    g.C0 = accumulate ? Cp : nullptr; g.ldc0 = lc; g.sC0_in = sc_in; g.sC0_out = sc_out;
    g.alpha = alpha;
    g.beta = accumulate ? cplx{1.0, 0.0} : cplx{0.0, 0.0};
    g.shift = nullptr;
    g.inner = inner; g.outer = outer;
    return launch_gemm(g, Prec::F64, s);
  };
  const cplx one{1.0, 0.0}, mone{-1.0, 0.0};
  const long long ws_in = w, ws_out = w * inner;

  SINV_HIP_RET(launch_inverse(A11, n1, lda, s_in, s_out, inner, outer, sub, info, tag, s));
  // T21 = A21 X11
  SINV_HIP_RET(gemm(n2, n1, n1, A21, lda, s_in, s_out, A11, lda, s_in, s_out, W21, n1, ws_in,
                    ws_out, one, false));
  // T12 = X11 A12
  SINV_HIP_RET(gemm(n1, n2, n1, A11, lda, s_in, s_out, A12, lda, s_in, s_out, W12, n2, ws_in,
                    ws_out, one, false));
  // S = A22 - A21 T12
  SINV_HIP_RET(gemm(n2, n2, n1, A21, lda, s_in, s_out, W12, n2, ws_in, ws_out, A22, lda, s_in,
                    s_out, mone, true));
  SINV_HIP_RET(launch_inverse(A22, n2, lda, s_in, s_out, inner, outer, sub, info, tag, s));
  // A12 = -T12 Y
  SINV_HIP_RET(gemm(n1, n2, n2, W12, n2, ws_in, ws_out, A22, lda, s_in, s_out, A12, lda, s_in,
                    s_out, mone, false));
  // A21 = -Y T21
  SINV_HIP_RET(gemm(n2, n1, n2, A22, lda, s_in, s_out, W21, n1, ws_in, ws_out, A21, lda, s_in,
                    s_out, mone, false));
  // A11 = X11 - A12 T21
  SINV_HIP_RET(gemm(n1, n1, n2, A12, lda, s_in, s_out, W21, n1, ws_in, ws_out, A11, lda, s_in,
                    s_out, mone, true));
  return hipSuccess;
}

}  // namespace sinv

extern "C" size_t sinv_zinv_workspace_bytes(int n, int batch) {
  if (n <= 0 || batch <= 0) return 0;
  return sinv::inverse_workspace_elems(n, batch) * sizeof(sinv::cplx);
}

extern "C" int sinv_zinv_batched(void* A, int n, int lda, long long strideA, int batch, void* work,
                                 size_t work_bytes, unsigned long long* info, void* stream) {
  using namespace sinv;
  if (A == nullptr || n < 0 || batch < 0 || lda < n) {
    set_error("sinv_zinv_batched: invalid argument");
    return SINV_E_INVAL;
  }
  if (work_bytes < sinv_zinv_workspace_bytes(n, batch) || (n > kLeaf && work == nullptr)) {
    set_error("sinv_zinv_batched: workspace too small (%zu < %zu)", work_bytes,
              sinv_zinv_workspace_bytes(n, batch));
    return SINV_E_INVAL;
  }
  if (n == 0 || batch == 0) return SINV_OK;
  hipStream_t s = (hipStream_t)stream;
  if (info != nullptr) SINV_HIP_TRY(launch_fill_u64(info, 1, SINV_INFO_OK, s));
  SINV_HIP_TRY(launch_inverse((cplx*)A, n, lda, strideA, 0, batch, 1, (cplx*)work, info, 0, s));
  return SINV_OK;
}

extern "C" int sinv_leaf_force_variant(int variant) {
  using namespace sinv;
  if (variant != SINV_LEAF_AUTO && variant != SINV_LEAF_LDS && variant != SINV_LEAF_ROWS8 &&
      variant != SINV_LEAF_PANEL8 && variant != SINV_LEAF_PANEL16 &&
      variant != SINV_LEAF_PANEL16_FLAGS && variant != SINV_LEAF_PANEL8_FLAGS) {
    set_error("sinv_leaf_force_variant: unknown variant %d", variant);
    return SINV_E_INVAL;
  }
  g_leaf_variant = variant;
  return SINV_OK;
}

extern "C" int sinv_leaf_spin_limit(int polls) {
  using namespace sinv;
  if (polls < -1) {
    set_error("sinv_leaf_spin_limit: polls must be >= 0, or -1 for the default");
    return SINV_E_INVAL;
  }
  g_leaf_spin_limit = polls < 0 ? kLeafSpinLimit : polls;
  return SINV_OK;
}
