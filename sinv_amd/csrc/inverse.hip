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
// Larger blocks (64 < n <= kMaxInverseN): Gauss-Jordan elimination with
// partial pivoting over all n rows (exact first maximum of |re| + |im| in
// position order, as izamax), blocked by columns: tall panels of 64, 32, 16
// or 8 columns are eliminated over all rows by a register-resident kernel
// (gj_flow_kernel up to 1024 rows, gj_step_kernel above and for 321 .. 384
// rows in launches of 64 or more blocks, the L2-resident
// gj_panel_global_kernel above 2048 rows), and each panel's elimination
// reaches the other columns through batched GEMMs on FP64 MFMA.  Cost 8 n^3
// real flops either way.
#include <algorithm>
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

// ---------------------------------------------------------------------------
// Blocks larger than the leaf: Gauss-Jordan elimination with partial pivoting
// over all n rows, blocked by columns.
//
// Rows never move.  Row r carries pos[r], its position in the row order that a
// swapping elimination would have.  The pivot of column k is the first
// maximum of |re| + |im|, in position order, among the rows with pos >= k (the
// rows not yet used as pivots): izamax's choice inside LAPACK's zgetrf, which
// numpy.linalg.inv calls.  The pivot row takes position k and the row that
// held position k takes the pivot's old position.
//
// Eliminating a set of columns J with pivot rows R = piv(J) in place (the
// eliminated columns replaced by the inverse's columns) changes every other
// column c as
//     A[:, c] <- A[:, J] A[R, c] + (A[:, c] with rows R cleared)
// with A[:, J] the eliminated panel.  A column range is therefore eliminated
// recursively,
//     GJ(J1);  B = A[piv(J1), J2], A[piv(J1), J2] = 0;  A[:, J2] += A[:, J1] B
//     GJ(J2);  B = A[piv(J2), J1], A[piv(J2), J1] = 0;  A[:, J1] += A[:, J2] B
// with the same pivots as column-by-column elimination.  The products are
// batched GEMMs on FP64 MFMA (n x n2 x n1 each); a range of at most one panel
// is eliminated over all n rows by a tall-panel kernel.  Panels and products
// together are n^3 complex multiply-adds (8 n^3 flops).  The inverse is then
// E[piv[r]][pos[s]] at (r, s), E the eliminated matrix: one permuted copy.

// Tall panel in registers: 8 waves; lane l holds rows l + 64 s (s < M), wave
// v holds the panel's columns v R .. v R + R - 1.  The wave owning a column
// finds its pivot (per-lane maximum over its M rows, then a wave-wide maximum
// of the modulus bits in three 32-bit DPP passes -- high word, low word, then
// the smallest position among exact ties), forms the multipliers
// g = a_ik / pivot (-1 / pivot on the pivot row, whose entries are replaced
// rather than updated) and publishes them with the pivot row's index; after
// one barrier per R steps every other wave applies the R rank-1 updates
// a <- keep a - g P to its columns, reading the pivot row's entries P from
// its own registers.  Every element sees the same operations in the same
// order as in column-by-column elimination.
constexpr int kGjWaves = 8;

// Entries of the row at (lane plane, slot) in this wave's R columns, read
// with readlane under branches on the wave-uniform slot.
template <int M, int R>
__device__ __forceinline__ void gj_row(const cplx (&a)[M][R], int slot, int plane, cplx (&P)[R]) {
#pragma unroll
  for (int s = 0; s < M; ++s)
    if (s == slot) {
#pragma unroll
      for (int r = 0; r < R; ++r)
        P[r] = cplx{readlane_f64(a[s][r].re, plane), readlane_f64(a[s][r].im, plane)};
    }
}

// keep a - g P, keep = 0 on the pivot row (whose entries are replaced:
// g = -1 / pivot there, so the row becomes P / pivot) and 1 elsewhere: six
// FP64 operations, the 0/1 factor folded into an FMA instead of selects
__device__ __forceinline__ cplx gj_upd(cplx a, cplx g, cplx P, double keep) {
  const double ti = g.im * P.im, tr = g.im * P.re;
  return cplx{fma(-g.re, P.re, fma(a.re, keep, ti)), fma(-g.re, P.im, fma(a.im, keep, -tr))};
}

// Pivot of global column kg, held in this wave's column kr: the first maximum
// of |re| + |im| in position order among the rows with pos >= kg.  Returns
// its lane and slot (row = lane + 64 slot), its position and whether its
// modulus is zero.  NaN moduli do not take part; a column with no candidate
// keeps the row at position kg.  All lanes of the wave must be active.
template <int M, int R>
__device__ __forceinline__ void gj_pivot(const cplx (&a)[M][R], const int (&pos)[M], int kr,
                                         int kg, int n, int lane, int& lp, int& sp, int& pp,
                                         bool& zero) {
  double mb = -1.0;
  int pb = 0x7fffffff, sb = 0;
#pragma unroll
  for (int s = 0; s < M; ++s) {
    const cplx f = a[s][kr];
    const double m = fabs(f.re) + fabs(f.im);
    const bool cand = (lane + 64 * s < n) & (pos[s] >= kg) & (m == m);
    const bool better = cand & ((m > mb) | ((m == mb) & (pos[s] < pb)));
    mb = better ? m : mb;
    pb = better ? pos[s] : pb;
    sb = better ? s : sb;
  }
  // a non-negative double orders like its bit pattern: compare the high
  // words (bit 31 set marks a candidate), then the low words, then positions
  const unsigned long long bits = (unsigned long long)__double_as_longlong(mb);
  const unsigned hi = (unsigned)(bits >> 32), lo = (unsigned)bits;
  const unsigned k1 = mb >= 0.0 ? (hi | 0x80000000u) : 0u;
  const unsigned t1 = (unsigned)__builtin_amdgcn_readlane((int)wave_umax(k1), 63);
  zero = false;
  if (t1 == 0u) {
    int hs = -1;
#pragma unroll
    for (int s = 0; s < M; ++s) hs = ((lane + 64 * s < n) & (pos[s] == kg)) ? s : hs;
    lp = (int)__builtin_ctzll(__ballot(hs >= 0));
    sp = __builtin_amdgcn_readlane(hs, lp);
    pp = kg;
    return;
  }
  bool tie = k1 == t1;
  unsigned long long bal = __ballot(tie);
  if (__builtin_popcountll(bal) > 1) {
    const unsigned t2 = (unsigned)__builtin_amdgcn_readlane((int)wave_umax(tie ? lo : 0u), 63);
    tie = tie & (lo == t2);
    bal = __ballot(tie);
    if (__builtin_popcountll(bal) > 1) {
      const unsigned k3 = tie ? ~(unsigned)pb : 0u;
      const unsigned t3 = (unsigned)__builtin_amdgcn_readlane((int)wave_umax(k3), 63);
      bal = __ballot(tie & (k3 == t3));
    }
  }
  lp = (int)__builtin_ctzll(bal);
  sp = __builtin_amdgcn_readlane(sb, lp);
  pp = __builtin_amdgcn_readlane(pb, lp);
  zero = t1 == 0x80000000u && (unsigned)__builtin_amdgcn_readlane((int)lo, lp) == 0u;
}

// The owner wave's R elimination steps of panel kb (columns c0 + kb R + kr):
// pivot, multipliers g = a_ik / pivot (-1 / pivot on the pivot row) published
// to gb with the pivot row, own columns updated, positions moved (the pivot
// row takes position kg, the row that held kg the pivot's old position).
template <int M, int R, bool FULL>
__device__ __forceinline__ void gj_factor(cplx (&a)[M][R], int (&pos)[M], int kb, int c0, int wc,
                                          int n, int lane, cplx* gb, int* pb, int* ppb,
                                          int* pivl, bool& singular) {
#pragma unroll
  for (int kr = 0; kr < R; ++kr) {
    const int k = kb * R + kr;
    if (FULL || k < wc) {
      const int kg = c0 + k;
      int lp, sp, pp;
      bool zero;
      gj_pivot<M, R>(a, pos, kr, kg, n, lane, lp, sp, pp, zero);
      singular |= zero;
      const int p = lp + 64 * sp;
      cplx P[R];
      gj_row<M, R>(a, sp, lp, P);
      const cplx d = crecip_uniform(P[kr]);
      cplx g[M];
#pragma unroll
      for (int s = 0; s < M; ++s) {
        const cplx gf = cmul(a[s][kr], d);
        const bool is_p = lane + 64 * s == p;
        g[s] = cplx{is_p ? -d.re : gf.re, is_p ? -d.im : gf.im};
        gb[(kr * M + s) * 64 + lane] = g[s];
        a[s][kr] = cplx{-g[s].re, -g[s].im};  // the inverse's column: -g, d on the pivot row
      }
#pragma unroll
      for (int j = 1; j < R; ++j) {
        const int r = (kr + j) % R;  // the next pivot column first
#pragma unroll
        for (int s = 0; s < M; ++s)
          a[s][r] = gj_upd(a[s][r], g[s], P[r], lane + 64 * s == p ? 0.0 : 1.0);
      }
#pragma unroll
      for (int s = 0; s < M; ++s) pos[s] = pos[s] == kg ? pp : (lane + 64 * s == p ? kg : pos[s]);
      if (lane == 0) {
        pb[kr] = p;
        ppb[kr] = pp;
        pivl[k] = p;
      }
    }
  }
}

// Every other wave: the R published steps of panel kb applied to its columns.
template <int M, int R, bool FULL>
__device__ __forceinline__ void gj_apply(cplx (&a)[M][R], int (&pos)[M], int kb, int c0, int wc,
                                         int lane, const cplx* gb, const int* pb,
                                         const int* ppb) {
#pragma unroll
  for (int kr = 0; kr < R; ++kr) {
    const int k = kb * R + kr;
    if (FULL || k < wc) {
      const int kg = c0 + k;
      const int p = __builtin_amdgcn_readfirstlane(pb[kr]);
      const int pp = __builtin_amdgcn_readfirstlane(ppb[kr]);
      cplx P[R];
      gj_row<M, R>(a, p >> 6, p & 63, P);
#pragma unroll
      for (int s = 0; s < M; ++s) {
        const cplx g = gb[(kr * M + s) * 64 + lane];
        const bool is_p = lane + 64 * s == p;
#pragma unroll
        for (int r = 0; r < R; ++r) a[s][r] = gj_upd(a[s][r], g, P[r], is_p ? 0.0 : 1.0);
      }
#pragma unroll
      for (int s = 0; s < M; ++s) pos[s] = pos[s] == kg ? pp : (lane + 64 * s == p ? kg : pos[s]);
    }
  }
}

// Eliminates columns [c0, c0 + wc) (wc <= 8 R) of every n x n matrix over all
// its rows, n <= 64 M.  ipos / ipiv (n ints per matrix) carry the row
// positions between launches (initialised here when c0 == 0) and receive the
// pivot row of each eliminated column.
template <int M, int R>
__global__ __launch_bounds__(64 * kGjWaves) void gj_panel_kernel(
    cplx* A, int n, int lda, long long s_in, long long s_out, int inner, int c0, int wc,
    int* ipos, int* ipiv, unsigned long long* info, int tag) {
  __shared__ cplx gbuf[2][R * M * 64];  // multipliers of the panel's steps, two panels
  __shared__ int pbuf[2][R];             // pivot row of each step
  __shared__ int ppbuf[2][R];            // its position before the step
  __shared__ int pivl[kGjWaves * R];     // pivot rows of the launch's columns
  const int z = blockIdx.x;
  const int zo = z / inner, zi = z - zo * inner;
  A += zi * s_in + zo * s_out;
  ipos += (long long)z * n;
  ipiv += (long long)z * n;
  const int lane = threadIdx.x & 63;
  const int v = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);

  cplx a[M][R];
  int pos[M];
#pragma unroll
  for (int s = 0; s < M; ++s) {
    const int row = lane + 64 * s;
    const cplx* src = A + (long long)row * lda + c0 + v * R;
#pragma unroll
    for (int r = 0; r < R; ++r) {
      if (row < n && v * R + r < wc) {
        const double2 t = *reinterpret_cast<const double2*>(src + r);
        a[s][r] = cplx{t.x, t.y};
      } else {
        a[s][r] = cplx{0.0, 0.0};
      }
    }
    pos[s] = row < n ? (c0 == 0 ? row : ipos[row]) : -1;
  }

  bool singular = false;
  for (int kb = 0; kb < kGjWaves && kb * R < wc; ++kb) {
    const int buf = kb & 1;
    const bool full = kb * R + R <= wc;
    if (v == kb) {
      // wave priorities follow the critical path: the factoring wave first,
      // then the next panel's owner while it applies this panel
      __builtin_amdgcn_s_setprio(3);
      if (full)
        gj_factor<M, R, true>(a, pos, kb, c0, wc, n, lane, gbuf[buf], pbuf[buf], ppbuf[buf], pivl,
                              singular);
      else
        gj_factor<M, R, false>(a, pos, kb, c0, wc, n, lane, gbuf[buf], pbuf[buf], ppbuf[buf],
                               pivl, singular);
    }
    __syncthreads();
    if (v != kb) {
      if (v == kb + 1)
        __builtin_amdgcn_s_setprio(2);
      else
        __builtin_amdgcn_s_setprio(0);
      if (full)
        gj_apply<M, R, true>(a, pos, kb, c0, wc, lane, gbuf[buf], pbuf[buf], ppbuf[buf]);
      else
        gj_apply<M, R, false>(a, pos, kb, c0, wc, lane, gbuf[buf], pbuf[buf], ppbuf[buf]);
    }
  }

#pragma unroll
  for (int s = 0; s < M; ++s) {
    const int row = lane + 64 * s;
    if (row < n) {
      cplx* dst = A + (long long)row * lda + c0 + v * R;
#pragma unroll
      for (int r = 0; r < R; ++r)
        if (v * R + r < wc) *reinterpret_cast<double2*>(dst + r) = double2{a[s][r].re, a[s][r].im};
      if (v == 0) ipos[row] = pos[s];
    }
  }
  if ((int)threadIdx.x < wc) ipiv[c0 + threadIdx.x] = pivl[threadIdx.x];
  if (lane == 0 && singular && info != nullptr)
    atomicMin(info, ((unsigned long long)(unsigned)tag << 32) | (unsigned long long)(unsigned)z);
}

// Step form of the tall-panel kernel: the same layout (8 waves, lane l holds
// rows l + 64 s, wave v columns v R .. v R + R - 1) with one barrier per
// column instead of per panel.  The owner of column k finds the pivot and
// publishes the multipliers g and the pivot row index; after the barrier every
// wave, the owner included, applies step k to its columns with one code path
// (the pivot column itself is taken as cleared with a unit pivot-row entry,
// which writes the inverse's column -g there).  The critical path per column
// is search, publish, barrier and the next owner's update; the wave that owns
// the next column applies at raised priority.
template <int M, int R>
__global__ __launch_bounds__(64 * kGjWaves) void gj_step_kernel(
    cplx* A, int n, int lda, long long s_in, long long s_out, int inner, int c0, int wc,
    int* ipos, int* ipiv, unsigned long long* info, int tag) {
  __shared__ cplx gbuf[2][M * 64];  // multipliers of the step, two steps
  __shared__ int pinfo[2][2];        // pivot row and its position before the step
  __shared__ int pivl[kGjWaves * R];
  const int z = blockIdx.x;
  const int zo = z / inner, zi = z - zo * inner;
  A += zi * s_in + zo * s_out;
  ipos += (long long)z * n;
  ipiv += (long long)z * n;
  const int lane = threadIdx.x & 63;
  const int v = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);

  cplx a[M][R];
  int pos[M];
#pragma unroll
  for (int s = 0; s < M; ++s) {
    const int row = lane + 64 * s;
    const cplx* src = A + (long long)row * lda + c0 + v * R;
#pragma unroll
    for (int r = 0; r < R; ++r) {
      if (row < n && v * R + r < wc) {
        const double2 t = *reinterpret_cast<const double2*>(src + r);
        a[s][r] = cplx{t.x, t.y};
      } else {
        a[s][r] = cplx{0.0, 0.0};
      }
    }
    pos[s] = row < n ? (c0 == 0 ? row : ipos[row]) : -1;
  }

  bool singular = false;
  for (int kb = 0; kb * R < wc; ++kb) {
#pragma unroll
    for (int kr = 0; kr < R; ++kr) {
      const int k = kb * R + kr;
      if (k >= wc) break;  // uniform
      const int buf = k & 1;
      const int kg = c0 + k;
      if (v == kb) {
        __builtin_amdgcn_s_setprio(3);
        int lp, sp, pp;
        bool zero;
        gj_pivot<M, R>(a, pos, kr, kg, n, lane, lp, sp, pp, zero);
        singular |= zero;
        const int p = lp + 64 * sp;
        cplx piv{0.0, 0.0};
#pragma unroll
        for (int s = 0; s < M; ++s)
          if (s == sp) piv = cplx{readlane_f64(a[s][kr].re, lp), readlane_f64(a[s][kr].im, lp)};
        const cplx d = crecip_uniform(piv);
#pragma unroll
        for (int s = 0; s < M; ++s) {
          const cplx gf = cmul(a[s][kr], d);
          const bool is_p = lane + 64 * s == p;
          gbuf[buf][s * 64 + lane] = cplx{is_p ? -d.re : gf.re, is_p ? -d.im : gf.im};
        }
        if (lane == 0) {
          pinfo[buf][0] = p;
          pinfo[buf][1] = pp;
          pivl[k] = p;
        }
      }
      __syncthreads();
      // every wave: step k on its columns, the next column's owner at raised priority
      if (v == (k + 1) / R)
        __builtin_amdgcn_s_setprio(2);
      else
        __builtin_amdgcn_s_setprio(0);
      const int p = __builtin_amdgcn_readfirstlane(pinfo[buf][0]);
      const int pp = __builtin_amdgcn_readfirstlane(pinfo[buf][1]);
      cplx P[R];
      gj_row<M, R>(a, p >> 6, p & 63, P);
      const bool mine = v == kb;
#pragma unroll
      for (int s = 0; s < M; ++s) {
        const cplx g = gbuf[buf][s * 64 + lane];
        const double keep = lane + 64 * s == p ? 0.0 : 1.0;
#pragma unroll
        for (int r = 0; r < R; ++r) {
          // the pivot column: cleared, with a unit pivot-row entry (writes -g)
          const bool pc = r == kr && mine;
          a[s][r] = gj_upd(a[s][r], g, pc ? cplx{1.0, 0.0} : P[r], pc ? 0.0 : keep);
        }
      }
#pragma unroll
      for (int s = 0; s < M; ++s)
        pos[s] = pos[s] == kg ? pp : (lane + 64 * s == p ? kg : pos[s]);
    }
  }

#pragma unroll
  for (int s = 0; s < M; ++s) {
    const int row = lane + 64 * s;
    if (row < n) {
      cplx* dst = A + (long long)row * lda + c0 + v * R;
#pragma unroll
      for (int r = 0; r < R; ++r)
        if (v * R + r < wc) *reinterpret_cast<double2*>(dst + r) = double2{a[s][r].re, a[s][r].im};
      if (v == 0) ipos[row] = pos[s];
    }
  }
  __syncthreads();
  if ((int)threadIdx.x < wc) ipiv[c0 + threadIdx.x] = pivl[threadIdx.x];
  if (lane == 0 && singular && info != nullptr)
    atomicMin(info, ((unsigned long long)(unsigned)tag << 32) | (unsigned long long)(unsigned)z);
}

// Flow form of the tall-panel kernel (the default up to 1024 rows): 8 waves,
// lane l holds rows l + 64 s (s < M); the panel's columns are dealt round
// robin, wave v holding columns v + 8 j (j < R), so consecutive steps are
// produced by different waves.  Steps are handed from wave to wave through
// LDS slots and flags instead of barriers:
//  * Step k (column k, owned by wave k % 8) is published in slot k % NS: the
//    multipliers g (0 on the pivot row), the pivot row's index and a sequence
//    flag (k + 1, release); every wave applies it (acquire) and then counts
//    itself out of the slot, and the producer of step k + NS waits for all
//    eight before reusing it.
//  * The wave that produces step k + 1 applies step k to that column first,
//    then searches and publishes step k + 1, then applies step k to its
//    other columns; the other waves' updates are off that path.
//  * The update is a <- a - g P on every row: the pivot row keeps its values
//    unscaled (g = 0 there) and is scaled by 1 / pivot when the panel is
//    written back; linear updates commute with that scaling.  Four FP64 FMAs
//    per element.  The pivot column becomes -g (1 on the pivot row), scaled
//    the same way.
//  * Every wave tracks the rows already used as pivots in one 64-bit scalar
//    mask per slot, so the pivot search needs no positions: the largest
//    modulus is one max over 32-bit keys (the high word of |re| + |im|); only
//    exact ties read positions, to take the first in position order, as
//    izamax does.  Row positions live once in LDS, moved by the producer.
// Every wait is bounded (spin_limit polls); a wave that runs out stops and
// the matrix is reported through info like a singular one.
template <int M>
struct GjFlow {
  static constexpr int NS = M <= 8 ? 4 : 3;  // slots: at most 48 KB of multipliers
};
constexpr int kGjFlowMaxM = 16;

__device__ __forceinline__ bool gjf_wait_geq(const int* flag, int target, int spin_limit) {
  int polls = 0;
  while (__hip_atomic_load(flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) < target) {
    if (++polls > spin_limit) return false;
    __builtin_amdgcn_s_sleep(1);
  }
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
  return true;
}

// Step k (global column kg = c0 + k) from this wave's column j.
template <int M, int R, int NS>
__device__ __forceinline__ void gjf_produce(cplx (&a)[M][R], int j, int k, int c0, int n,
                                            int lane, const unsigned long long (&used)[M],
                                            cplx* gs, int* pinfo, int* seq, int* cnt, int* spos,
                                            int* rowat, cplx* dv, int spin_limit,
                                            bool& singular, bool& stuck) {
  const int kg = c0 + k;
  const int slot = k % NS;
  const int freed = __hip_atomic_load(&cnt[slot], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
  // keys: the high word of the modulus (bit 31 marks an available row)
  unsigned k1[M];
#pragma unroll
  for (int s = 0; s < M; ++s) {
    const cplx f = a[s][j];
    const double m = fabs(f.re) + fabs(f.im);
    const unsigned long long b = (unsigned long long)__double_as_longlong(m);
    const bool ok = (((used[s] >> lane) & 1ull) == 0ull) & (m == m);
    k1[s] = ok ? ((unsigned)(b >> 32) | 0x80000000u) : 0u;
  }
  unsigned km = k1[0];
#pragma unroll
  for (int s = 1; s < M; ++s) km = max(km, k1[s]);
  const unsigned t1 = (unsigned)__builtin_amdgcn_readlane((int)wave_umax(km), 63);
  int p, pp;
  if (t1 == 0u) {  // no candidate: keep the row at position kg
    p = __builtin_amdgcn_readfirstlane(rowat[kg]);
    pp = kg;
  } else {
    unsigned long long e[M];
    int tot = 0;
#pragma unroll
    for (int s = 0; s < M; ++s) {
      e[s] = __ballot(k1[s] == t1);
      tot += __builtin_popcountll(e[s]);
    }
    if (tot > 1) {  // equal high words: the low words decide, then positions
      unsigned lo[M];
      unsigned k2m = 0u;
#pragma unroll
      for (int s = 0; s < M; ++s) {
        const double m = fabs(a[s][j].re) + fabs(a[s][j].im);
        lo[s] = (unsigned)(unsigned long long)__double_as_longlong(m);
        k2m = max(k2m, k1[s] == t1 ? lo[s] : 0u);
      }
      const unsigned t2 = (unsigned)__builtin_amdgcn_readlane((int)wave_umax(k2m), 63);
      tot = 0;
#pragma unroll
      for (int s = 0; s < M; ++s) {
        e[s] = __ballot((k1[s] == t1) & (lo[s] == t2));
        tot += __builtin_popcountll(e[s]);
      }
      if (tot > 1) {
        unsigned k3[M];
        unsigned k3m = 0u;
#pragma unroll
        for (int s = 0; s < M; ++s) {
          const bool tie = ((e[s] >> lane) & 1ull) != 0ull;
          const int row = lane + 64 * s;
          k3[s] = tie && row < n ? ~(unsigned)spos[row] : 0u;
          k3m = max(k3m, k3[s]);
        }
        const unsigned t3 = (unsigned)__builtin_amdgcn_readlane((int)wave_umax(k3m), 63);
#pragma unroll
        for (int s = 0; s < M; ++s) e[s] = __ballot((k3[s] == t3) & (((e[s] >> lane) & 1ull) != 0ull));
      }
    }
    int sp = 0, lp = 0;
#pragma unroll
    for (int s = M - 1; s >= 0; --s)
      if (e[s] != 0ull) {
        sp = s;
        lp = (int)__builtin_ctzll(e[s]);
      }
    p = lp + 64 * sp;
    pp = __builtin_amdgcn_readfirstlane(spos[p]);
  }
  // positions: the pivot row takes position kg, the row there the pivot's
  // old position (issued now; the reciprocal below covers their latency)
  if (lane == 0) {
    const int q = rowat[kg];
    spos[q] = pp;
    rowat[pp] = q;
    spos[p] = kg;
    rowat[kg] = p;
  }
  const int sp = p >> 6, lp = p & 63;
  cplx piv{0.0, 0.0};
#pragma unroll
  for (int s = 0; s < M; ++s)
    if (s == sp) piv = cplx{readlane_f64(a[s][j].re, lp), readlane_f64(a[s][j].im, lp)};
  if (t1 != 0u && piv.re == 0.0 && piv.im == 0.0) singular = true;  // the largest modulus is zero
  const cplx d = crecip_uniform(piv);
  if (freed < kGjWaves) {
    if (!gjf_wait_geq(&cnt[slot], kGjWaves, spin_limit)) {
      stuck = true;
      return;
    }
  } else {
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
  }
  cplx* gsl = gs + slot * (M * 64);
#pragma unroll
  for (int s = 0; s < M; ++s) {
    cplx g = cmul(a[s][j], d);
    if (s == sp && lane == lp) g = cplx{0.0, 0.0};
    gsl[s * 64 + lane] = g;
    a[s][j] = cplx{-g.re, -g.im};
    if (s == sp && lane == lp) a[s][j] = cplx{1.0, 0.0};
  }
  if (lane == 0) {
    __hip_atomic_store(&cnt[slot], 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    pinfo[slot] = p;
    dv[k] = d;
  }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
  if (lane == 0) __hip_atomic_store(&seq[slot], k + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

__device__ __forceinline__ cplx gjf_upd(cplx a, cplx g, cplx P) {
  return cplx{fma(-g.re, P.re, fma(g.im, P.im, a.re)), fma(-g.re, P.im, fma(-g.im, P.re, a.im))};
}

template <int M, int R>
__global__ __launch_bounds__(64 * kGjWaves) void gj_flow_kernel(
    cplx* A, int n, int lda, long long s_in, long long s_out, int inner, int c0, int wc,
    int* ipos, int* ipiv, unsigned long long* info, int tag, int spin_limit) {
  constexpr int NS = GjFlow<M>::NS;
  __shared__ cplx gs[NS * M * 64];  // multipliers of the published steps
  __shared__ int pinfo[NS];         // pivot row of each slot's step
  __shared__ int seq[NS];           // step published in the slot, plus one
  __shared__ int cnt[NS];           // waves done with the slot's step
  __shared__ int spos[64 * M];      // position of each row
  __shared__ int rowat[64 * M];     // row at each position
  __shared__ cplx dv[kGjWaves * R]; // 1 / pivot of each step
  const int z = blockIdx.x;
  const int zo = z / inner, zi = z - zo * inner;
  A += zi * s_in + zo * s_out;
  ipos += (long long)z * n;
  ipiv += (long long)z * n;
  const int lane = threadIdx.x & 63;
  const int v = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);

  for (int r = threadIdx.x; r < n; r += 64 * kGjWaves) {
    const int ps = c0 == 0 ? r : ipos[r];
    spos[r] = ps;
    rowat[ps] = r;
  }
  if ((int)threadIdx.x < NS) {
    seq[threadIdx.x] = 0;
    cnt[threadIdx.x] = kGjWaves;
  }
  cplx a[M][R];
#pragma unroll
  for (int s = 0; s < M; ++s) {
    const int row = lane + 64 * s;
#pragma unroll
    for (int j = 0; j < R; ++j) {
      const int c = v + kGjWaves * j;
      if (row < n && c < wc) {
        const double2 t = *reinterpret_cast<const double2*>(A + (long long)row * lda + c0 + c);
        a[s][j] = cplx{t.x, t.y};
      } else {
        a[s][j] = cplx{0.0, 0.0};
      }
    }
  }
  __syncthreads();
  unsigned long long used[M];  // rows outside the matrix or already pivots
#pragma unroll
  for (int s = 0; s < M; ++s) {
    const int row = lane + 64 * s;
    used[s] = __ballot(row >= n || spos[row] < c0);
  }

  bool singular = false, stuck = false;
  if (v == 0) {
    __builtin_amdgcn_s_setprio(3);
    gjf_produce<M, R, NS>(a, 0, 0, c0, n, lane, used, gs, pinfo, seq, cnt, spos, rowat, dv,
                          spin_limit, singular, stuck);
  }
#pragma unroll
  for (int j = 0; j < R; ++j) {
    if (kGjWaves * j >= wc || stuck) break;  // uniform
    for (int vv = 0; vv < kGjWaves; ++vv) {
      const int k = kGjWaves * j + vv;
      if (k >= wc || stuck) break;  // uniform
      const int slot = k % NS;
      const bool mine = v == vv;
      const bool last = vv == kGjWaves - 1;  // step k + 1 is column j + 1 of wave 0
      const bool prod = k + 1 < wc && v == (last ? 0 : vv + 1);
      if (prod)
        __builtin_amdgcn_s_setprio(3);
      else if (v == (last ? 1 : (vv + 2) % kGjWaves))
        __builtin_amdgcn_s_setprio(2);
      else
        __builtin_amdgcn_s_setprio(0);
      if (!gjf_wait_geq(&seq[slot], k + 1, spin_limit)) {
        stuck = true;
        break;
      }
      const int p = __builtin_amdgcn_readfirstlane(pinfo[slot]);
      const cplx* gsl = gs + slot * (M * 64);
      cplx g[M];
#pragma unroll
      for (int s = 0; s < M; ++s) g[s] = gsl[s * 64 + lane];
      const int sp = p >> 6, lp = p & 63;
#pragma unroll
      for (int s = 0; s < M; ++s)
        if (s == sp) used[s] |= 1ull << lp;
      cplx P[R];
      gj_row<M, R>(a, sp, lp, P);
      // column j (the owner's was replaced when the step was produced), then
      // column j + 1, each followed by the next step's production when it is
      // this wave's; then the other columns
      if (!mine) {
#pragma unroll
        for (int s = 0; s < M; ++s) a[s][j] = gjf_upd(a[s][j], g[s], P[j]);
      }
      if (prod && !last)
        gjf_produce<M, R, NS>(a, j, k + 1, c0, n, lane, used, gs, pinfo, seq, cnt, spos, rowat,
                              dv, spin_limit, singular, stuck);
      if (j + 1 < R) {
#pragma unroll
        for (int s = 0; s < M; ++s) a[s][j + 1] = gjf_upd(a[s][j + 1], g[s], P[j + 1]);
        if (prod && last)
          gjf_produce<M, R, NS>(a, j + 1, k + 1, c0, n, lane, used, gs, pinfo, seq, cnt, spos,
                                rowat, dv, spin_limit, singular, stuck);
      }
#pragma unroll
      for (int r = 0; r < R; ++r) {
        if (r == j || r == j + 1) continue;
#pragma unroll
        for (int s = 0; s < M; ++s) a[s][r] = gjf_upd(a[s][r], g[s], P[r]);
      }
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
      if (lane == 0)
        __hip_atomic_fetch_add(&cnt[slot], 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    }
  }
  __builtin_amdgcn_s_setprio(0);
  __syncthreads();

  // write-back: rows that were pivots of this launch take their 1 / pivot
#pragma unroll
  for (int s = 0; s < M; ++s) {
    const int row = lane + 64 * s;
    if (row < n) {
      const int q = spos[row] - c0;
      const bool scale = q >= 0 && q < wc;
      const cplx d = scale ? dv[q] : cplx{1.0, 0.0};
      cplx* dst = A + (long long)row * lda + c0 + v;
#pragma unroll
      for (int j = 0; j < R; ++j)
        if (v + kGjWaves * j < wc) {
          const cplx o = scale ? cmul(a[s][j], d) : a[s][j];
          *reinterpret_cast<double2*>(dst + kGjWaves * j) = double2{o.re, o.im};
        }
    }
  }
  // pivots from the row at each position: a permutation even when a wave
  // stopped early, so that every later launch indexes rows inside the matrix
  for (int r = threadIdx.x; r < n; r += 64 * kGjWaves) ipos[r] = spos[r];
  if ((int)threadIdx.x < wc) ipiv[c0 + threadIdx.x] = rowat[c0 + threadIdx.x];
  if (lane == 0 && (singular || stuck) && info != nullptr)
    atomicMin(info, ((unsigned long long)(unsigned)tag << 32) | (unsigned long long)(unsigned)z);
}

// Panels of blocks above 64 M = 2048 rows: the panel (16 columns) stays in
// global memory, where it is L2-resident, and is walked 64 rows at a time per
// step (16 lanes per row, lane j on column j; one workgroup per matrix).  The
// pivot search of the next column runs inside the current step's update pass
// on the values it writes.  Positions live in LDS (pos and its inverse), so n
// is limited by LDS to kMaxInverseN (64 KB of positions).
constexpr int kGjGlobalW = 16;

__device__ __forceinline__ double2 shfl_c(double2 v, int src) {
  return double2{__shfl(v.x, src), __shfl(v.y, src)};
}

// Wave winner of (modulus, position, row) candidates -> slot w of the block's
// table (bits 0 = no candidate).  All lanes active.
__device__ __forceinline__ void gj_wave_best(double mb, int pb, int rb, int lane, int w,
                                             unsigned long long* wbits, int* wpos, int* wrow) {
  const unsigned long long bits = (unsigned long long)__double_as_longlong(mb);
  const unsigned hi = (unsigned)(bits >> 32), lo = (unsigned)bits;
  const unsigned k1 = mb >= 0.0 ? (hi | 0x80000000u) : 0u;
  const unsigned t1 = (unsigned)__builtin_amdgcn_readlane((int)wave_umax(k1), 63);
  int winner = -1;
  if (t1 != 0u) {
    bool tie = k1 == t1;
    unsigned long long bal = __ballot(tie);
    if (__builtin_popcountll(bal) > 1) {
      const unsigned t2 = (unsigned)__builtin_amdgcn_readlane((int)wave_umax(tie ? lo : 0u), 63);
      tie = tie & (lo == t2);
      bal = __ballot(tie);
      if (__builtin_popcountll(bal) > 1) {
        const unsigned k3 = tie ? ~(unsigned)pb : 0u;
        const unsigned t3 = (unsigned)__builtin_amdgcn_readlane((int)wave_umax(k3), 63);
        bal = __ballot(tie & (k3 == t3));
      }
    }
    winner = (int)__builtin_ctzll(bal);
  }
  const unsigned long long wb =
      winner < 0 ? 0ull
                 : (1ull << 63) |
                       ((unsigned long long)(unsigned)__builtin_amdgcn_readlane((int)hi, winner)
                        << 32) |
                       (unsigned long long)(unsigned)__builtin_amdgcn_readlane((int)lo, winner);
  const int wp = winner < 0 ? 0x7fffffff : __builtin_amdgcn_readlane(pb, winner);
  const int wr = winner < 0 ? -1 : __builtin_amdgcn_readlane(rb, winner);
  if (lane == 0) {
    wbits[w] = wb;
    wpos[w] = wp;
    wrow[w] = wr;
  }
}

// Thread 0: the block's pivot of global column kg from the wave winners (the
// largest modulus, then the smallest position; none: the row at position kg),
// its 1/pivot and whether the pivot is zero, into LDS.
__device__ __forceinline__ void gj_pick(const cplx* A, int lda, int c0, int kg, int nw,
                                        const unsigned long long* wbits, const int* wpos,
                                        const int* wrow, const int* rowat, int* s_p, int* s_pp,
                                        cplx* s_d, int* s_zero) {
  unsigned long long bb = 0ull;
  int bp = 0x7fffffff, br = -1;
  for (int w = 0; w < nw; ++w)
    if (wbits[w] > bb || (wbits[w] == bb && bb != 0ull && wpos[w] < bp)) {
      bb = wbits[w];
      bp = wpos[w];
      br = wrow[w];
    }
  if (bb == 0ull) {
    br = rowat[kg];
    bp = kg;
  }
  const double2 v = *reinterpret_cast<const double2*>(A + (long long)br * lda + kg);
  *s_p = br;
  *s_pp = bp;
  *s_zero = (v.x == 0.0 && v.y == 0.0) ? 1 : 0;
  *s_d = crecip(cplx{v.x, v.y});
  (void)c0;
}

__global__ __launch_bounds__(1024) void gj_panel_global_kernel(
    cplx* A, int n, int lda, long long s_in, long long s_out, int inner, int c0, int wc,
    int* ipos, int* ipiv, unsigned long long* info, int tag) {
  extern __shared__ int gj_lds[];  // pos[n], then rowat[n] (the row at each position)
  int* spos = gj_lds;
  int* rowat = gj_lds + n;
  __shared__ cplx prow[kGjGlobalW];
  __shared__ unsigned long long wbits[16];
  __shared__ int wpos[16], wrow[16];
  __shared__ int s_p, s_pp, s_zero;
  __shared__ cplx s_d;
  const int z = blockIdx.x;
  const int zo = z / inner, zi = z - zo * inner;
  A += zi * s_in + zo * s_out;
  ipos += (long long)z * n;
  ipiv += (long long)z * n;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int j = tid & 15, grp = tid >> 4;
  const int src0 = lane & ~15;

  for (int r = tid; r < n; r += 1024) {
    const int p = c0 == 0 ? r : ipos[r];
    spos[r] = p;
    rowat[p] = r;
  }
  __syncthreads();
  {  // pivot of the first column
    double mb = -1.0;
    int pb = 0x7fffffff, rb = -1;
    if (j == 0)
      for (int r = grp; r < n; r += 64) {
        const double2 f = *reinterpret_cast<const double2*>(A + (long long)r * lda + c0);
        const double m = fabs(f.x) + fabs(f.y);
        const int ps = spos[r];
        const bool better = (ps >= c0) & (m == m) & ((m > mb) | ((m == mb) & (ps < pb)));
        mb = better ? m : mb;
        pb = better ? ps : pb;
        rb = better ? r : rb;
      }
    gj_wave_best(mb, pb, rb, lane, w, wbits, wpos, wrow);
    __syncthreads();
    if (tid == 0) gj_pick(A, lda, c0, c0, 16, wbits, wpos, wrow, rowat, &s_p, &s_pp, &s_d, &s_zero);
    __syncthreads();
  }

  bool singular = false;
  for (int k = 0; k < wc; ++k) {
    const int kg = c0 + k;
    const int p = s_p, pp = s_pp;
    const cplx d = s_d;
    singular |= s_zero != 0;
    if (tid < wc) prow[tid] = A[(long long)p * lda + c0 + tid];
    __syncthreads();
    const bool next = k + 1 < wc;
    double mb = -1.0;
    int pb = 0x7fffffff, rb = -1;
    const cplx P = j == k ? cplx{1.0, 0.0} : (j < wc ? prow[j] : cplx{0.0, 0.0});
    for (int r = grp; r < n; r += 64) {
      double2* ap = reinterpret_cast<double2*>(A + (long long)r * lda + c0 + j);
      const double2 av = j < wc ? *ap : double2{0.0, 0.0};
      const double2 f = shfl_c(av, src0 | k);
      const bool is_p = r == p;
      const cplx gf = cmul(cplx{f.x, f.y}, d);
      const cplx g{is_p ? -d.re : gf.re, is_p ? -d.im : gf.im};
      const double keep = is_p ? 0.0 : 1.0;
      const cplx a0 = j == k ? cplx{0.0, 0.0} : cplx{av.x, av.y};
      const cplx nv{keep * a0.re - (g.re * P.re - g.im * P.im),
                    keep * a0.im - (g.re * P.im + g.im * P.re)};
      if (j < wc) *ap = double2{nv.re, nv.im};
      if (next && j == k + 1) {  // candidate for the next column, with its new position
        const int ps = spos[r];
        const int np = ps == kg ? pp : (is_p ? kg : ps);
        const double m = fabs(nv.re) + fabs(nv.im);
        const bool better = (np > kg) & (m == m) & ((m > mb) | ((m == mb) & (np < pb)));
        mb = better ? m : mb;
        pb = better ? np : pb;
        rb = better ? r : rb;
      }
    }
    if (next) gj_wave_best(mb, pb, rb, lane, w, wbits, wpos, wrow);
    __syncthreads();
    if (tid == 0) {
      ipiv[kg] = p;
      const int q = rowat[kg];
      spos[q] = pp;
      rowat[pp] = q;
      spos[p] = kg;
      rowat[kg] = p;
      if (next) gj_pick(A, lda, c0, kg + 1, 16, wbits, wpos, wrow, rowat, &s_p, &s_pp, &s_d, &s_zero);
    }
    __syncthreads();
  }
  for (int r = tid; r < n; r += 1024) ipos[r] = spos[r];
  if (tid == 0 && singular && info != nullptr)
    atomicMin(info, ((unsigned long long)(unsigned)tag << 32) | (unsigned long long)(unsigned)z);
}

// B[z][t][c] = A[z][piv[k0 + t]][c0 + c], and that entry of A cleared
// (t < K, c < N; B compact K x N per matrix).
__global__ __launch_bounds__(256) void gj_extract_kernel(cplx* A, int lda, long long s_in,
                                                         long long s_out, int inner,
                                                         const int* ipiv, int n, int k0, int K,
                                                         int c0, int N, cplx* B) {
  const int z = blockIdx.y;
  const int zo = z / inner, zi = z - zo * inner;
  A += zi * s_in + zo * s_out;
  ipiv += (long long)z * n + k0;
  const long long total = (long long)K * N;
  double2* Bz = reinterpret_cast<double2*>(B + (long long)z * total);
  for (long long e = (long long)blockIdx.x * 256 + threadIdx.x; e < total;
       e += (long long)gridDim.x * 256) {
    const int t = (int)(e / N), c = (int)(e - (long long)t * N);
    const int row = min(max(ipiv[t], 0), n - 1);  // a pivot row; clamped as a guard
    double2* src = reinterpret_cast<double2*>(A + (long long)row * lda + c0 + c);
    Bz[e] = *src;
    *src = double2{0.0, 0.0};
  }
}

// W[z] = A[z] (compact n x n copy)
__global__ __launch_bounds__(256) void gj_copy_kernel(const cplx* A, int n, int lda,
                                                      long long s_in, long long s_out, int inner,
                                                      cplx* W) {
  const int z = blockIdx.y;
  const int zo = z / inner, zi = z - zo * inner;
  A += zi * s_in + zo * s_out;
  const long long total = (long long)n * n;
  double2* Wz = reinterpret_cast<double2*>(W + (long long)z * total);
  for (long long e = (long long)blockIdx.x * 256 + threadIdx.x; e < total;
       e += (long long)gridDim.x * 256) {
    const int r = (int)(e / n), c = (int)(e - (long long)r * n);
    Wz[e] = *reinterpret_cast<const double2*>(A + (long long)r * lda + c);
  }
}

// D[z][r][s] = S[z][piv[r]][pos[s]]: the inverse from the eliminated matrix
__global__ __launch_bounds__(256) void gj_permute_kernel(const cplx* S, int lds, long long ss_in,
                                                         long long ss_out, cplx* D, int ldd,
                                                         long long sd_in, long long sd_out, int n,
                                                         int inner, const int* ipiv,
                                                         const int* ipos) {
  const int r = blockIdx.x, z = blockIdx.y;
  const int zo = z / inner, zi = z - zo * inner;
  const int* pos = ipos + (long long)z * n;
  // pivots and positions are permutations of 0 .. n - 1; clamped as a guard
  const int row = min(max(ipiv[(long long)z * n + r], 0), n - 1);
  const double2* src =
      reinterpret_cast<const double2*>(S + zi * ss_in + zo * ss_out + (long long)row * lds);
  double2* dst = reinterpret_cast<double2*>(D + zi * sd_in + zo * sd_out + (long long)r * ldd);
  for (int c = threadIdx.x; c < n; c += 256) dst[c] = src[min(max(pos[c], 0), n - 1)];
}

// Tall-panel configuration for n rows: M rows per lane (the smallest
// instantiated M >= n / 64) and R columns per wave, so that the panel holds
// at most 16 complex values per lane up to 1024 rows (32 above); M = 0 selects
// the L2-resident kernel.
struct GjShape {
  int M, R;
};
// Panel kernel choice: the flow kernel (default, up to 1024 rows; the step
// kernel above), the per-column step kernel or the per-panel one
// (SINV_GJ_PANEL=flow / step / panel in the environment, or
// sinv_gj_force_variant); each has its own widest panels that fit its
// registers without spilling.
int g_gj_variant = -1;  // SINV_GJ_FLOW / SINV_GJ_STEP / SINV_GJ_PANEL; -1 until first read
int gj_variant() {
  if (g_gj_variant < 0) {
    const char* e = getenv("SINV_GJ_PANEL");
    int v = SINV_GJ_FLOW;
    if (e != nullptr && strcmp(e, "panel") == 0) v = SINV_GJ_PANEL;
    if (e != nullptr && strcmp(e, "step") == 0) v = SINV_GJ_STEP;
    g_gj_variant = v;
  }
  return g_gj_variant;
}

// The flow kernel holds at most 4 columns per wave from 6 row slots on
// (registers); at 321 .. 384 rows and 64 or more matrices per launch the
// step kernel's 64-column panels are taken instead, as the wider panels make
// the sweep products cheaper than the narrower panels' faster steps save.
int gj_kind(int n, long long nb) {
  const int kind = gj_variant();
  if (kind == SINV_GJ_FLOW && (n + 63) / 64 == 6 && nb >= 64) return SINV_GJ_STEP;
  return kind;
}

GjShape gj_shape(int n, long long nb) {
  static constexpr int kMs[] = {2, 3, 4, 5, 6, 7, 8, 10, 12, 14, 16, 20, 24, 28, 32};
  const int m = (n + 63) / 64;
  const int kind = gj_kind(n, nb);
  for (int M : kMs)
    if (M >= m) {
      if (kind == SINV_GJ_FLOW && M <= kGjFlowMaxM)
        return GjShape{M, M <= 5 ? 8 : (M <= 8 ? 4 : (M <= 12 ? 2 : 1))};
      if (kind != SINV_GJ_PANEL)
        return GjShape{M, M <= 6 ? 8 : (M <= 8 ? 4 : (M <= 14 ? 2 : 1))};
      return GjShape{M, M <= 2 ? 8 : (M <= 4 ? 4 : (M <= 8 ? 2 : 1))};
    }
  return GjShape{0, 0};
}
int gj_width(int n, long long nb) {
  const GjShape g = gj_shape(n, nb);
  return g.M > 0 ? kGjWaves * g.R : kGjGlobalW;
}

template <int M, int R, int KIND>
void gj_panel_mr(cplx* A, int n, int lda, long long s_in, long long s_out, int inner, unsigned nb,
                 int c0, int wc, int* ipos, int* ipiv, unsigned long long* info, int tag,
                 hipStream_t s) {
  if constexpr (KIND == SINV_GJ_FLOW && M <= kGjFlowMaxM)
    hipLaunchKernelGGL((gj_flow_kernel<M, R>), dim3(nb), dim3(64 * kGjWaves), 0, s, A, n, lda,
                       s_in, s_out, inner, c0, wc, ipos, ipiv, info, tag, g_leaf_spin_limit);
  else if constexpr (KIND != SINV_GJ_PANEL)
    hipLaunchKernelGGL((gj_step_kernel<M, R>), dim3(nb), dim3(64 * kGjWaves), 0, s, A, n, lda,
                       s_in, s_out, inner, c0, wc, ipos, ipiv, info, tag);
  else
    hipLaunchKernelGGL((gj_panel_kernel<M, R>), dim3(nb), dim3(64 * kGjWaves), 0, s, A, n, lda,
                       s_in, s_out, inner, c0, wc, ipos, ipiv, info, tag);
}

hipError_t launch_gj_panel(cplx* A, int n, int lda, long long s_in, long long s_out, int inner,
                           int outer, int c0, int wc, int* ipos, int* ipiv,
                           unsigned long long* info, int tag, hipStream_t s) {
  const unsigned nb = (unsigned)(inner * outer);
  const int pid = prof_open(s);
#define SINV_GJ_CASE(M_, R_, KIND_)                                                     \
  case M_ * 100 + R_:                                                                  \
    gj_panel_mr<M_, R_, KIND_>(A, n, lda, s_in, s_out, inner, nb, c0, wc, ipos, ipiv,   \
                               info, tag, s);                                          \
    break;
#define SINV_GJ_CASES(KIND_)      \
  SINV_GJ_CASE(2, 8, KIND_)       \
  SINV_GJ_CASE(3, 8, KIND_)       \
  SINV_GJ_CASE(4, 8, KIND_)       \
  SINV_GJ_CASE(5, 8, KIND_)       \
  SINV_GJ_CASE(6, 8, KIND_)       \
  SINV_GJ_CASE(7, 4, KIND_)       \
  SINV_GJ_CASE(8, 4, KIND_)       \
  SINV_GJ_CASE(10, 2, KIND_)      \
  SINV_GJ_CASE(12, 2, KIND_)      \
  SINV_GJ_CASE(14, 2, KIND_)      \
  SINV_GJ_CASE(16, 1, KIND_)      \
  SINV_GJ_CASE(20, 1, KIND_)      \
  SINV_GJ_CASE(24, 1, KIND_)      \
  SINV_GJ_CASE(28, 1, KIND_)      \
  SINV_GJ_CASE(32, 1, KIND_)
  const GjShape gs = gj_shape(n, nb);
  const int kind = gj_kind(n, nb);
  if (gs.M == 0) {
    hipLaunchKernelGGL(gj_panel_global_kernel, dim3(nb), dim3(1024), 2 * n * sizeof(int), s, A,
                       n, lda, s_in, s_out, inner, c0, wc, ipos, ipiv, info, tag);
  } else if (kind == SINV_GJ_FLOW) {
    switch (gs.M * 100 + gs.R) {
      SINV_GJ_CASE(2, 8, SINV_GJ_FLOW)
      SINV_GJ_CASE(3, 8, SINV_GJ_FLOW)
      SINV_GJ_CASE(4, 8, SINV_GJ_FLOW)
      SINV_GJ_CASE(5, 8, SINV_GJ_FLOW)
      SINV_GJ_CASE(6, 4, SINV_GJ_FLOW)
      SINV_GJ_CASE(7, 4, SINV_GJ_FLOW)
      SINV_GJ_CASE(8, 4, SINV_GJ_FLOW)
      SINV_GJ_CASE(10, 2, SINV_GJ_FLOW)
      SINV_GJ_CASE(12, 2, SINV_GJ_FLOW)
      SINV_GJ_CASE(14, 1, SINV_GJ_FLOW)
      SINV_GJ_CASE(16, 1, SINV_GJ_FLOW)
      SINV_GJ_CASE(20, 1, SINV_GJ_FLOW)
      SINV_GJ_CASE(24, 1, SINV_GJ_FLOW)
      SINV_GJ_CASE(28, 1, SINV_GJ_FLOW)
      SINV_GJ_CASE(32, 1, SINV_GJ_FLOW)
      default:
        return hipErrorInvalidValue;
    }
  } else if (kind == SINV_GJ_STEP) {
    switch (gs.M * 100 + gs.R) {
      SINV_GJ_CASES(SINV_GJ_STEP)
      default:
        return hipErrorInvalidValue;
    }
  } else {
    switch (gs.M * 100 + gs.R) {
      SINV_GJ_CASE(2, 8, SINV_GJ_PANEL)
      SINV_GJ_CASE(3, 4, SINV_GJ_PANEL)
      SINV_GJ_CASE(4, 4, SINV_GJ_PANEL)
      SINV_GJ_CASE(5, 2, SINV_GJ_PANEL)
      SINV_GJ_CASE(6, 2, SINV_GJ_PANEL)
      SINV_GJ_CASE(7, 2, SINV_GJ_PANEL)
      SINV_GJ_CASE(8, 2, SINV_GJ_PANEL)
      SINV_GJ_CASE(10, 1, SINV_GJ_PANEL)
      SINV_GJ_CASE(12, 1, SINV_GJ_PANEL)
      SINV_GJ_CASE(14, 1, SINV_GJ_PANEL)
      SINV_GJ_CASE(16, 1, SINV_GJ_PANEL)
      SINV_GJ_CASE(20, 1, SINV_GJ_PANEL)
      SINV_GJ_CASE(24, 1, SINV_GJ_PANEL)
      SINV_GJ_CASE(28, 1, SINV_GJ_PANEL)
      SINV_GJ_CASE(32, 1, SINV_GJ_PANEL)
      default:
        return hipErrorInvalidValue;
    }
  }
#undef SINV_GJ_CASES
#undef SINV_GJ_CASE
  prof_close(pid, s, PROF_INV_PANEL, 8.0 * n * (double)wc * wc * nb,
             2.0 * sizeof(cplx) * n * (double)wc * nb);
  return hipGetLastError();
}

// The elimination of columns [k0, k0 + K) applied to columns [cN, cN + N):
// B = A[piv, cN..] (rows cleared), then A[:, cN..] += A[:, k0..] B.
hipError_t gj_sweep(cplx* A, int n, int lda, long long s_in, long long s_out, int inner,
                    int outer, int k0, int K, int cN, int N, const int* ipiv, cplx* B,
                    hipStream_t s) {
  const long long nb = (long long)inner * outer;
  const long long total = (long long)K * N;
  const unsigned gx = (unsigned)std::min<long long>((total + 255) / 256, 64);
  hipLaunchKernelGGL(gj_extract_kernel, dim3(gx, (unsigned)nb), dim3(256), 0, s, A, lda, s_in,
                     s_out, inner, ipiv, n, k0, K, cN, N, B);
  SINV_HIP_RET(hipGetLastError());
  GemmArgs g{};
  g.M = n; g.N = N; g.K = K; g.opA = OP_N; g.opB = OP_N;
  g.A = A + k0; g.lda = lda; g.sA_in = s_in; g.sA_out = s_out;
  g.B = B; g.ldb = N; g.sB_in = total; g.sB_out = total * inner;
  g.C = A + cN; g.ldc = lda; g.sC_in = s_in; g.sC_out = s_out;
  g.C0 = A + cN; g.ldc0 = lda; g.sC0_in = s_in; g.sC0_out = s_out;
  g.alpha = cplx{1.0, 0.0};
  g.beta = cplx{1.0, 0.0};
  g.shift = nullptr;
  g.inner = inner; g.outer = outer;
  return launch_gemm(g, Prec::F64, s);
}

hipError_t gj_range(cplx* A, int n, int lda, long long s_in, long long s_out, int inner,
                    int outer, int c0, int c1, int wp, cplx* B, int* ipos, int* ipiv,
                    unsigned long long* info, int tag, hipStream_t s) {
  const int nc = c1 - c0;
  if (nc <= wp)
    return launch_gj_panel(A, n, lda, s_in, s_out, inner, outer, c0, nc, ipos, ipiv, info, tag,
                           s);
  const int n1 = std::max(wp, (nc / 2) / wp * wp);
  const int mid = c0 + n1, n2 = nc - n1;
  SINV_HIP_RET(gj_range(A, n, lda, s_in, s_out, inner, outer, c0, mid, wp, B, ipos, ipiv, info,
                        tag, s));
  SINV_HIP_RET(gj_sweep(A, n, lda, s_in, s_out, inner, outer, c0, n1, mid, n2, ipiv, B, s));
  SINV_HIP_RET(gj_range(A, n, lda, s_in, s_out, inner, outer, mid, c1, wp, B, ipos, ipiv, info,
                        tag, s));
  return gj_sweep(A, n, lda, s_in, s_out, inner, outer, mid, n2, c0, n1, ipiv, B, s);
}

}  // namespace

// elimination: the eliminated-matrix copy (also the sweeps' B operand), then
// positions and pivots
static size_t gj_workspace_elems(int n, long long nb) {
  return (size_t)n * n * (size_t)nb + ((size_t)n * (size_t)nb + 1) / 2 + 1;
}

size_t inverse_workspace_elems(int n, int batch, bool newton) {
  if (n <= LEAF) return 0;
  // the Newton-Schulz buffers follow the elimination's
  return gj_workspace_elems(n, batch) + (newton ? newton_workspace_elems(n) : 0);
}

static bool use_newton(bool newton, int n, long long nb) {
  return newton && n > LEAF && nb <= kNewtonMaxBatch;
}

namespace {

hipError_t launch_leaf(cplx* A, int n, int lda, long long s_in, long long s_out, int inner,
                       int outer, unsigned long long* info, int tag, hipStream_t s) {
  const int pid = prof_open(s);
  const dim3 grid((unsigned)(inner * outer));
  int v = leaf_variant();
  if (v == SINV_LEAF_AUTO)
    v = grid.x >= (unsigned)kLeafPanel8MinGrid ? SINV_LEAF_PANEL8_FLAGS : SINV_LEAF_PANEL16_FLAGS;
  switch (v) {
    case SINV_LEAF_LDS:
      hipLaunchKernelGGL(inv_leaf_kernel, grid, dim3(256), 0, s, A, n, lda, s_in, s_out, inner,
                         info, tag);
      break;
    case SINV_LEAF_ROWS8:
      hipLaunchKernelGGL(inv_leaf_kernel_rows<8>, grid, dim3(512), 0, s, A, n, lda, s_in, s_out,
                         inner, info, tag);
      break;
    case SINV_LEAF_PANEL8:
      hipLaunchKernelGGL(inv_leaf_kernel_panel<8>, grid, dim3(512), 0, s, A, n, lda, s_in, s_out,
                         inner, info, tag, g_leaf_spin_limit);
      break;
    case SINV_LEAF_PANEL16_FLAGS:
      hipLaunchKernelGGL((inv_leaf_kernel_panel<16, true>), grid, dim3(1024), 0, s, A, n, lda,
                         s_in, s_out, inner, info, tag, g_leaf_spin_limit);
      break;
    case SINV_LEAF_PANEL8_FLAGS:
      hipLaunchKernelGGL((inv_leaf_kernel_panel<8, true>), grid, dim3(512), 0, s, A, n, lda, s_in,
                         s_out, inner, info, tag, g_leaf_spin_limit);
      break;
    default:
      hipLaunchKernelGGL(inv_leaf_kernel_panel<16>, grid, dim3(1024), 0, s, A, n, lda, s_in,
                         s_out, inner, info, tag, g_leaf_spin_limit);
  }
  prof_close(pid, s, PROF_INV_LEAF, 8.0 * n * (double)n * n * inner * (double)outer,
             2.0 * sizeof(cplx) * n * (double)n * inner * (double)outer);
  return hipGetLastError();
}

// Blocked elimination of A in place (n > kLeaf); leaves positions and pivots
// in the workspace after the n^2-per-matrix copy region W.
hipError_t gj_eliminate(cplx* A, int n, int lda, long long s_in, long long s_out, int inner,
                        int outer, cplx* work, int** ipos, int** ipiv, unsigned long long* info,
                        int tag, hipStream_t s) {
  const long long nb = (long long)inner * outer;
  *ipos = reinterpret_cast<int*>(work + (long long)n * n * nb);
  *ipiv = *ipos + (long long)n * nb;
  return gj_range(A, n, lda, s_in, s_out, inner, outer, 0, n, gj_width(n, nb), work, *ipos, *ipiv,
                  info, tag, s);
}

}  // namespace

hipError_t launch_inverse(cplx* A, int n, int lda, long long s_in, long long s_out, int inner,
                          int outer, cplx* work, unsigned long long* info, int tag,
                          hipStream_t s, bool newton) {
  if (n <= 0 || inner <= 0 || outer <= 0) return hipSuccess;
  if (n <= LEAF) return launch_leaf(A, n, lda, s_in, s_out, inner, outer, info, tag, s);
  if (n > kMaxInverseN || work == nullptr) return hipErrorInvalidValue;
  const long long nb = (long long)inner * outer;
  if (use_newton(newton, n, nb))
    return launch_newton_inverse(A, n, lda, s_in, s_out, inner, outer, A, lda, s_in, s_out, true,
                                 work, gj_workspace_elems(n, nb), info, tag, s);
  int *ipos, *ipiv;
  SINV_HIP_RET(gj_eliminate(A, n, lda, s_in, s_out, inner, outer, work, &ipos, &ipiv, info, tag,
                            s));
  // the eliminated matrix is copied aside, then permuted back into A
  const long long total = (long long)n * n;
  const unsigned gx = (unsigned)std::min<long long>((total + 255) / 256, 256);
  const int pid = prof_open(s);
  hipLaunchKernelGGL(gj_copy_kernel, dim3(gx, (unsigned)nb), dim3(256), 0, s, A, n, lda, s_in,
                     s_out, inner, work);
  hipLaunchKernelGGL(gj_permute_kernel, dim3((unsigned)n, (unsigned)nb), dim3(256), 0, s, work, n,
                     total, total * inner, A, lda, s_in, s_out, n, inner, ipiv, ipos);
  prof_close(pid, s, PROF_OTHER, 0.0, 4.0 * sizeof(cplx) * total * nb);
  return hipGetLastError();
}

hipError_t launch_inverse_to(cplx* A, int n, int lda, long long s_in, long long s_out, int inner,
                             int outer, cplx* X, int ldx, long long sx_in, long long sx_out,
                             cplx* work, unsigned long long* info, int tag, hipStream_t s,
                             bool newton) {
  if (n <= 0 || inner <= 0 || outer <= 0) return hipSuccess;
  if (n <= LEAF) {
    SINV_HIP_RET(launch_axpby_shift(n, n, cplx{1.0, 0.0}, A, lda, s_in, s_out, X, ldx, sx_in,
                                    sx_out, nullptr, inner, outer, s));
    return launch_leaf(X, n, ldx, sx_in, sx_out, inner, outer, info, tag, s);
  }
  if (n > kMaxInverseN || work == nullptr) return hipErrorInvalidValue;
  const long long nb = (long long)inner * outer;
  if (use_newton(newton, n, nb))
    return launch_newton_inverse(A, n, lda, s_in, s_out, inner, outer, X, ldx, sx_in, sx_out,
                                 false, work, gj_workspace_elems(n, nb), info, tag, s);
  int *ipos, *ipiv;
  SINV_HIP_RET(gj_eliminate(A, n, lda, s_in, s_out, inner, outer, work, &ipos, &ipiv, info, tag,
                            s));
  const int pid = prof_open(s);
  hipLaunchKernelGGL(gj_permute_kernel, dim3((unsigned)n, (unsigned)nb), dim3(256), 0, s, A, lda,
                     s_in, s_out, X, ldx, sx_in, sx_out, n, inner, ipiv, ipos);
  prof_close(pid, s, PROF_OTHER, 0.0, 2.0 * sizeof(cplx) * (double)n * n * nb);
  return hipGetLastError();
}

}  // namespace sinv

extern "C" size_t sinv_zinv_workspace_bytes(int n, int batch) {
  if (n <= 0 || batch <= 0) return 0;
  return sinv::inverse_workspace_elems(n, batch,
                                       sinv::inverse_method() == SINV_INV_NEWTON) *
         sizeof(sinv::cplx);
}

extern "C" int sinv_zinv_batched(void* A, int n, int lda, long long strideA, int batch, void* work,
                                 size_t work_bytes, unsigned long long* info, void* stream) {
  using namespace sinv;
  if (A == nullptr || n < 0 || batch < 0 || lda < n) {
    set_error("sinv_zinv_batched: invalid argument");
    return SINV_E_INVAL;
  }
  if (n > kMaxInverseN) {
    set_error("sinv_zinv_batched: n = %d exceeds the largest supported block (%d)", n,
              kMaxInverseN);
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
  SINV_HIP_TRY(launch_inverse((cplx*)A, n, lda, strideA, 0, batch, 1, (cplx*)work, info, 0, s,
                              inverse_method() == SINV_INV_NEWTON));
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

extern "C" int sinv_gj_force_variant(int variant) {
  using namespace sinv;
  if (variant != SINV_GJ_STEP && variant != SINV_GJ_PANEL && variant != SINV_GJ_FLOW) {
    set_error("sinv_gj_force_variant: unknown variant %d", variant);
    return SINV_E_INVAL;
  }
  g_gj_variant = variant;
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
