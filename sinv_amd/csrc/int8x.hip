// FP64-accurate complex GEMM on the int8 matrix cores (sliced-integer
// emulation in the manner of Ozaki, Ogita and Oishi's error-free splitting).
//
// C[z] = alpha * op(A[z]) op(B[z]) + beta * C0[z] + shift[z] * I, complex128
// in and out, the same contract as launch_gemm.
//
// Method.  Each row r of op(A) (each column of op(B)) and each real part
// p in {re, im, re + im} gets a power-of-two scale 2^e with max |x| / 2^e <=
// 127/128 over the row (the frexp exponent of the maximum, one higher when its
// mantissa exceeds 127/128), and x becomes the integer
// N = rint(x 2^(8S-1-e)), |N| <= 127 * 2^(8S-8), written with S balanced
// base-256 digits: the low S - 1 in [-128, 127] and a top digit in
// [-127, 127], x = 2^e (sum_t q_t 2^(-7-8t) + rem), |rem| <= 2^(-8S).  The
// digits are the bytes of m = N + B_S, B_S = 128 (1 + 256 + ... +
// 256^(S-2)): biasing each low digit by 128 makes it the byte q_t + 128 of m
// (so the digit's own byte is that byte XOR 0x80), and the top byte of m is
// the top digit.  The digit products q_t(A) q_u(B) summed over k are exact in
// int32 on v_mfma_i32_16x16x64_i8 (|sum| <= S Kp 2^14 < 2^31); pairs with the
// same t + u share one accumulator, and pairs with t + u > S - 1 are dropped.
// After the k loop each accumulator is converted to double and scaled by
// 2^(-14-8(t+u)) 2^(eA_r + eB_c).  Per term the dropped pairs (S - 1 of size
// 2^(-8S) at most, then geometrically smaller) and the two remainders beyond S
// digits give an error below 4.1 (S + 4) 2^(-8S) max_k|a_rk| max_k|b_kc|
// (scales are under 2 (128/127) times the maxima), so one real dot product
// errs by at most that times K (S = 6: 2^-42.6 K, typically far less as the
// dropped digits are not aligned); an FP64 dot product has 2^-53 K sum|a||b|
// as its bound.  The complex
// product uses the 3M form as the FP64 kernel does: T1 = Ar Br, T2 = Ai Bi,
// T3 = (Ar + Ai)(Br + Bi), C = (T1 - T2) + i (T3 - T1 - T2).
//
// Data flow per launch: per operand, a split writes the digits of op(A) and
// op(B) (int8, blocked by 64-wide k chunks, k and rows padded to multiples of
// 64 with zeros) and their scales into the caller's workspace; an operand with
// a zero batch stride is split once.  For K <= 512 one kernel reads each row
// once (its maximum and its digits from the same registers); for longer rows a
// row-maximum pass over 64 x 64 pieces precedes the digit pass.  The product
// kernel stages 64 x 64 digit tiles of all S digits of op(A) and op(B) through
// LDS, runs the three real products one after another over k, and folds each
// into double accumulators.  Two versions: one LDS buffer with two workgroups
// per CU (the default), and two LDS buffers with one workgroup per CU
// (selectable; identical bits).
//
// int8 MFMA 16x16x64 operands: lane l supplies 16 consecutive k of row l % 16
// (A) or column l % 16 (B) starting at k = 16 (l / 16).  Any assignment of k to
// (lane group, byte) gives the same sum as long as A and B use the same one.
// Accumulator item v of lane l holds D[4 (l / 16) + v][l % 16].
#include <cstdint>
#include <initializer_list>

#include "sinv_core.hpp"

namespace sinv {

namespace {

typedef int i4v __attribute__((ext_vector_type(4)));

constexpr int kTile = 64;   // workgroup tile (rows and columns)
constexpr int kChunk = 64;  // k per LDS chunk (one MFMA's k)
constexpr int kRowB = 80;   // LDS row stride in bytes (64 + 16 pad: conflict-free b128 reads)
constexpr int kMaxKp = 16384;  // S * Kp * 2^14 must stay below 2^31 (S <= 7)
constexpr int kSplitMaxChunks = 8;  // single-pass split up to K = 512

inline long long round_up(long long x, long long m) { return (x + m - 1) / m * m; }

// Which operand copies are distinct: bit 1 = inner stride non-zero, bit 0 =
// outer stride non-zero.  The slice index zi of batch entry (z_in, z_out).
__host__ __device__ inline int zmode(long long s_in, long long s_out) {
  return (s_in != 0 ? 2 : 0) | (s_out != 0 ? 1 : 0);
}
__host__ __device__ inline long long zcount(int mode, int inner, int outer) {
  return mode == 0 ? 1 : mode == 1 ? outer : mode == 2 ? inner : (long long)inner * outer;
}
__host__ __device__ inline long long zindex(int mode, int z_in, int z_out, int inner) {
  return mode == 0 ? 0 : mode == 1 ? z_out : mode == 2 ? z_in : (long long)z_out * inner + z_in;
}

// op(X) seen as R "rows" of K elements: element (r, k) at X + r*rs + k*ks,
// imaginary part negated when conj.  Digit t of part p of element (r, k) of
// distinct copy zi is stored blocked by 64-wide k chunks, so that one chunk's
// 64 x 64 digit tile is 4 KB of consecutive bytes:
//   data + zi*3*S*Rp*Kp + (((p*nkc + k/64)*S + t)*Rp + r)*64 + pos(k%64)
// (nkc = Kp/64), where within a chunk k = 4 j + q (q < 4, j < 16) sits at byte
// pos = 16 q + j: the split's threads hold k interleaved by 4 (so that a load
// instruction reads runs of each row) and write 16 consecutive bytes.  Both
// operands use the same order, and a dot product does not depend on the order
// of its terms.  Scales at scale + (zi*3 + p)*Rp + r (2^e, 0 for an all-zero
// row).
struct SliceArgs {
  const cplx* X;
  long long s_in, s_out;
  int inner, mode;
  int R, K, Rp, Kp;
  long long rs, ks;
  double conj;
  int8_t* data;
  double* scale;
};

__device__ __forceinline__ void parts_of(const cplx& x, double conj, double v[3]) {
  const double im = conj * x.im;
  v[0] = x.re;
  v[1] = im;
  v[2] = x.re + im;
}

__device__ __forceinline__ const cplx* slice_src(const SliceArgs& a, long long zi) {
  long long z_in = 0, z_out = 0;
  if (a.mode == 1) z_out = zi;
  else if (a.mode == 2) z_in = zi;
  else if (a.mode == 3) { z_out = zi / a.inner; z_in = zi - z_out * a.inner; }
  return a.X + z_in * a.s_in + z_out * a.s_out;
}

// scale of a row from its maximum magnitude: 2^e with max / 2^e <= 127/128
// (frexp gives max = f 2^e, f in [0.5, 1); e + 1 when f > 127/128, so the top
// digit stays in int8).  Returns the exponent shift 8S - 1 - e that makes the
// row's integers; an all-zero row gets scale 0 and shift 0 (its digits are 0).
template <int S>
__device__ __forceinline__ int row_shift(double mx, double& sc) {
  if (!(mx > 0.0)) {
    sc = 0.0;
    return 0;
  }
  int e;
  const double f = frexp(mx, &e);
  if (f > 127.0 / 128.0) ++e;
  sc = ldexp(1.0, e);
  return 8 * S - 1 - e;
}

// m = rint(x 2^sh) + B_S (the digits as bytes, see the method above).  For
// S <= 6, |rint(x 2^sh)| < 2^47 and one addition does it: x 2^sh + (1.5 2^52 +
// B_S) rounds to the nearest integer in [2^52, 2^53) (ties to even; B_S is
// even), whose 52 significand bits are 2^51 + m, so the low 48 bits of the
// double's bit pattern are m mod 2^48.  S = 7 (|N| < 2^55) goes through int64.
template <int S>
__device__ __forceinline__ unsigned long long biased_integer(double x, int sh) {
  static_assert(S >= 2 && S <= 7, "the biased integer must fit 56 bits");
  constexpr unsigned long long bias = 0x8080808080808080ull >> (8 * (9 - S));
  if constexpr (S <= 6) {
    constexpr double magic = 6755399441055744.0 + (double)bias;  // 1.5 * 2^52 + B_S
    return (unsigned long long)__double_as_longlong(__builtin_ldexp(x, sh) + magic);
  } else {
    return (unsigned long long)((long long)rint(__builtin_ldexp(x, sh)) + (long long)bias);
  }
}

// Digit planes of four consecutive elements: word t holds digit t (t = 0 the
// top digit) of elements 0..3 in bytes 0..3; digit t is byte S - 1 - t of m,
// XOR 0x80 for the biased low digits.  A 4 x 4 byte transpose per 32-bit half
// of m with v_perm_b32 (perm(s0, s1, sel): result byte i is byte sel_i of the
// eight bytes s1[0..3], s0[0..3]).
__device__ __forceinline__ void transpose4(unsigned a, unsigned b, unsigned c, unsigned d,
                                           unsigned out[4]) {
  const unsigned ab01 = __builtin_amdgcn_perm(b, a, 0x05010400u);  // a0 b0 a1 b1
  const unsigned ab23 = __builtin_amdgcn_perm(b, a, 0x07030602u);  // a2 b2 a3 b3
  const unsigned cd01 = __builtin_amdgcn_perm(d, c, 0x05010400u);
  const unsigned cd23 = __builtin_amdgcn_perm(d, c, 0x07030602u);
  out[0] = __builtin_amdgcn_perm(cd01, ab01, 0x05040100u);  // a0 b0 c0 d0
  out[1] = __builtin_amdgcn_perm(cd01, ab01, 0x07060302u);  // a1 b1 c1 d1
  out[2] = __builtin_amdgcn_perm(cd23, ab23, 0x05040100u);
  out[3] = __builtin_amdgcn_perm(cd23, ab23, 0x07060302u);
}

template <int S>
__device__ __forceinline__ void plane_words(const unsigned long long m[4], unsigned w[S]) {
  unsigned lo[4], hi[4];
  transpose4((unsigned)m[0], (unsigned)m[1], (unsigned)m[2], (unsigned)m[3], lo);
  transpose4((unsigned)(m[0] >> 32), (unsigned)(m[1] >> 32), (unsigned)(m[2] >> 32),
             (unsigned)(m[3] >> 32), hi);
#pragma unroll
  for (int t = 0; t < S; ++t) {
    const int byte = S - 1 - t;
    const unsigned v = byte < 4 ? lo[byte] : hi[byte - 4];
    w[t] = t > 0 ? v ^ 0x80808080u : v;
  }
}

// byte offset of digit t, part p, row r, k chunk kc (k%64 added by the caller)
__host__ __device__ __forceinline__ size_t dig_off(int S, int nkc, int Rp, int p, int kc, int t,
                                                   int r) {
  return ((((size_t)p * nkc + kc) * S + t) * Rp + r) * 64;
}

// The digits of a thread's 16 elements k = 64 kc + 4 j + kq (j < 16) of row
// r, with exponent shifts sh[p]: 16 consecutive bytes per plane at byte
// 16 kq of the chunk's row (the in-chunk order of the layout above).
template <int S>
__device__ __forceinline__ void write_digits(const SliceArgs& a, long long zi, int nkc, int kc,
                                             int kq, int r, const cplx x[16], const int sh[3]) {
  int8_t* base = a.data + (size_t)zi * 3 * S * a.Rp * a.Kp + kq * 16;
#pragma unroll
  for (int p = 0; p < 3; ++p) {
    unsigned words[4][S];
#pragma unroll
    for (int j4 = 0; j4 < 4; ++j4) {
      unsigned long long m[4];
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        double v[3];
        parts_of(x[j4 * 4 + j], a.conj, v);
        m[j] = biased_integer<S>(v[p], sh[p]);
      }
      plane_words<S>(m, words[j4]);
    }
#pragma unroll
    for (int t = 0; t < S; ++t) {
      i4v v4;
      v4.x = (int)words[0][t];
      v4.y = (int)words[1][t];
      v4.z = (int)words[2][t];
      v4.w = (int)words[3][t];
      *reinterpret_cast<i4v*>(base + dig_off(S, nkc, a.Rp, p, kc, t, r)) = v4;
    }
  }
}

__device__ __forceinline__ void row_max16(const cplx x[16], double conj, double mx[3]) {
#pragma unroll
  for (int j = 0; j < 16; ++j) {
    double v[3];
    parts_of(x[j], conj, v);
#pragma unroll
    for (int p = 0; p < 3; ++p) mx[p] = fmax(mx[p], fabs(v[p]));
  }
}

// Single pass for rows of at most 8 chunks (K <= 512): grid (Rp / 16,
// distinct copies), 64 nkc threads; wave kc takes chunk kc of 16 rows, thread
// (rl, kq) the 16 elements k = 64 kc + 4 j + kq of row 16 blockIdx.x + rl,
// held in registers from the maximum to the digits:
//   KC = true (k-contiguous storage, ks == 1): rl = lane / 4, kq = lane % 4,
//     so each load instruction reads 64 consecutive bytes of 16 rows;
//   KC = false (row-contiguous storage, rs == 1): rl = lane % 16,
//     kq = lane / 16: 256 consecutive bytes (16 rows) at each of 4 k.
// A row's maximum: over the thread's 16 elements, the row's 4 threads (lane
// exchanges), then the waves (LDS).  A wave's digit stores cover 16 whole
// 64-byte rows (1 KB) per plane.
template <int S, bool KC>
__global__ __launch_bounds__(64 * kSplitMaxChunks) void split_rows_kernel(SliceArgs a, int nkc) {
  __shared__ double part[kSplitMaxChunks][3][16];
  __shared__ int shift[3][16];
  const int lane = threadIdx.x & 63, kc = threadIdx.x >> 6;
  const int rl = KC ? lane >> 2 : lane & 15;
  const int kq = KC ? lane & 3 : lane >> 4;
  const int r = blockIdx.x * 16 + rl;
  const long long zi = blockIdx.y;
  const cplx* X = slice_src(a, zi);
  const bool live = r < a.R;
  cplx x[16];
#pragma unroll
  for (int j = 0; j < 16; ++j) {
    const int k = kc * 64 + 4 * j + kq;
    x[j] = (live && k < a.K) ? X[(long long)r * a.rs + (long long)k * a.ks] : cplx{0.0, 0.0};
  }
  double mx[3] = {0.0, 0.0, 0.0};
  row_max16(x, a.conj, mx);
#pragma unroll
  for (int p = 0; p < 3; ++p) {
    mx[p] = fmax(mx[p], __shfl_xor(mx[p], KC ? 1 : 16));
    mx[p] = fmax(mx[p], __shfl_xor(mx[p], KC ? 2 : 32));
  }
  if (kq == 0) {
#pragma unroll
    for (int p = 0; p < 3; ++p) part[kc][p][rl] = mx[p];
  }
  __syncthreads();
  if (threadIdx.x < 48) {
    const int p = threadIdx.x >> 4, row = threadIdx.x & 15;
    double m = 0.0;
    for (int w = 0; w < nkc; ++w) m = fmax(m, part[w][p][row]);
    double sc;
    shift[p][row] = row_shift<S>(m, sc);
    a.scale[(zi * 3 + p) * a.Rp + blockIdx.x * 16 + row] = sc;
  }
  __syncthreads();
  const int sh[3] = {shift[0][rl], shift[1][rl], shift[2][rl]};
  write_digits<S>(a, zi, nkc, kc, kq, r, x, sh);
}

// Two passes for longer rows, over 64 x 64 pieces of op(X) (rows
// 64 blockIdx.x + [0, 64), k in 64 blockIdx.y + [0, 64)), grid
// (Rp / 64, Kp / 64, distinct copies), 256 threads; thread (rl, kq) takes row
// 64 blockIdx.x + rl and the 16 elements k = 64 kc + 4 j + kq:
//   KC = false (row-contiguous storage, rs == 1): rl = lane, kq = wave, so a
//     wave's loads are runs along the rows;
//   KC = true (k-contiguous storage, ks == 1): rl = tid / 4, kq = tid % 4, so
//     each load instruction reads 64 consecutive bytes of each of 64 rows.
//  1. piece_max: the piece's row maxima per part -> pmax[(zi*nkc + kc)*3 + p][r]
//  2. slice_piece: each row's scale from its nkc partial maxima, then the
//     piece's digits.
__device__ __forceinline__ double* pmax_at(double* pmax, long long zi, int nkc, int kc, int p,
                                           int Rp) {
  return pmax + ((zi * nkc + kc) * 3 + p) * (long long)Rp;
}

template <bool KC>
__device__ __forceinline__ void piece_coords(int& rl, int& kq) {
  const int tid = threadIdx.x;
  rl = KC ? tid >> 2 : tid & 63;
  kq = KC ? tid & 3 : tid >> 6;
}

template <bool KC>
__device__ __forceinline__ void load_piece(const SliceArgs& a, const cplx* X, int r, int kc,
                                           int kq, cplx x[16]) {
  const bool live = r < a.R;
#pragma unroll
  for (int j = 0; j < 16; ++j) {
    const int k = kc * 64 + 4 * j + kq;
    x[j] = (live && k < a.K) ? X[(long long)r * a.rs + (long long)k * a.ks] : cplx{0.0, 0.0};
  }
}

template <bool KC>
__global__ __launch_bounds__(256) void piece_max_kernel(SliceArgs a, double* pmax, int nkc) {
  __shared__ double part[4][3][64];
  int rl, kq;
  piece_coords<KC>(rl, kq);
  const int r = blockIdx.x * 64 + rl;
  const int kc = blockIdx.y;
  const long long zi = blockIdx.z;
  cplx x[16];
  load_piece<KC>(a, slice_src(a, zi), r, kc, kq, x);
  double mx[3] = {0.0, 0.0, 0.0};
  row_max16(x, a.conj, mx);
#pragma unroll
  for (int p = 0; p < 3; ++p) part[kq][p][rl] = mx[p];
  __syncthreads();
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (w < 3) {
    const double m = fmax(fmax(part[0][w][lane], part[1][w][lane]),
                          fmax(part[2][w][lane], part[3][w][lane]));
    pmax_at(pmax, zi, nkc, kc, w, a.Rp)[blockIdx.x * 64 + lane] = m;
  }
}

template <int S, bool KC>
__global__ __launch_bounds__(256) void slice_piece_kernel(SliceArgs a, const double* pmax,
                                                          int nkc) {
  int rl, kq;
  piece_coords<KC>(rl, kq);
  const int r = blockIdx.x * 64 + rl;
  const int kc = blockIdx.y;
  const long long zi = blockIdx.z;
  int sh[3];
#pragma unroll
  for (int p = 0; p < 3; ++p) {
    double m = 0.0;
    for (int c = 0; c < nkc; ++c)
      m = fmax(m, pmax_at(const_cast<double*>(pmax), zi, nkc, c, p, a.Rp)[r]);
    double sc;
    sh[p] = row_shift<S>(m, sc);
    if (kc == 0 && kq == 0) a.scale[(zi * 3 + p) * a.Rp + r] = sc;
  }
  cplx x[16];
  load_piece<KC>(a, slice_src(a, zi), r, kc, kq, x);
  write_digits<S>(a, zi, nkc, kc, kq, r, x, sh);
}

struct ProdArgs {
  const int8_t* a_data;
  const double* a_scale;
  int a_mode, a_Rp;
  const int8_t* b_data;
  const double* b_scale;
  int b_mode, b_Rp;
  int Kp;
};

// product kernel (two LDS buffers, one workgroup per CU): 256 threads = 4
// waves (2 x 2), workgroup tile 64 x 64, wave tile 32 x 32 = 2 x 2 MFMA tiles;
// S digit planes per operand; the next chunk's global loads are issued before
// the current chunk's MFMAs and stored to the other buffer after them.
template <int S>
__global__ __launch_bounds__(256) void gemm_i8x_kernel(GemmArgs g, ProdArgs o, int tiles_m,
                                                       int tiles_n) {
  __shared__ __attribute__((aligned(16))) int8_t lds[2][2][S][kTile][kRowB];

  // XCD-aware order: workgroup bid runs on XCD bid % 8; give each XCD a
  // contiguous range of tiles (neighbouring batch entries share its L2)
  int bid = blockIdx.x;
  const int total = gridDim.x;
  if ((total & 7) == 0) bid = (bid & 7) * (total >> 3) + (bid >> 3);
  const int tiles = tiles_m * tiles_n;
  const int z = bid / tiles;
  const int t = bid - z * tiles;
  const int z_out = z / g.inner, z_in = z - z_out * g.inner;
  const int m0 = (t / tiles_n) * kTile, n0 = (t % tiles_n) * kTile;

  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave >> 1, wn = wave & 1;
  const int Kp = o.Kp, nk = Kp / kChunk, nc = 3 * nk;

  const long long za = zindex(o.a_mode, z_in, z_out, g.inner);
  const long long zb = zindex(o.b_mode, z_in, z_out, g.inner);
  // this thread's 16-byte piece of every digit tile: row tid / 4, bytes 16 (tid % 4)
  const int lr = tid >> 2, lc = (tid & 3) * 16;
  const int8_t* a_src = o.a_data + (size_t)za * 3 * S * o.a_Rp * Kp + (size_t)(m0 + lr) * 64 + lc;
  const int8_t* b_src = o.b_data + (size_t)zb * 3 * S * o.b_Rp * Kp + (size_t)(n0 + lr) * 64 + lc;

  i4v ra[S], rb[S];
  auto load = [&](int c) {
    const int p = c / nk, kc = c - p * nk;
#pragma unroll
    for (int d = 0; d < S; ++d) {
      ra[d] = *reinterpret_cast<const i4v*>(a_src + dig_off(S, nk, o.a_Rp, p, kc, d, 0));
      rb[d] = *reinterpret_cast<const i4v*>(b_src + dig_off(S, nk, o.b_Rp, p, kc, d, 0));
    }
  };
  auto store = [&](int buf) {
#pragma unroll
    for (int d = 0; d < S; ++d) {
      *reinterpret_cast<i4v*>(&lds[buf][0][d][lr][lc]) = ra[d];
      *reinterpret_cast<i4v*>(&lds[buf][1][d][lr][lc]) = rb[d];
    }
  };

  i4v acc[2][2][S];
  double cr[2][2][4], ci[2][2][4];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) {
#pragma unroll
      for (int d = 0; d < S; ++d) acc[i][j][d] = i4v{0, 0, 0, 0};
#pragma unroll
      for (int v = 0; v < 4; ++v) cr[i][j][v] = ci[i][j][v] = 0.0;
    }

  const int fr = lane & 15, fk = (lane >> 4) * 16;
  load(0);
  store(0);
  __syncthreads();
  // the three real products one after another; accumulators live only
  // inside one product's k loop
#pragma unroll 1
  for (int p = 0; p < 3; ++p) {
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j)
#pragma unroll
        for (int d = 0; d < S; ++d) acc[i][j][d] = i4v{0, 0, 0, 0};
#pragma unroll 1
    for (int kc = 0; kc < nk; ++kc) {
      const int c = p * nk + kc;
      const int buf = c & 1;
      if (c + 1 < nc) load(c + 1);
      i4v bf[2][S];
#pragma unroll
      for (int ct = 0; ct < 2; ++ct)
#pragma unroll
        for (int u = 0; u < S; ++u)
          bf[ct][u] =
              *reinterpret_cast<const i4v*>(&lds[buf][1][u][wn * 32 + ct * 16 + fr][fk]);
#pragma unroll
      for (int rt = 0; rt < 2; ++rt)
#pragma unroll
        for (int ta = 0; ta < S; ++ta) {
          // one digit plane of op(A) at a time, against every B digit it pairs with
          const i4v af =
              *reinterpret_cast<const i4v*>(&lds[buf][0][ta][wm * 32 + rt * 16 + fr][fk]);
#pragma unroll
          for (int ub = 0; ta + ub < S; ++ub)
#pragma unroll
            for (int ct = 0; ct < 2; ++ct)
              acc[rt][ct][ta + ub] = __builtin_amdgcn_mfma_i32_16x16x64_i8(
                  af, bf[ct][ub], acc[rt][ct][ta + ub], 0, 0, 0);
        }
      if (c + 1 < nc) store(buf ^ 1);
      __syncthreads();
    }
    // fold real product p into the complex accumulators (3M combination)
    const double* sa = o.a_scale + ((size_t)za * 3 + p) * o.a_Rp + m0 + wm * 32;
    const double* sb = o.b_scale + ((size_t)zb * 3 + p) * o.b_Rp + n0 + wn * 32;
    const double sr = p == 0 ? 1.0 : -1.0;  // sign of T_p in the real part (T3: 0)
    const double si = p == 2 ? 1.0 : -1.0;  // sign of T_p in the imaginary part
#pragma unroll
    for (int rt = 0; rt < 2; ++rt)
#pragma unroll
      for (int ct = 0; ct < 2; ++ct) {
        const double scol = sb[ct * 16 + fr];
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          double val = 0.0;
#pragma unroll
          for (int d = S - 1; d >= 0; --d)
            val += (double)acc[rt][ct][d][v] * __builtin_ldexp(1.0, -14 - 8 * d);
          val *= sa[rt * 16 + (lane >> 4) * 4 + v] * scol;
          if (p < 2) cr[rt][ct][v] += sr * val;
          ci[rt][ct][v] += si * val;
        }
      }
  }

  // epilogue: C = alpha T + beta C0 + shift I, in bounds only
  cplx* C = g.C + z_in * g.sC_in + z_out * g.sC_out;
  const cplx* C0 = g.C0 ? g.C0 + z_in * g.sC0_in + z_out * g.sC0_out : nullptr;
  const cplx sh = g.shift ? g.shift[z_out] : cplx{0.0, 0.0};
#pragma unroll
  for (int rt = 0; rt < 2; ++rt)
#pragma unroll
    for (int ct = 0; ct < 2; ++ct)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int row = m0 + wm * 32 + rt * 16 + (lane >> 4) * 4 + v;
        const int col = n0 + wn * 32 + ct * 16 + fr;
        if (row < g.M && col < g.N) {
          const double tr = cr[rt][ct][v], ti = ci[rt][ct][v];
          double re = g.alpha.re * tr - g.alpha.im * ti;
          double im = g.alpha.re * ti + g.alpha.im * tr;
          if (C0) {
            const cplx c0 = C0[(long long)row * g.ldc0 + col];
            re += g.beta.re * c0.re - g.beta.im * c0.im;
            im += g.beta.re * c0.im + g.beta.im * c0.re;
          }
          if (row == col) {
            re += sh.re;
            im += sh.im;
          }
          C[(long long)row * g.ldc + col] = cplx{re, im};
        }
      }
}

// Same product with a single LDS buffer (S x 64 rows per operand, so two
// workgroups share a CU) and registers capped for two waves per SIMD: each
// chunk is loaded, stored to LDS between two barriers and then multiplied;
// the other workgroup's MFMAs cover this one's load latency.  The default
// (variant 2) swizzles its 64-byte LDS rows, variant 1 pads them to 80 bytes.
template <int S, bool SWZ>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2)))
void gemm_i8x_o2_kernel(GemmArgs g, ProdArgs o, int tiles_m, int tiles_n) {
  // SWZ: 64-byte LDS rows whose four 16-byte pieces sit at slots q ^ f(r)
  // (f(r) = ((r >> 1) ^ (r >> 3)) & 3), conflict-free for the stores (four
  // lanes per row) and the fragment reads (sixteen rows per piece) alike;
  // otherwise 80-byte rows (conflict-free reads only)
  constexpr int kPitch = SWZ ? 64 : kRowB;
  __shared__ __attribute__((aligned(16))) int8_t lds[2][S][kTile][kPitch];

  int bid = blockIdx.x;
  const int total = gridDim.x;
  if ((total & 7) == 0) bid = (bid & 7) * (total >> 3) + (bid >> 3);
  const int tiles = tiles_m * tiles_n;
  const int z = bid / tiles;
  const int t = bid - z * tiles;
  const int z_out = z / g.inner, z_in = z - z_out * g.inner;
  const int m0 = (t / tiles_n) * kTile, n0 = (t % tiles_n) * kTile;

  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave >> 1, wn = wave & 1;
  const int Kp = o.Kp, nk = Kp / kChunk;

  const long long za = zindex(o.a_mode, z_in, z_out, g.inner);
  const long long zb = zindex(o.b_mode, z_in, z_out, g.inner);
  const int lr = tid >> 2, lc = (tid & 3) * 16;
  const int8_t* a_src = o.a_data + (size_t)za * 3 * S * o.a_Rp * Kp + (size_t)(m0 + lr) * 64 + lc;
  const int8_t* b_src = o.b_data + (size_t)zb * 3 * S * o.b_Rp * Kp + (size_t)(n0 + lr) * 64 + lc;
  const int fr = lane & 15, fk = (lane >> 4) * 16;
  auto swz = [](int r) { return ((r >> 1) ^ (r >> 3)) & 3; };  // of the row mod 16
  // store slot of this thread's piece, fragment slot of this lane's piece
  const int st = SWZ ? (((lc >> 4) ^ swz(lr & 15)) << 4) : lc;
  const int fs = SWZ ? (((fk >> 4) ^ swz(fr)) << 4) : fk;

  double cr[2][2][4], ci[2][2][4];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int v = 0; v < 4; ++v) cr[i][j][v] = ci[i][j][v] = 0.0;

#pragma unroll 1
  for (int p = 0; p < 3; ++p) {
    i4v acc[2][2][S];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j)
#pragma unroll
        for (int d = 0; d < S; ++d) acc[i][j][d] = i4v{0, 0, 0, 0};
#pragma unroll 1
    for (int kc = 0; kc < nk; ++kc) {
      {
        i4v ra[S], rb[S];
#pragma unroll
        for (int d = 0; d < S; ++d) {
          ra[d] = *reinterpret_cast<const i4v*>(a_src + dig_off(S, nk, o.a_Rp, p, kc, d, 0));
          rb[d] = *reinterpret_cast<const i4v*>(b_src + dig_off(S, nk, o.b_Rp, p, kc, d, 0));
        }
        __syncthreads();  // every wave is done with the previous chunk
#pragma unroll
        for (int d = 0; d < S; ++d) {
          *reinterpret_cast<i4v*>(&lds[0][d][lr][st]) = ra[d];
          *reinterpret_cast<i4v*>(&lds[1][d][lr][st]) = rb[d];
        }
        __syncthreads();
      }
      // one column tile at a time (its S digit fragments in registers); the
      // A fragments are read once per column tile
#pragma unroll
      for (int ct = 0; ct < 2; ++ct) {
        i4v bf[S];
#pragma unroll
        for (int u = 0; u < S; ++u)
          bf[u] = *reinterpret_cast<const i4v*>(&lds[1][u][wn * 32 + ct * 16 + fr][fs]);
#pragma unroll
        for (int ta = 0; ta < S; ++ta)
#pragma unroll
          for (int rt = 0; rt < 2; ++rt) {
            const i4v af =
                *reinterpret_cast<const i4v*>(&lds[0][ta][wm * 32 + rt * 16 + fr][fs]);
#pragma unroll
            for (int ub = 0; ta + ub < S; ++ub)
              acc[rt][ct][ta + ub] = __builtin_amdgcn_mfma_i32_16x16x64_i8(
                  af, bf[ub], acc[rt][ct][ta + ub], 0, 0, 0);
          }
      }
    }
    const double* sa = o.a_scale + ((size_t)za * 3 + p) * o.a_Rp + m0 + wm * 32;
    const double* sb = o.b_scale + ((size_t)zb * 3 + p) * o.b_Rp + n0 + wn * 32;
    const double sr = p == 0 ? 1.0 : -1.0;
    const double si = p == 2 ? 1.0 : -1.0;
#pragma unroll
    for (int rt = 0; rt < 2; ++rt)
#pragma unroll
      for (int ct = 0; ct < 2; ++ct) {
        const double scol = sb[ct * 16 + fr];
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          double val = 0.0;
#pragma unroll
          for (int d = S - 1; d >= 0; --d)
            val += (double)acc[rt][ct][d][v] * __builtin_ldexp(1.0, -14 - 8 * d);
          val *= sa[rt * 16 + (lane >> 4) * 4 + v] * scol;
          if (p < 2) cr[rt][ct][v] += sr * val;
          ci[rt][ct][v] += si * val;
        }
      }
  }

  cplx* C = g.C + z_in * g.sC_in + z_out * g.sC_out;
  const cplx* C0 = g.C0 ? g.C0 + z_in * g.sC0_in + z_out * g.sC0_out : nullptr;
  const cplx sh = g.shift ? g.shift[z_out] : cplx{0.0, 0.0};
#pragma unroll
  for (int rt = 0; rt < 2; ++rt)
#pragma unroll
    for (int ct = 0; ct < 2; ++ct)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int row = m0 + wm * 32 + rt * 16 + (lane >> 4) * 4 + v;
        const int col = n0 + wn * 32 + ct * 16 + fr;
        if (row < g.M && col < g.N) {
          const double tr = cr[rt][ct][v], ti = ci[rt][ct][v];
          double re = g.alpha.re * tr - g.alpha.im * ti;
          double im = g.alpha.re * ti + g.alpha.im * tr;
          if (C0) {
            const cplx c0 = C0[(long long)row * g.ldc0 + col];
            re += g.beta.re * c0.re - g.beta.im * c0.im;
            im += g.beta.re * c0.im + g.beta.im * c0.re;
          }
          if (row == col) {
            re += sh.re;
            im += sh.im;
          }
          C[(long long)row * g.ldc + col] = cplx{re, im};
        }
      }
}

int g_i8x_variant = -1;  // forced product kernel (-1: default)
constexpr int kI8xDefaultVariant = 2;

struct I8xPlan {
  int a_mode, b_mode;
  long long a_nz, b_nz;
  int Mp, Np, Kp;
  size_t a_data, a_scale, b_data, b_scale, pmax, total;  // byte offsets / total
};

inline size_t al256(size_t x) { return (x + 255) & ~(size_t)255; }

I8xPlan plan_i8x(const GemmArgs& g, int S) {
  I8xPlan p{};
  p.a_mode = zmode(g.sA_in, g.sA_out);
  p.b_mode = zmode(g.sB_in, g.sB_out);
  p.a_nz = zcount(p.a_mode, g.inner, g.outer);
  p.b_nz = zcount(p.b_mode, g.inner, g.outer);
  p.Mp = (int)round_up(g.M, kTile);
  p.Np = (int)round_up(g.N, kTile);
  p.Kp = (int)round_up(g.K, kChunk);
  size_t off = 0;
  p.a_data = off;
  off = al256(off + (size_t)p.a_nz * 3 * S * p.Mp * p.Kp);
  p.a_scale = off;
  off = al256(off + (size_t)p.a_nz * 3 * p.Mp * sizeof(double));
  p.b_data = off;
  off = al256(off + (size_t)p.b_nz * 3 * S * p.Np * p.Kp);
  p.b_scale = off;
  off = al256(off + (size_t)p.b_nz * 3 * p.Np * sizeof(double));
  // partial row maxima (one per 64-wide k chunk) of the two-pass split
  const size_t nkc = (size_t)p.Kp / kChunk;
  const size_t pa = (size_t)p.a_nz * nkc * 3 * p.Mp, pb = (size_t)p.b_nz * nkc * 3 * p.Np;
  p.pmax = off;
  if (nkc > (size_t)kSplitMaxChunks) off = al256(off + (pa > pb ? pa : pb) * sizeof(double));
  p.total = off + 256;  // room to align the base
  return p;
}

template <int S>
hipError_t launch_i8x_s(const GemmArgs& g, const I8xPlan& p, uint8_t* base, hipStream_t s,
                        bool reuse_a) {
  // the launch profiler times the digit split and the product separately
  int pid = prof_open(s);
  // op(A): rows m, k along K;  element (m, k) = A[m*lda + k] (N) or A[k*lda + m] (T, C)
  SliceArgs sa{};
  sa.X = g.A; sa.s_in = g.sA_in; sa.s_out = g.sA_out; sa.inner = g.inner; sa.mode = p.a_mode;
  sa.R = g.M; sa.K = g.K; sa.Rp = p.Mp; sa.Kp = p.Kp;
  sa.rs = g.opA == OP_N ? g.lda : 1; sa.ks = g.opA == OP_N ? 1 : g.lda;
  sa.conj = g.opA == OP_C ? -1.0 : 1.0;
  sa.data = reinterpret_cast<int8_t*>(base + p.a_data);
  sa.scale = reinterpret_cast<double*>(base + p.a_scale);
  // op(B): rows n (columns of op(B)); element (n, k) = B[k*ldb + n] (N) or B[n*ldb + k] (T, C)
  SliceArgs sb{};
  sb.X = g.B; sb.s_in = g.sB_in; sb.s_out = g.sB_out; sb.inner = g.inner; sb.mode = p.b_mode;
  sb.R = g.N; sb.K = g.K; sb.Rp = p.Np; sb.Kp = p.Kp;
  sb.rs = g.opB == OP_N ? 1 : g.ldb; sb.ks = g.opB == OP_N ? g.ldb : 1;
  sb.conj = g.opB == OP_C ? -1.0 : 1.0;
  sb.data = reinterpret_cast<int8_t*>(base + p.b_data);
  sb.scale = reinterpret_cast<double*>(base + p.b_scale);

  const int nkc = p.Kp / kChunk;
  double* pmax = reinterpret_cast<double*>(base + p.pmax);
  for (const SliceArgs* a : {&sa, &sb}) {
    if (a == &sa && reuse_a) continue;
    const unsigned nz = (unsigned)(a == &sa ? p.a_nz : p.b_nz);
    const bool kc_major = a->ks == 1;
    if (nkc <= kSplitMaxChunks) {
      const dim3 grid((unsigned)(a->Rp / 16), nz);
      if (kc_major)
        hipLaunchKernelGGL((split_rows_kernel<S, true>), grid, dim3(64 * nkc), 0, s, *a, nkc);
      else
        hipLaunchKernelGGL((split_rows_kernel<S, false>), grid, dim3(64 * nkc), 0, s, *a, nkc);
      SINV_HIP_RET(hipGetLastError());
      continue;
    }
    // the two operands' passes run one after the other on the stream, so they
    // share the partial-maximum buffer
    const dim3 grid((unsigned)(a->Rp / 64), (unsigned)nkc, nz);
    if (kc_major) {
      hipLaunchKernelGGL(piece_max_kernel<true>, grid, dim3(256), 0, s, *a, pmax, nkc);
      SINV_HIP_RET(hipGetLastError());
      hipLaunchKernelGGL((slice_piece_kernel<S, true>), grid, dim3(256), 0, s, *a, pmax, nkc);
    } else {
      hipLaunchKernelGGL(piece_max_kernel<false>, grid, dim3(256), 0, s, *a, pmax, nkc);
      SINV_HIP_RET(hipGetLastError());
      hipLaunchKernelGGL((slice_piece_kernel<S, false>), grid, dim3(256), 0, s, *a, pmax, nkc);
    }
    SINV_HIP_RET(hipGetLastError());
  }

  {
    // split: each distinct operand copy read once (16 B per complex element),
    // its digits (3 S B per padded element) and scales written
    const double an = reuse_a ? 0.0 : (double)p.a_nz;
    const double rd = 16.0 * (an * g.M * g.K + (double)p.b_nz * g.N * g.K);
    const double wr = 3.0 * S * (an * p.Mp + (double)p.b_nz * p.Np) * p.Kp +
                      24.0 * (an * p.Mp + (double)p.b_nz * p.Np);
    prof_close(pid, s, PROF_I8X_SPLIT, 0.0, rd + wr);
  }
  pid = prof_open(s);

  ProdArgs o{};
  o.a_data = sa.data; o.a_scale = sa.scale; o.a_mode = p.a_mode; o.a_Rp = p.Mp;
  o.b_data = sb.data; o.b_scale = sb.scale; o.b_mode = p.b_mode; o.b_Rp = p.Np;
  o.Kp = p.Kp;
  const int tm = p.Mp / kTile, tn = p.Np / kTile;
  const long long blocks = (long long)tm * tn * g.inner * g.outer;
  const int variant = g_i8x_variant >= 0 ? g_i8x_variant : kI8xDefaultVariant;
  if (variant == 0)
    hipLaunchKernelGGL(gemm_i8x_kernel<S>, dim3((unsigned)blocks), dim3(256), 0, s, g, o, tm, tn);
  else if (variant == 2)
    hipLaunchKernelGGL((gemm_i8x_o2_kernel<S, true>), dim3((unsigned)blocks), dim3(256), 0, s, g, o,
                       tm, tn);
  else
    hipLaunchKernelGGL((gemm_i8x_o2_kernel<S, false>), dim3((unsigned)blocks), dim3(256), 0, s, g, o,
                       tm, tn);
  SINV_HIP_RET(hipGetLastError());
  // product: the digit planes read once, C written (and C0 read) once
  const double nb = (double)g.inner * g.outer;
  const double digits = 3.0 * S * ((double)p.a_nz * p.Mp + (double)p.b_nz * p.Np) * p.Kp;
  prof_shape(pid, g.M, g.N, g.K, (long long)nb, g.opA | (g.opB << 2) | (g.C0 != nullptr ? 16 : 0));
  prof_close(pid, s, PROF_GEMM_I8X, 8.0 * g.M * (double)g.N * g.K * nb,
             digits + sizeof(cplx) * (double)g.M * g.N * nb * (g.C0 != nullptr ? 2.0 : 1.0));
  return hipSuccess;
}

}  // namespace

bool i8x_slices_ok(int S) { return S >= kI8xMinSlices && S <= kI8xMaxSlices; }

size_t gemm_i8x_workspace_bytes(const GemmArgs& g, int S) {
  if (!i8x_slices_ok(S) || g.M <= 0 || g.N <= 0 || g.K <= 0 || g.inner <= 0 || g.outer <= 0)
    return 0;
  return plan_i8x(g, S).total;
}

hipError_t launch_gemm_i8x(const GemmArgs& g, int S, void* work, size_t work_bytes,
                           hipStream_t s, bool reuse_a) {
  if (g.M <= 0 || g.N <= 0 || g.inner <= 0 || g.outer <= 0) return hipSuccess;
  if (g.K <= 0) return launch_gemm(g, Prec::F64, s);  // no product term: C = beta C0 + shift
  if (!i8x_slices_ok(S)) return hipErrorInvalidValue;
  const I8xPlan p = plan_i8x(g, S);
  if (p.Kp > kMaxKp) return hipErrorInvalidValue;
  if (work == nullptr || work_bytes < p.total) return hipErrorInvalidValue;
  const long long blocks = (long long)(p.Mp / kTile) * (p.Np / kTile) * g.inner * g.outer;
  if (blocks > 0x7fffffffLL || p.a_nz > 65535 || p.b_nz > 65535)
    return hipErrorInvalidConfiguration;
  uint8_t* base = (uint8_t*)(((uintptr_t)work + 255) & ~(uintptr_t)255);
  switch (S) {
    case 5: return launch_i8x_s<5>(g, p, base, s, reuse_a);
    case 6: return launch_i8x_s<6>(g, p, base, s, reuse_a);
    default: return launch_i8x_s<7>(g, p, base, s, reuse_a);
  }
}

}  // namespace sinv

using namespace sinv;

extern "C" size_t sinv_zgemm_int8x_workspace_bytes(int M, int N, int K, long long strideA,
                                                   long long strideB, int batch, int slices) {
  GemmArgs g{};
  g.M = M; g.N = N; g.K = K;
  g.sA_out = strideA; g.sB_out = strideB;
  g.inner = 1; g.outer = batch;
  return gemm_i8x_workspace_bytes(g, slices);
}

extern "C" int sinv_zgemm_int8x_batched(int M, int N, int K, int opA, int opB, sinv_c128 alpha,
                                        const void* A, int lda, long long strideA, const void* B,
                                        int ldb, long long strideB, sinv_c128 beta,
                                        const void* C0, int ldc0, long long strideC0, void* C,
                                        int ldc, long long strideC, const void* shift, int batch,
                                        int slices, void* work, size_t work_bytes, void* stream) {
  if (M < 0 || N < 0 || K < 0 || batch < 0 || opA < 0 || opA > 2 || opB < 0 || opB > 2 ||
      !i8x_slices_ok(slices) || strideA < 0 || strideB < 0 || strideC < 0 || strideC0 < 0) {
    set_error("sinv_zgemm_int8x_batched: invalid argument");
    return SINV_E_INVAL;
  }
  if (M == 0 || N == 0 || batch == 0) return SINV_OK;
  if (A == nullptr || B == nullptr || C == nullptr || (K == 0 && C0 == nullptr)) {
    set_error("sinv_zgemm_int8x_batched: null operand");
    return SINV_E_INVAL;
  }
  const int ra = opA == OP_N ? K : M, rb = opB == OP_N ? N : K;
  if (lda < (ra > 1 ? ra : 1) || ldb < (rb > 1 ? rb : 1) || ldc < N || (C0 && ldc0 < N)) {
    set_error("sinv_zgemm_int8x_batched: leading dimension too small");
    return SINV_E_INVAL;
  }
  if ((long long)K > kMaxKp) {
    set_error("sinv_zgemm_int8x_batched: K = %d exceeds %d (int32 digit sums)", K, kMaxKp);
    return SINV_E_INVAL;
  }
  GemmArgs g{};
  g.M = M; g.N = N; g.K = K; g.opA = opA; g.opB = opB;
  g.A = (const cplx*)A; g.lda = lda; g.sA_out = strideA;
  g.B = (const cplx*)B; g.ldb = ldb; g.sB_out = strideB;
  g.C0 = (const cplx*)C0; g.ldc0 = ldc0; g.sC0_out = strideC0;
  g.C = (cplx*)C; g.ldc = ldc; g.sC_out = strideC;
  g.alpha = cplx{alpha.re, alpha.im};
  g.beta = cplx{beta.re, beta.im};
  g.shift = (const cplx*)shift;
  g.inner = 1; g.outer = batch;
  const size_t need = K > 0 ? gemm_i8x_workspace_bytes(g, slices) : 0;
  if (K > 0 && (work == nullptr || work_bytes < need)) {
    set_error("sinv_zgemm_int8x_batched: workspace too small (%zu < %zu)", work_bytes, need);
    return SINV_E_INVAL;
  }
  SINV_HIP_TRY(launch_gemm_i8x(g, slices, work, work_bytes, (hipStream_t)stream));
  return SINV_OK;
}

extern "C" int sinv_int8x_force_variant(int variant) {
  if (variant < -1 || variant > 2) {
    set_error("sinv_int8x_force_variant: variant must be -1, 0, 1 or 2");
    return SINV_E_INVAL;
  }
  sinv::g_i8x_variant = variant;
  return SINV_OK;
}
