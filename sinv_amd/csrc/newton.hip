// K1 for the latency-bound case: Newton-Schulz inverse of a few blocks at a
// time, and the column-dominance test that decides when it applies.
//
// At one to four blocks per launch, the pivoted elimination's column-serial
// pivot chain runs on one CU per block while the rest of the chip idles (a
// 512 block takes ~1.4 ms).  For a strictly column diagonally dominant block
// M (sum_{i != j} |m_ij| < |m_jj| for every column j) the Newton-Schulz
// iteration
//     X_{k+1} = X_k (2I - M X_k),   I - M X_{k+1} = (I - M X_k)^2
// converges from the Jacobi start X_0 = D^-1 (D = diag(M)), since
// ||I - M D^-1||_1 is the dominance ratio, below 1.  Each step is two FP64
// GEMMs that spread over the whole chip (K2), so a 512 block converged in two
// steps costs four 512^3 products.  The first step, X_1 = 2 D^-1 - D^-1 M
// D^-1, is elementwise.
//
// Acceptance: after the product T = 2I - M X_k a check reduces
// max |(T - I)_ij| (|re| + |im|) per instance; once it is at most 1e-9 the
// step's second product gives X_{k+1} with ||I - M X_{k+1}||_2 <=
// (n 1e-9)^2 (rounding aside), and the instance's later products leave at
// once (GemmArgs::skip).  An instance still above the bound after the
// iteration limit is reported through info with SINV_INFO_NOT_CONVERGED set
// (its output is then undefined; the caller repeats with the pivoted
// elimination).  NaN and Inf (a zero diagonal entry) never pass the check.
//
// Dominance is inherited: eliminating a pivot from a strictly column
// dominant matrix leaves a Schur complement whose every column has a
// dominance ratio no larger than before (for column j the off-diagonal sum
// and the diagonal modulus both drop by |a_j1 a_1j| / |a_11| at most), so
// every block RGF, BCR, PSR or PDIV inverts from such a matrix is dominant
// with a ratio at most the matrix's; sinv_column_dominance measures it.
#include <algorithm>
#include <climits>

#include "sinv_core.hpp"

namespace sinv {

namespace {

thread_local int g_inv_method = SINV_INV_PIVOTED;
thread_local int g_newton_iter = 6;

constexpr double kNewtonTol = 1e-9;

__device__ __forceinline__ cplx ld_c(const cplx* p) {
  const double2 v = *reinterpret_cast<const double2*>(p);
  return cplx{v.x, v.y};
}

__device__ __forceinline__ void st_c(cplx* p, double re, double im) {
  *reinterpret_cast<double2*>(p) = double2{re, im};
}

// 1/p by the conjugate over |p|^2 (inputs here are scaled like matrix
// entries; a zero p gives Inf/NaN, which the acceptance check rejects)
__device__ __forceinline__ cplx crcp(cplx p) {
  const double d = 1.0 / (p.re * p.re + p.im * p.im);
  return cplx{p.re * d, -p.im * d};
}

__device__ __forceinline__ cplx cmul(cplx a, cplx b) {
  return cplx{a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}

// block-wide maximum of 64-bit keys (256 threads)
__device__ __forceinline__ unsigned long long block_umax(unsigned long long v) {
  __shared__ unsigned long long red[256];
  red[threadIdx.x] = v;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      const unsigned long long o = red[threadIdx.x + w];
      if (o > red[threadIdx.x]) red[threadIdx.x] = o;
    }
    __syncthreads();
  }
  return red[0];
}

// order-preserving key of a non-negative double (NaN above +Inf)
__device__ __forceinline__ unsigned long long nonneg_key(double x) {
  return (unsigned long long)__double_as_longlong(x);
}

// X_1 = 2 D^-1 - D^-1 M D^-1; block (x, 0) of each instance also resets the
// instance's iteration state.
__global__ __launch_bounds__(256) void ns_init_kernel(const cplx* M, int n, int ldm, long long sm_in,
                                                      long long sm_out, int inner, cplx* X,
                                                      int ldx, long long sx_in, long long sx_out,
                                                      cplx* two, int* conv, unsigned* cnt,
                                                      unsigned long long* dmax) {
  const int z = blockIdx.y;
  const int zo = z / inner, zi = z - zo * inner;
  M += zi * sm_in + zo * sm_out;
  X += zi * sx_in + zo * sx_out;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    two[z] = cplx{2.0, 0.0};
    conv[z] = INT_MAX;
    cnt[z] = 0u;
    dmax[z] = 0ull;
  }
  const long long total = (long long)n * n;
  for (long long e = (long long)blockIdx.x * 256 + threadIdx.x; e < total;
       e += (long long)gridDim.x * 256) {
    const int r = (int)(e / n), c = (int)(e - (long long)r * n);
    const cplx dr = crcp(ld_c(M + (long long)r * ldm + r));
    const cplx dc = crcp(ld_c(M + (long long)c * ldm + c));
    const cplx m = ld_c(M + (long long)r * ldm + c);
    cplx x = cmul(cmul(dr, m), dc);
    x.re = -x.re;
    x.im = -x.im;
    if (r == c) {
      x.re += 2.0 * dr.re;
      x.im += 2.0 * dr.im;
    }
    st_c(X + (long long)r * ldx + c, x.re, x.im);
  }
}

// Residual check of step j on T = 2I - M X_j (contiguous n x n per instance):
// max |T - I| per instance over all of the instance's workgroups; the last
// workgroup to finish marks the instance converged at j when the maximum is
// within kNewtonTol and clears the reduction state for the next step.
__global__ __launch_bounds__(256) void ns_check_kernel(const cplx* T, int n, int j,
                                                       unsigned long long tol_key, int* conv,
                                                       unsigned* cnt,
                                                       unsigned long long* dmax) {
  const int z = blockIdx.y;
  if (conv[z] < j) return;  // converged at an earlier step: uniform per instance
  T += (long long)z * n * n;
  const long long total = (long long)n * n;
  unsigned long long m = 0ull;
  for (long long e = (long long)blockIdx.x * 256 + threadIdx.x; e < total;
       e += (long long)gridDim.x * 256) {
    const int r = (int)(e / n), c = (int)(e - (long long)r * n);
    const cplx t = ld_c(T + e);
    const double v = fabs(t.re - (r == c ? 1.0 : 0.0)) + fabs(t.im);
    const unsigned long long k = nonneg_key(v);
    if (k > m) m = k;
  }
  m = block_umax(m);
  __shared__ bool last;
  if (threadIdx.x == 0) {
    atomicMax(dmax + z, m);
    __threadfence();
    last = atomicAdd(cnt + z, 1u) == gridDim.x - 1;
  }
  __syncthreads();
  if (last && threadIdx.x == 0) {
    __threadfence();
    const unsigned long long d = atomicAdd(dmax + z, 0ull);
    if (d <= tol_key) conv[z] = j;
    atomicExch(dmax + z, 0ull);
    atomicExch(cnt + z, 0u);
  }
}

// After the last step: an instance accepted at an odd step holds its result
// in Y (step j writes Y for odd j, X for even j); copy it to X.  Instances
// never accepted are reported through info.
__global__ __launch_bounds__(256) void ns_finish_kernel(cplx* X, int n, int ldx, long long sx_in,
                                                        long long sx_out, int inner, const cplx* Y,
                                                        const int* conv,
                                                        unsigned long long* info, int tag) {
  const int z = blockIdx.y;
  const int c = conv[z];
  if (c == INT_MAX) {
    if (blockIdx.x == 0 && threadIdx.x == 0 && info != nullptr)
      atomicMin(info, SINV_INFO_NOT_CONVERGED |
                          ((unsigned long long)(unsigned)tag << 32) | (unsigned long long)(unsigned)z);
    return;
  }
  if ((c & 1) == 0) return;
  const int zo = z / inner, zi = z - zo * inner;
  X += zi * sx_in + zo * sx_out;
  Y += (long long)z * n * n;
  const long long total = (long long)n * n;
  for (long long e = (long long)blockIdx.x * 256 + threadIdx.x; e < total;
       e += (long long)gridDim.x * 256) {
    const int r = (int)(e / n), cc = (int)(e - (long long)r * n);
    *reinterpret_cast<double2*>(X + (long long)r * ldx + cc) =
        *reinterpret_cast<const double2*>(Y + e);
  }
}

__global__ __launch_bounds__(256) void ns_copy_kernel(const cplx* A, int n, int lda, long long s_in,
                                                      long long s_out, int inner, cplx* W) {
  const int z = blockIdx.y;
  const int zo = z / inner, zi = z - zo * inner;
  A += zi * s_in + zo * s_out;
  W += (long long)z * n * n;
  const long long total = (long long)n * n;
  for (long long e = (long long)blockIdx.x * 256 + threadIdx.x; e < total;
       e += (long long)gridDim.x * 256) {
    const int r = (int)(e / n), c = (int)(e - (long long)r * n);
    *reinterpret_cast<double2*>(W + e) =
        *reinterpret_cast<const double2*>(A + (long long)r * lda + c);
  }
}

// Column dominance ratio of block column c of every instance: workgroup
// (column tile, z) takes 64 columns of one block column, its four waves
// splitting the rows four ways (a wave reads 64 consecutive columns of a row),
// partial sums combined through LDS.  rho holds the keys of non-negative
// doubles (atomic maximum; cleared beforehand).
__global__ __launch_bounds__(256) void dominance_kernel(
    const cplx* D, long long sd, const cplx* U, long long su, const cplx* L, long long sl,
    int nblocks, int blocksize, cplx dscale, cplx oscale, const cplx* shift,
    unsigned long long* rho) {
  // blockIdx.x = c * tiles + column tile of block column c
  const int z = blockIdx.y;
  const int bc = blocksize;
  const int tiles = (bc + 63) / 64;
  const int c = (int)blockIdx.x / tiles;
  const int j0 = ((int)blockIdx.x % tiles) * 64;
  const long long b2 = (long long)bc * bc;
  const cplx* Dc = D + z * sd + c * b2;
  // A_{c-1,c} = upper[c-1] and A_{c+1,c} = lower[c], where they exist
  const cplx* Uc = c > 0 ? U + z * su + (c - 1) * b2 : nullptr;
  const cplx* Lc = c < nblocks - 1 ? L + z * sl + c * b2 : nullptr;
  const double ds = sqrt(dscale.re * dscale.re + dscale.im * dscale.im);
  const double os = sqrt(oscale.re * oscale.re + oscale.im * oscale.im);
  const int j = j0 + (int)(threadIdx.x & 63);
  const int ph = (int)(threadIdx.x >> 6);
  double off = 0.0, coup = 0.0;
  if (j < bc) {
    for (int i = ph; i < bc; i += 4) {
      if (i == j) continue;
      const cplx a = ld_c(Dc + (long long)i * bc + j);
      off += sqrt(a.re * a.re + a.im * a.im);
    }
    if (Uc)
      for (int i = ph; i < bc; i += 4) {
        const cplx a = ld_c(Uc + (long long)i * bc + j);
        coup += sqrt(a.re * a.re + a.im * a.im);
      }
    if (Lc)
      for (int i = ph; i < bc; i += 4) {
        const cplx a = ld_c(Lc + (long long)i * bc + j);
        coup += sqrt(a.re * a.re + a.im * a.im);
      }
  }
  __shared__ double part[4][64];
  part[ph][threadIdx.x & 63] = ds * off + os * coup;
  __syncthreads();
  unsigned long long m = 0ull;
  if (ph == 0 && j < bc) {
    const double tot = part[0][threadIdx.x] + part[1][threadIdx.x] + part[2][threadIdx.x] +
                       part[3][threadIdx.x];
    const cplx sh = shift ? shift[z] : cplx{0.0, 0.0};
    const cplx dj = ld_c(Dc + (long long)j * bc + j);
    const double are = dscale.re * dj.re - dscale.im * dj.im + sh.re;
    const double aim = dscale.re * dj.im + dscale.im * dj.re + sh.im;
    const double ad = sqrt(are * are + aim * aim);
    double ratio = ad > 0.0 ? tot / ad : __builtin_inf();
    if (ratio != ratio) ratio = __builtin_inf();
    m = nonneg_key(ratio);
  }
  m = block_umax(m);
  if (threadIdx.x == 0) atomicMax(rho + z, m);
}

}  // namespace

int inverse_method() { return g_inv_method; }
int newton_max_iter() { return g_newton_iter; }

// Workspace (complex elements from `work`, after the elimination's region of
// base elements): Y (n^2 per instance), the in-place path's copy of M (n^2 per
// instance), the shift constant 2 per instance, then per-instance ints: the
// step at which the instance was accepted, the check's workgroup counter, and
// the residual maximum (64-bit).  T takes the elimination's n^2 region.
size_t newton_workspace_elems(int n) {
  return 2 * (size_t)n * n * kNewtonMaxBatch + 2 * (size_t)kNewtonMaxBatch;
}

hipError_t launch_newton_inverse(const cplx* M, int n, int ldm, long long sm_in, long long sm_out,
                                 int inner, int outer, cplx* X, int ldx, long long sx_in,
                                 long long sx_out, bool in_place, cplx* work, size_t base,
                                 unsigned long long* info, int tag, hipStream_t s) {
  const int nb = inner * outer;
  const long long n2 = (long long)n * n;
  cplx* T = work;
  cplx* Y = work + base;
  cplx* Mc = Y + n2 * kNewtonMaxBatch;
  cplx* two = Mc + n2 * kNewtonMaxBatch;
  int* conv = reinterpret_cast<int*>(two + kNewtonMaxBatch);
  unsigned* cnt = reinterpret_cast<unsigned*>(conv + kNewtonMaxBatch);
  unsigned long long* dmax = reinterpret_cast<unsigned long long*>(cnt + kNewtonMaxBatch);
  const unsigned gx = (unsigned)std::min<long long>((n2 + 255) / 256, 256);
  const unsigned gchk = (unsigned)std::min<long long>((n2 + 1023) / 1024, 128);
  const int pid = prof_open(s);
  if (in_place) {
    hipLaunchKernelGGL(ns_copy_kernel, dim3(gx, (unsigned)nb), dim3(256), 0, s, M, n, ldm, sm_in,
                       sm_out, inner, Mc);
    M = Mc;
    ldm = n;
    sm_in = n2;
    sm_out = n2 * inner;
  }
  hipLaunchKernelGGL(ns_init_kernel, dim3(gx, (unsigned)nb), dim3(256), 0, s, M, n, ldm, sm_in,
                     sm_out, inner, X, ldx, sx_in, sx_out, two, conv, cnt, dmax);
  prof_close(pid, s, PROF_INV_NEWTON, 0.0, 2.0 * sizeof(cplx) * (double)n2 * nb);
  SINV_HIP_RET(hipGetLastError());
  const unsigned long long tol_key =
      (unsigned long long)__builtin_bit_cast(long long, kNewtonTol);
  const int kmax = g_newton_iter;
  for (int j = 1; j <= kmax; ++j) {
    const bool odd = (j & 1) != 0;
    cplx* cur = odd ? X : Y;
    const int ldcur = odd ? ldx : n;
    const long long scur_in = odd ? sx_in : n2, scur_out = odd ? sx_out : n2 * inner;
    cplx* nxt = odd ? Y : X;
    const int ldnxt = odd ? n : ldx;
    const long long snxt_in = odd ? n2 : sx_in, snxt_out = odd ? n2 * inner : sx_out;
    // T = 2I - M X_j
    GemmArgs g{};
    g.M = g.N = g.K = n;
    g.opA = g.opB = OP_N;
    g.A = M; g.lda = ldm; g.sA_in = sm_in; g.sA_out = sm_out;
    g.B = cur; g.ldb = ldcur; g.sB_in = scur_in; g.sB_out = scur_out;
    g.C = T; g.ldc = n; g.sC_in = n2; g.sC_out = n2 * inner;
    g.alpha = cplx{-1.0, 0.0};
    g.beta = cplx{0.0, 0.0};
    g.shift = two;
    g.inner = inner;
    g.outer = outer;
    g.skip = conv;
    g.skip_iter = j;
    g.prof_as = PROF_INV_NEWTON + 1;
    SINV_HIP_RET(launch_gemm(g, Prec::F64, s));
    const int pc = prof_open(s);
    hipLaunchKernelGGL(ns_check_kernel, dim3(gchk, (unsigned)nb), dim3(256), 0, s, T, n, j,
                       tol_key, conv, cnt, dmax);
    prof_close(pc, s, PROF_INV_NEWTON, 0.0, sizeof(cplx) * (double)n2 * nb);
    SINV_HIP_RET(hipGetLastError());
    // X_{j+1} = X_j T
    GemmArgs h{};
    h.M = h.N = h.K = n;
    h.opA = h.opB = OP_N;
    h.A = cur; h.lda = ldcur; h.sA_in = scur_in; h.sA_out = scur_out;
    h.B = T; h.ldb = n; h.sB_in = n2; h.sB_out = n2 * inner;
    h.C = nxt; h.ldc = ldnxt; h.sC_in = snxt_in; h.sC_out = snxt_out;
    h.alpha = cplx{1.0, 0.0};
    h.beta = cplx{0.0, 0.0};
    h.inner = inner;
    h.outer = outer;
    h.skip = conv;
    h.skip_iter = j;
    h.prof_as = PROF_INV_NEWTON + 1;
    SINV_HIP_RET(launch_gemm(h, Prec::F64, s));
  }
  const int pf = prof_open(s);
  hipLaunchKernelGGL(ns_finish_kernel, dim3(gx, (unsigned)nb), dim3(256), 0, s, X, n, ldx, sx_in,
                     sx_out, inner, Y, conv, info, tag);
  // the algorithmic work of the inverse (8 n^3 per block) is booked here once
  prof_close(pf, s, PROF_INV_NEWTON, 8.0 * (double)n2 * n * nb, 2.0 * sizeof(cplx) * (double)n2 * nb);
  return hipGetLastError();
}

}  // namespace sinv

using namespace sinv;

extern "C" int sinv_inverse_method(int method, int max_iter) {
  if ((method != SINV_INV_PIVOTED && method != SINV_INV_NEWTON) || max_iter < 0 ||
      max_iter > 16) {
    set_error("sinv_inverse_method: unknown method %d or iteration limit %d outside 0..16",
              method, max_iter);
    return -1;
  }
  const int prev = g_inv_method;
  g_inv_method = method;
  if (max_iter > 0) g_newton_iter = max_iter;
  return prev;
}

extern "C" int sinv_column_dominance(const sinv_rgf_desc* d, double* rho, void* stream) {
  if (d == nullptr || rho == nullptr || d->nblocks <= 0 || d->batch <= 0 || d->diag == nullptr ||
      d->bsizes != nullptr || d->blocksize <= 0 ||
      (d->nblocks > 1 && (d->upper == nullptr || d->lower == nullptr)) || d->diag_stride < 0 ||
      d->upper_stride < 0 || d->lower_stride < 0) {
    set_error("sinv_column_dominance: invalid descriptor");
    return SINV_E_INVAL;
  }
  const BlockLayout L = make_block_layout(d->nblocks, d->blocksize, d->bsizes);
  if (!L.ok) {
    set_error("sinv_column_dominance: invalid block sizes");
    return SINV_E_INVAL;
  }
  hipStream_t s = (hipStream_t)stream;
  unsigned long long* keys = reinterpret_cast<unsigned long long*>(rho);
  SINV_HIP_TRY(launch_fill_u64(keys, d->batch, 0ull, s));
  const cplx* D = (const cplx*)d->diag;
  const cplx* U = (const cplx*)d->upper;
  const cplx* Lo = (const cplx*)d->lower;
  const cplx ds{d->dscale.re, d->dscale.im}, os{d->oscale.re, d->oscale.im};
  const cplx* sh = (const cplx*)d->shift;
  const unsigned tiles = (unsigned)((d->blocksize + 63) / 64);
  hipLaunchKernelGGL(dominance_kernel, dim3((unsigned)L.N * tiles, (unsigned)d->batch), dim3(256),
                     0, s, D, d->diag_stride, U, d->upper_stride, Lo, d->lower_stride, L.N,
                     d->blocksize, ds, os, sh, keys);
  SINV_HIP_TRY(hipGetLastError());
  return SINV_OK;
}
