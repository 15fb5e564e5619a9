// Internal declarations shared by the HIP translation units of libsinv_hip.
//
// Storage convention: complex128 is two interleaved doubles (numpy / torch
// layout), blocks are row-major with a leading dimension.  Every launcher here
// is stream-ordered on the caller's hipStream_t and allocates nothing.
#pragma once

#include <hip/hip_runtime.h>
#include <cstddef>
#include <cstdint>
#include <vector>

#include "../../include/sinv.h"

namespace sinv {

struct cplx {
  double re, im;
};

enum Op : int { OP_N = 0, OP_T = 1, OP_C = 2 };

// C[z] = alpha * op(A[z]) * op(B[z]) + beta * C0[z] + shift[z] * I
//   op(A) is M x K, op(B) is K x N, C is M x N.
//   Batch index z = z_outer * inner + z_inner addresses
//     X[z] = X + z_inner * sX_in + z_outer * sX_out   (element strides)
//   so one launch covers (energies x block rows) without pointer arrays.
//   shift (device, indexed by z_outer) may be null; C0 may alias C.
struct GemmArgs {
  int M, N, K;
  int opA, opB;
  const cplx* A;  int lda;  long long sA_in, sA_out;
  const cplx* B;  int ldb;  long long sB_in, sB_out;
  const cplx* C0; int ldc0; long long sC0_in, sC0_out;
  cplx* C;        int ldc;  long long sC_in, sC_out;
  cplx alpha, beta;
  const cplx* shift;
  int inner, outer;  // batch = inner * outer
  // optional per-instance early exit: instance z (= z_outer * inner + z_inner)
  // is skipped when skip[z] < skip_iter (iterative inverse: instances that
  // have converged stop); null = never skip.  FP64 / FP32 kernels only.
  const int* skip;
  int skip_iter;
  // launch-profiler kind of this product + 1, recorded with no algorithmic
  // flops (the iterative inverse's products); 0: a GEMM record with 8 MNK
  int prof_as;
};

enum class Prec : int { F64 = 0, F32 = 1 };

// K2: batched complex GEMM on MFMA (FP64 16x16x4 for F64, FP32 16x16x4 for F32).
hipError_t launch_gemm(const GemmArgs& g, Prec p, hipStream_t s);

// K2 in FP64-accurate sliced-int8 form (int8x.hip): the same product from S
// balanced base-256 digit planes per operand on the int8 matrix cores, exact
// int32 sums.  The workspace holds the digit planes and row scales of both
// operands.  reuse_a: op(A)'s digits and scales are already in the workspace
// (the previous call on this workspace and stream had the same op(A), M, K and
// batch layout), so its split is skipped.
constexpr int kI8xMinSlices = 5, kI8xMaxSlices = 7, kI8xDefaultSlices = 6;
bool i8x_slices_ok(int S);
size_t gemm_i8x_workspace_bytes(const GemmArgs& g, int S);
hipError_t launch_gemm_i8x(const GemmArgs& g, int S, void* work, size_t work_bytes,
                           hipStream_t s, bool reuse_a = false);

// C[z] = a * D[z] + shift[z_outer] * I  (elementwise, b x b blocks), batched as GemmArgs.
hipError_t launch_axpby_shift(int M, int N, cplx a, const cplx* D, int ldd, long long sD_in,
                              long long sD_out, cplx* C, int ldc, long long sC_in,
                              long long sC_out, const cplx* shift, int inner, int outer,
                              hipStream_t s);

// C[z] = alpha op(S[z]) + shift[z_outer] I  (op = T or C), M x N output from an
// N x M input; the shift is conjugated when conj_shift is set.
hipError_t launch_transpose(int M, int N, int op, const cplx* S, int lds, long long sS_in,
                            long long sS_out, cplx* C, int ldc, long long sC_in,
                            long long sC_out, int inner, int outer, hipStream_t s,
                            cplx alpha = cplx{1.0, 0.0}, const cplx* shift = nullptr,
                            bool conj_shift = false);

// K1: in-place batched inverse of n x n blocks at A + z_in*sA_in + z_out*sA_out.
// Gauss-Jordan elimination with partial pivoting over all n rows: one
// register-resident workgroup per block for n <= kLeaf, column-blocked with
// tall panels and GEMM sweeps for kLeaf < n <= kMaxInverseN (workspace from
// inverse_workspace_elems).  A zero pivot records key = (tag << 32 | z) into
// info (atomic min).
constexpr int kLeaf = 64;
constexpr int kMaxInverseN = 8192;
// Newton-Schulz iteration (newton = true, for drivers whose thread selected
// SINV_INV_NEWTON with sinv_inverse_method): launches of kLeaf < n blocks
// with at most kNewtonMaxBatch instances run X <- X (2I - M X) from the
// Jacobi start instead of the elimination; instances whose residual does not
// fall below the acceptance bound within the iteration limit are reported
// through info with SINV_INFO_NOT_CONVERGED set in the key.  The workspace
// must then come from inverse_workspace_elems(n, batch, true).
constexpr int kNewtonMaxBatch = 4;
int inverse_method();
int newton_max_iter();
size_t inverse_workspace_elems(int n, int batch, bool newton = false);
// Newton-Schulz inverse of the nb = inner * outer <= kNewtonMaxBatch blocks M
// into X (newton.hip); its buffers follow the elimination's `base` elements
// of `work` (newton_workspace_elems(n) more).  in_place: M aliases X and is
// copied aside first.
size_t newton_workspace_elems(int n);
hipError_t launch_newton_inverse(const cplx* M, int n, int ldm, long long sm_in, long long sm_out,
                                 int inner, int outer, cplx* X, int ldx, long long sx_in,
                                 long long sx_out, bool in_place, cplx* work, size_t base,
                                 unsigned long long* info, int tag, hipStream_t s);
hipError_t launch_inverse(cplx* A, int n, int lda, long long sA_in, long long sA_out, int inner,
                          int outer, cplx* work, unsigned long long* info, int tag, hipStream_t s,
                          bool newton = false);
// Out of place: A is destroyed (it is the elimination's working copy) and
// its inverse is written to X.  Same workspace as launch_inverse.
hipError_t launch_inverse_to(cplx* A, int n, int lda, long long sA_in, long long sA_out, int inner,
                             int outer, cplx* X, int ldx, long long sX_in, long long sX_out,
                             cplx* work, unsigned long long* info, int tag, hipStream_t s,
                             bool newton = false);

// Fill n elements with a value (used for NaN off-diagonals and info reset).
hipError_t launch_fill(cplx* p, long long n, cplx v, hipStream_t s);
hipError_t launch_fill_u64(unsigned long long* p, long long n, unsigned long long v, hipStream_t s);

// Packed block-tridiagonal layout with per-block sizes b_i: diagonal block i
// at doff[i] (b_i x b_i), upper block i (b_i x b_{i+1}) and lower block i
// (b_{i+1} x b_i) at ooff[i]; dsize / osize are the per-instance totals.
struct BlockLayout {
  int N = 0;
  std::vector<int> b;
  std::vector<long long> doff, ooff;
  long long dsize = 0, osize = 0;
  int bmax = 0;
  bool ok = false;
};

inline BlockLayout make_block_layout(int N, int blocksize, const int* bsizes) {
  BlockLayout L;
  if (N <= 0) return L;
  L.N = N;
  L.b.resize(N);
  for (int i = 0; i < N; ++i) {
    L.b[i] = bsizes ? bsizes[i] : blocksize;
    if (L.b[i] <= 0) return L;
    if (L.b[i] > L.bmax) L.bmax = L.b[i];
  }
  L.doff.resize(N);
  L.ooff.resize(N > 1 ? N - 1 : 0);
  long long d = 0, o = 0;
  for (int i = 0; i < N; ++i) {
    L.doff[i] = d;
    d += (long long)L.b[i] * L.b[i];
    if (i < N - 1) {
      L.ooff[i] = o;
      o += (long long)L.b[i] * L.b[i + 1];
    }
  }
  L.dsize = d;
  L.osize = o;
  L.ok = true;
  return L;
}

// Fork/join of independent launch sequences onto library-owned side streams
// (default priority, non-blocking, created once per device on first use).
// The constructor makes nside side streams wait for the work already queued
// on `main`; join() (or the destructor) makes `main` wait for everything
// queued on them since.  Stream capture follows the fork and the join like
// any cross-stream event dependency.  Callers that fork concurrently share
// the side streams: their side work serialises, their results do not mix.
class StreamFork {
 public:
  static constexpr int kMaxSide = 2;
  StreamFork(hipStream_t main, int nside);
  ~StreamFork() { join(); }
  StreamFork(const StreamFork&) = delete;
  StreamFork& operator=(const StreamFork&) = delete;
  hipStream_t side(int i) const { return side_[i]; }
  hipError_t status() const { return err_; }
  hipError_t join();

 private:
  hipStream_t main_;
  int n_ = 0;
  hipStream_t side_[kMaxSide] = {};
  hipEvent_t fork_ = nullptr;
  bool joined_ = false;
  hipError_t err_ = hipSuccess;
};

// thread-local last error text for sinv_last_error()
void set_error(const char* fmt, ...);

// launch profiler (prof.hip): prof_open before a launch returns an id (-1 when
// off), prof_close after it records the kernel kind and its algorithmic flops
// and bytes (the operands read and results written once each, as stored).
enum ProfKind : int {
  PROF_GEMM_F64 = 0,
  PROF_GEMM_F32 = 1,
  PROF_INV_LEAF = 2,
  PROF_OTHER = 3,
  PROF_GEMM_I8X = 4,
  PROF_INV_PANEL = 5,
  PROF_INV_NEWTON = 6,
  PROF_I8X_SPLIT = 7
};
int prof_open(hipStream_t s);
void prof_close(int id, hipStream_t s, int kind, double flops, double bytes);
// shape of a profiled GEMM launch (kept in the per-launch records)
void prof_shape(int id, int M, int N, int K, long long batch, int flags);

}  // namespace sinv

#define SINV_HIP_TRY(expr)                                                        \
  do {                                                                            \
    hipError_t _e = (expr);                                                       \
    if (_e != hipSuccess) {                                                       \
      ::sinv::set_error("%s:%d: %s", __FILE__, __LINE__, hipGetErrorString(_e)); \
      return SINV_E_HIP;                                                          \
    }                                                                             \
  } while (0)

#define SINV_HIP_RET(expr)                \
  do {                                    \
    hipError_t _e = (expr);               \
    if (_e != hipSuccess) return _e;      \
  } while (0)
