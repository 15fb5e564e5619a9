// Dense <-> block-tridiagonal converters on the device (matrices_utils.py:115-189).
//
// Pure data movement, HBM-bound: each thread moves one complex128 (16 B) with
// a vector load and store; consecutive threads walk a block row, so reads
// and writes are coalesced.  A batch of matrices is the outer grid dimension.
//
//  * dense_to_blocks: gather the 3N - 2 blocks of the tridiagonal band into
//    the stacked layout diag (N,b,b), upper (N-1,b,b), lower (N-1,b,b).
//  * blocks_to_dense: one pass over the whole n x n output; an element in the
//    band is read from its block, every other element is written as zero (so
//    no separate fill pass touches the output twice).
//  * scatter_blocks: copy extra b x b blocks to given block positions (the
//    fill-in blocks of block cyclic reduction's dense result).
#include "sinv_core.hpp"

namespace sinv {

namespace {

__device__ __forceinline__ void mov16(const cplx* src, cplx* dst) {
  *reinterpret_cast<double2*>(dst) = *reinterpret_cast<const double2*>(src);
}

// grid: x over (3N - 2) * b * b elements, y over the batch
__global__ void dense_to_blocks_kernel(const cplx* A, int lda, long long sA, int N, int b,
                                       cplx* D, cplx* U, cplx* L) {
  const long long bb = (long long)b * b;
  const long long total = (3LL * N - 2) * bb;
  const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= total) return;
  const long long z = blockIdx.y;
  const int t = (int)(e / bb);
  const int w = (int)(e - (long long)t * bb);
  const int r = w / b, c = w - (w / b) * b;
  int br, bc;
  cplx* dst;
  if (t < N) {
    br = bc = t;
    dst = D + z * N * bb + (long long)t * bb;
  } else if (t < 2 * N - 1) {
    br = t - N;
    bc = br + 1;
    dst = U + z * (N - 1) * bb + (long long)br * bb;
  } else {
    bc = t - (2 * N - 1);
    br = bc + 1;
    dst = L + z * (N - 1) * bb + (long long)bc * bb;
  }
  mov16(A + z * sA + (long long)(br * b + r) * lda + (bc * b + c), dst + w);
}

// grid: x over n * n elements, y over the batch
__global__ void blocks_to_dense_kernel(const cplx* D, const cplx* U, const cplx* L, int N, int b,
                                       cplx* A, int lda, long long sA) {
  const long long n = (long long)N * b;
  const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n * n) return;
  const long long z = blockIdx.y;
  const int row = (int)(e / n), col = (int)(e - (long long)row * n);
  const int br = row / b, bc = col / b;
  const int r = row - br * b, c = col - bc * b;
  const long long bb = (long long)b * b;
  cplx* dst = A + z * sA + (long long)row * lda + col;
  if (br == bc)
    mov16(D + z * N * bb + (long long)br * bb + (long long)r * b + c, dst);
  else if (bc == br + 1)
    mov16(U + z * (N - 1) * bb + (long long)br * bb + (long long)r * b + c, dst);
  else if (br == bc + 1)
    mov16(L + z * (N - 1) * bb + (long long)bc * bb + (long long)r * b + c, dst);
  else
    *reinterpret_cast<double2*>(dst) = double2{0.0, 0.0};
}

// grid: x over count * b * b elements, y over the batch; block k goes to block
// position (rows[k], cols[k]) (device arrays); positions outside the
// nblocks x nblocks block grid are skipped
__global__ void scatter_blocks_kernel(const cplx* S, long long sS, int count, const int* rows,
                                      const int* cols, int nblocks, int b, cplx* A, int lda,
                                      long long sA) {
  const long long bb = (long long)b * b;
  const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (long long)count * bb) return;
  const long long z = blockIdx.y;
  const int k = (int)(e / bb);
  if (rows[k] < 0 || rows[k] >= nblocks || cols[k] < 0 || cols[k] >= nblocks) return;
  const int w = (int)(e - (long long)k * bb);
  const int r = w / b, c = w - (w / b) * b;
  mov16(S + z * sS + e, A + z * sA + (long long)(rows[k] * b + r) * lda + (cols[k] * b + c));
}

constexpr int kThreads = 256;

bool grid_ok(long long elems, int batch, dim3& grid) {
  const long long nb = (elems + kThreads - 1) / kThreads;
  if (nb <= 0 || nb > 0x7fffffffLL || batch <= 0 || batch > 65535) return false;
  grid = dim3((unsigned)nb, (unsigned)batch);
  return true;
}

}  // namespace

}  // namespace sinv

using namespace sinv;

extern "C" int sinv_dense_to_blocks(const void* A, int n, int lda, long long strideA, int b,
                                    int batch, void* diag, void* upper, void* lower,
                                    void* stream) {
  if (A == nullptr || diag == nullptr || n <= 0 || b <= 0 || batch <= 0 || lda < n ||
      n % b != 0 || strideA < 0) {
    set_error("sinv_dense_to_blocks: invalid argument (n must be a multiple of b)");
    return SINV_E_INVAL;
  }
  const int N = n / b;
  if (N > 1 && (upper == nullptr || lower == nullptr)) {
    set_error("sinv_dense_to_blocks: upper/lower required for more than one block");
    return SINV_E_INVAL;
  }
  dim3 grid;
  if (!grid_ok((3LL * N - 2) * b * b, batch, grid)) {
    set_error("sinv_dense_to_blocks: size out of range");
    return SINV_E_INVAL;
  }
  hipLaunchKernelGGL(dense_to_blocks_kernel, grid, dim3(kThreads), 0, (hipStream_t)stream,
                     (const cplx*)A, lda, strideA, N, b, (cplx*)diag, (cplx*)upper,
                     (cplx*)lower);
  SINV_HIP_TRY(hipGetLastError());
  return SINV_OK;
}

extern "C" int sinv_blocks_to_dense(const void* diag, const void* upper, const void* lower,
                                    int nblocks, int b, int batch, void* A, int lda,
                                    long long strideA, void* stream) {
  const long long n = (long long)nblocks * b;
  if (diag == nullptr || A == nullptr || nblocks <= 0 || b <= 0 || batch <= 0 || lda < n ||
      strideA < 0 || (nblocks > 1 && (upper == nullptr || lower == nullptr))) {
    set_error("sinv_blocks_to_dense: invalid argument");
    return SINV_E_INVAL;
  }
  dim3 grid;
  if (!grid_ok(n * n, batch, grid)) {
    set_error("sinv_blocks_to_dense: size out of range");
    return SINV_E_INVAL;
  }
  hipLaunchKernelGGL(blocks_to_dense_kernel, grid, dim3(kThreads), 0, (hipStream_t)stream,
                     (const cplx*)diag, (const cplx*)upper, (const cplx*)lower, nblocks, b,
                     (cplx*)A, lda, strideA);
  SINV_HIP_TRY(hipGetLastError());
  return SINV_OK;
}

extern "C" int sinv_scatter_blocks(const void* blocks, long long strideS, int count,
                                   const int* rows, const int* cols, int nblocks, int b,
                                   int batch, void* A, int lda, long long strideA, void* stream) {
  if (count == 0) return SINV_OK;
  if (blocks == nullptr || rows == nullptr || cols == nullptr || A == nullptr || count < 0 ||
      nblocks <= 0 || b <= 0 || batch <= 0 || lda < (long long)nblocks * b || strideA < 0 ||
      strideS < 0) {
    set_error("sinv_scatter_blocks: invalid argument");
    return SINV_E_INVAL;
  }
  dim3 grid;
  if (!grid_ok((long long)count * b * b, batch, grid)) {
    set_error("sinv_scatter_blocks: size out of range");
    return SINV_E_INVAL;
  }
  hipLaunchKernelGGL(scatter_blocks_kernel, grid, dim3(kThreads), 0, (hipStream_t)stream,
                     (const cplx*)blocks, strideS, count, rows, cols, nblocks, b, (cplx*)A, lda,
                     strideA);
  SINV_HIP_TRY(hipGetLastError());
  return SINV_OK;
}
