/*
 * sinv.h -- C ABI of libsinv_hip, the MI355X (gfx950) selected-inversion hot
 * path for block-tridiagonal matrices.
 *
 * The reference (Nano-TCAD/SINV) is pure Python/numpy and has no C layer; each
 * entry point below names the reference function whose arithmetic it replaces.
 * Callers are the Python host layer (sinv_amd, via ctypes) or any C caller.
 *
 * Conventions
 *   - complex128 = two interleaved doubles (numpy/torch layout); blocks are
 *     row-major b x b; stacked layout diag (N,b,b), upper[i] = A_{i,i+1},
 *     lower[i] = A_{i+1,i} (src/sinv/utils/matrices_utils.py:141-151).
 *   - All pointers except descriptors and block-size tables are DEVICE
 *     pointers.  `stream` is a hipStream_t (NULL = default stream).
 *   - Nothing here allocates: the caller passes a workspace whose size comes
 *     from the matching *_workspace_bytes query.  Calls are asynchronous and
 *     stream-ordered; `info` is written on the device.
 *   - Return codes: SINV_OK, SINV_E_INVAL (bad arguments, nothing launched),
 *     SINV_E_HIP (a HIP runtime error; text via sinv_last_error()).
 *     Singular pivots are reported through `info` (device), not the return
 *     code, because the call does not synchronise: info holds
 *     (block << 32 | instance) of the first block whose elimination met a
 *     zero pivot (or whose hand-off wait ran out), or SINV_INFO_OK.
 */
#ifndef SINV_H_
#define SINV_H_

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SINV_OK 0
#define SINV_E_SINGULAR 1
#define SINV_E_INVAL 2
#define SINV_E_HIP 3

#define SINV_INFO_OK (~0ULL)

#define SINV_PREC_FP64 0  /* complex128 GEMMs on FP64 MFMA, FP64 inverses   */
#define SINV_PREC_MIXED 1 /* complex64 GEMMs on FP32 MFMA, FP64 inverses    */
#define SINV_PREC_INT8X 2 /* complex128 GEMMs from 6 balanced base-256 int8  */
                          /* digit planes per operand on the int8 matrix     */
                          /* cores (exact int32 digit sums, error <= 2^-42.6 */
                          /* K max|row| max|col| per real product), FP64     */
                          /* inverses; taken by the                          */
                          /* RGF and adjoint drivers; sinv_zgemm_batched     */
                          /* rejects it (sinv_zgemm_int8x_batched instead)   */
#define SINV_PREC_MIXED_REFINED 3 /* RGF only: complex64 forward Schur GEMMs, FP64 */
                          /* inverses, one FP64 refinement step of each g_i  */
                          /* against M_i formed in FP64 (R = M_i g_i - I,     */
                          /* g_i -= g_i R), FP64 backward sweep              */

#define SINV_OP_N 0
#define SINV_OP_T 1
#define SINV_OP_C 2

typedef struct sinv_c128 {
  double re, im;
} sinv_c128;

/* Library version, major*10000 + minor*100 + patch. */
int sinv_version(void);
/* Text of the last error on the calling thread ("" if none). */
const char* sinv_last_error(void);

/* ------------------------------------------------------------------------
 * RGF selected inversion, batched over instances.
 * Replaces rgf.rgf (src/sinv/algorithms/rgf/rgf.py:17-64): forward sweep
 * rgf.py:43-48, backward sweep rgf.py:50-62, `sym_mat` / `save_off_diag`
 * keyword semantics of rgf.py:18-20 (off-diagonals NaN when not saved,
 * rgf.py:41,55; plain transpose when sym_mat, rgf.py:58).
 *
 * Instance z has blocks
 *   A_ii      = dscale * diag[z][i] + shift[z] * I
 *   A_{i,i+1} = oscale * upper[z][i],  A_{i+1,i} = oscale * lower[z][i]
 * so A(E) = (E + i eta) I - H is dscale = oscale = -1, shift = E + i eta with
 * H shared (stride 0).  Plain input A is dscale = oscale = 1, shift = NULL.
 *
 * Block sizes: uniform `blocksize` when `bsizes` is NULL; otherwise bsizes
 * (host, nblocks entries) gives b_i and blocks are packed back to back:
 * diag block i at sum_{j<i} b_j^2, upper/lower block i (b_i x b_{i+1} /
 * b_{i+1} x b_i) at sum_{j<i} b_j b_{j+1}.  The *_stride fields are the
 * element distance between instances (0 = shared by all instances).
 * Outputs g_* are instance-major with the packed per-instance size.
 * ------------------------------------------------------------------------ */
typedef struct sinv_rgf_desc {
  int nblocks;
  int blocksize;
  const int* bsizes;
  int batch;
  int sym_mat;
  int save_off_diag;
  int precision;
  const void* diag;  long long diag_stride;
  const void* upper; long long upper_stride;
  const void* lower; long long lower_stride;
  sinv_c128 dscale;
  sinv_c128 oscale;
  const void* shift;
  void* g_diag;
  void* g_upper;
  void* g_lower;
  unsigned long long* info;
} sinv_rgf_desc;

size_t sinv_rgf_workspace_bytes(const sinv_rgf_desc* d);
int sinv_rgf_batched(const sinv_rgf_desc* d, void* work, size_t work_bytes, void* stream);

/* Allocation-free variant of sinv_rgf_batched.  No counterpart in the
 * reference (it allocates through numpy); it is for callers that cannot
 * allocate device memory: code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc.  Same descriptor, arithmetic and results
 * as sinv_rgf_batched; the workspace is a library-held device array of
 * SINV_NOALLOC_SCRATCH_BYTES on each GPU (held until the process ends), so a
 * call whose sinv_rgf_workspace_bytes exceeds that returns SINV_E_INVAL with
 * the sizes in sinv_last_error().  Nothing is allocated.  Every call on one
 * GPU shares that GPU's array: calls that overlap on a GPU (other streams or
 * threads) race and must be serialised by the caller. */
#define SINV_NOALLOC_SCRATCH_BYTES (64ull << 20)
int sinv_rgf_batched_noalloc(const sinv_rgf_desc* d, void* stream);

/* ------------------------------------------------------------------------
 * Adjoint of the RGF selected inversion, batched over instances.
 * No counterpart in the reference (it has no gradients); it differentiates
 * the map computed by rgf.py:40-62.  Given grad_G on the block-tridiagonal
 * pattern (torch's convention for a real loss: dL/dRe + i dL/dIm), returns
 * grad_A = -G^H grad_G G^H restricted to the pattern, for the same instance
 * convention as sinv_rgf_desc (uniform `blocksize`, or per-block sizes in
 * `bsizes` with the packed layout described there).  For
 * A(E) = (E + i eta) I - H the gradient of H is -grad_A.
 * checkpoint_every = k > 0 stores the left-connected blocks only every k
 * blocks and recomputes each k-block segment during the backward sweep
 * (0 = store all).  gbar_* strides may be 0 (one grad_G for every instance).
 * The sweep over A^H passes through G_ii^H, so a non-NULL g_diag receives
 * the selected inverse's diagonal blocks as well (a loss and its gradient
 * from one call, without a separate forward solve).
 * ------------------------------------------------------------------------ */
typedef struct sinv_rgf_adjoint_desc {
  int nblocks;
  int blocksize;
  const int* bsizes;
  int batch;
  int precision;
  int checkpoint_every;
  const void* diag;  long long diag_stride;
  const void* upper; long long upper_stride;
  const void* lower; long long lower_stride;
  sinv_c128 dscale;
  sinv_c128 oscale;
  const void* shift;
  const void* gbar_diag;  long long gbar_diag_stride;
  const void* gbar_upper; long long gbar_upper_stride;
  const void* gbar_lower; long long gbar_lower_stride;
  void* abar_diag;
  void* abar_upper;
  void* abar_lower;
  void* g_diag;  /* optional: also the diagonal blocks G_ii of the forward solve */
  unsigned long long* info;
} sinv_rgf_adjoint_desc;

size_t sinv_rgf_adjoint_workspace_bytes(const sinv_rgf_adjoint_desc* d);
int sinv_rgf_adjoint_batched(const sinv_rgf_adjoint_desc* d, void* work, size_t work_bytes,
                             void* stream);

/* ------------------------------------------------------------------------
 * BCR-serial selected inversion, batched over instances.
 * Replaces bcr_s.bcr_serial (src/sinv/algorithms/bcr/bcr_serial.py:19-62):
 * reduce_bcr/reduce (:173-216, :66-169), invert_block (:220-249),
 * produce_bcr with corner/center production (:387-466, :253-383).  The
 * identity padding of bcr_utils.py:35-71 is skipped (padded rows contribute
 * exact zeros).  Instance blocks as in sinv_rgf_desc (uniform block size,
 * A_ii = dscale*diag + shift*I, couplings = oscale*upper/lower).
 * Outputs: diagonal / upper / lower blocks of G; optionally (g_fill != NULL)
 * the other blocks the reference's dense result holds -- G between the
 * active rows r_m = m*2^l + 2^l - 1 of each level l >= 1, per instance
 * level by level: G[r_c, r_{c+1}] for c < n_l - 1, then G[r_{c+1}, r_c];
 * sinv_bcr_fill_blocks(N) blocks per instance.
 * ------------------------------------------------------------------------ */
typedef struct sinv_bcr_desc {
  int nblocks;
  int blocksize;
  int batch;
  const void* diag;  long long diag_stride;
  const void* upper; long long upper_stride;
  const void* lower; long long lower_stride;
  sinv_c128 dscale;
  sinv_c128 oscale;
  const void* shift;
  void* g_diag;
  void* g_upper;
  void* g_lower;
  void* g_fill;
  unsigned long long* info;
} sinv_bcr_desc;

size_t sinv_bcr_workspace_bytes(const sinv_bcr_desc* d);
long long sinv_bcr_fill_blocks(int nblocks);
int sinv_bcr_batched(const sinv_bcr_desc* d, void* work, size_t work_bytes, void* stream);
/* BCR-parallel's per-rank work on one GPU (bcr_parallel.py:21-78): the same
 * reduction and production, with every level operation issued as one launch
 * per partition over the rows that partition owns -- npart partitions of the
 * padded block rows split as bcr_utils.py:133-184 does among ranks (equal
 * strides, the last takes the rest); a partition reduces its kept rows
 * (bcr_parallel.py:227-256) and produces its rows (:1004-1015).  Each
 * operation family is issued for all partitions before the next, which is
 * where the ranks exchange halo rows.  Results equal sinv_bcr_batched's bit
 * for bit; same workspace.  SINV_E_INVAL unless 1 <= npart <= padded rows. */
int sinv_bcr_parallel_batched(const sinv_bcr_desc* d, int npart, void* work, size_t work_bytes,
                              void* stream);

/* ------------------------------------------------------------------------
 * PSR selected inversion on one GPU, batched over instances.
 * Replaces psr_s.psr_seqsolve (src/sinv/algorithms/psr/psr_seqsolve.py:20-72)
 * with `npart` partitions in place of MPI ranks: partition-local Schur
 * reduction (:127-309) and production (:693-898) as batched kernels over
 * partitions, the 2(npart-1)-block reduced system (:313-383, :470-491)
 * solved by RGF.  Partitioning as psr_utils.divide_matrix (:61-99); needs
 * npart >= 2 and nblocks >= 3 npart (psr_utils.check_input :44-55).
 * With npart == 2 and first_rows = N/2 this is RGF-2sided
 * (src/sinv/algorithms/rgf/rgf2sided.py:18-274): the two partitions' sweeps
 * are its half-sweeps and the 2-block reduced system its connection step; it
 * then needs only 1 <= first_rows < nblocks.
 * Instance convention as sinv_rgf_desc (uniform block size).  Outputs the
 * diagonal, upper and lower blocks of G.
 * ------------------------------------------------------------------------ */
typedef struct sinv_psr_desc {
  int nblocks;
  int blocksize;
  int batch;
  int npart;
  int first_rows;  /* 0: divide_matrix's split; else rows of partition 0 (npart == 2 only) */
  const void* diag;  long long diag_stride;
  const void* upper; long long upper_stride;
  const void* lower; long long lower_stride;
  sinv_c128 dscale;
  sinv_c128 oscale;
  const void* shift;
  void* g_diag;
  void* g_upper;
  void* g_lower;
  unsigned long long* info;
} sinv_psr_desc;

size_t sinv_psr_workspace_bytes(const sinv_psr_desc* d);
int sinv_psr_batched(const sinv_psr_desc* d, void* work, size_t work_bytes, void* stream);

/* ------------------------------------------------------------------------
 * PDIV (P-Division) selected inversion with npart partitions on one device.
 * Replaces pdiv_localmap.py:27-90 (pdiv_localmap on every rank) with the
 * partitioning of pdiv_utils.py:84-170: npart must be a power of 2 and at
 * most nblocks; partition 0 takes the remainder N - m (npart - 1),
 * m = N // npart.  Each partition is inverted densely (invert_partition,
 * pdiv_localmap.py:94-115) and its selected blocks produced from that inverse
 * (produce_matrix_elements :1194-1241), batched over partitions and
 * instances; the coupling between partitions (the reference's map updates)
 * is a Woodbury correction from two sweeps of b x b boundary blocks.
 * Outputs: the block-tridiagonal part of A^-1, diag (batch,N,b,b),
 * upper/lower (batch,N-1,b,b), contiguous.  Strides of 0 share the input.
 * ------------------------------------------------------------------------ */
typedef struct sinv_pdiv_desc {
  int nblocks;
  int blocksize;
  int batch;
  int npart;
  const void* diag;  long long diag_stride;
  const void* upper; long long upper_stride;
  const void* lower; long long lower_stride;
  void* g_diag;
  void* g_upper;
  void* g_lower;
  unsigned long long* info;
} sinv_pdiv_desc;

size_t sinv_pdiv_workspace_bytes(const sinv_pdiv_desc* d);
int sinv_pdiv_batched(const sinv_pdiv_desc* d, void* work, size_t work_bytes, void* stream);

/* ------------------------------------------------------------------------
 * Batched in-place complex128 block inverse (kernel K1).
 * Replaces the numpy.linalg.inv calls of the hot path (rgf.py:44,48;
 * bcr_serial.py:121,148,249,302,379).  Gauss-Jordan elimination with partial
 * pivoting over all n rows on |re| + |im| (the modulus of LAPACK's izamax,
 * which numpy.linalg.inv's zgetrf uses): one workgroup per block for n <= 64;
 * above, blocked by columns, tall panels eliminated over all rows by one
 * workgroup each and their effect on the other columns applied by batched
 * FP64 MFMA products (n <= 8192; the workspace holds one n x n copy per
 * block plus row positions and pivots).
 * Matrix z is at A + z*strideA (complex elements), leading dimension lda.
 * ------------------------------------------------------------------------ */
size_t sinv_zinv_workspace_bytes(int n, int batch);
int sinv_zinv_batched(void* A, int n, int lda, long long strideA, int batch, void* work,
                      size_t work_bytes, unsigned long long* info, void* stream);

/* Leaf kernel selection (no counterpart in the reference; for tests and
 * measurements).  SINV_LEAF_AUTO: the panel kernel with flag hand-off
 * between waves, 16 waves of 4 columns for launches of fewer than 512
 * blocks, else 8 waves of 8 columns (pivot: the earliest row within 2^-15 of
 * the column maximum); SINV_LEAF_PANEL16 / PANEL8: the same with one barrier
 * per panel.  SINV_LEAF_ROWS8: one barrier per elimination step, pivot =
 * LAPACK izamax's exact first maximum.  SINV_LEAF_LDS: the LDS-resident
 * kernel (exact first maximum).  Process-wide, unsynchronised: set it before
 * launching the work it applies to.  The environment variable
 * SINV_LEAF_KERNEL (panel16f, panel8f, panel16, panel8, rows8, lds) sets the
 * initial choice. */
#define SINV_LEAF_AUTO 1
#define SINV_LEAF_LDS 0
#define SINV_LEAF_ROWS8 2
#define SINV_LEAF_PANEL8 3
#define SINV_LEAF_PANEL16 4
#define SINV_LEAF_PANEL16_FLAGS 5
#define SINV_LEAF_PANEL8_FLAGS 6
int sinv_leaf_force_variant(int variant);
/* Tall-panel kernel of the blocked inverse for blocks above 64 (no
 * counterpart in the reference; for tests and measurements).  SINV_GJ_FLOW
 * (default): steps handed between waves through LDS flags, the next pivot
 * column updated and searched first, pivot rows scaled at write-back (the
 * step kernel serves blocks above 1024).  SINV_GJ_STEP: one barrier per
 * eliminated column.  Both use panels of 64 columns up to 384 rows, 32 up to
 * 512, 16 up to 896, 8 above.  SINV_GJ_PANEL: one barrier per wave's group
 * of columns, the owner factoring them alone, panels of 64 columns up to 128
 * rows, 32 up to 256, 16 up to 512, 8 above.  All pivot identically (the
 * exact first maximum of |re| + |im| in position order) and agree to
 * rounding.  Process-wide, unsynchronised; SINV_GJ_PANEL=flow / step / panel
 * in the environment sets the initial choice. */
#define SINV_GJ_STEP 0
#define SINV_GJ_PANEL 1
#define SINV_GJ_FLOW 2
int sinv_gj_force_variant(int variant);
/* Bound on a leaf or flow tall-panel wave's wait for a hand-off (no counterpart in the
 * reference; for failure-handling tests).  polls = -1 restores the default
 * (2^22); a wave that exhausts it stops and its block reports as singular
 * through info, so the kernel always exits. */
int sinv_leaf_spin_limit(int polls);

/* Inverse method of the calling thread (no counterpart in the reference:
 * numpy.linalg.inv always eliminates).  SINV_INV_PIVOTED (default): the
 * elimination above.  SINV_INV_NEWTON: launches of blocks above 64 with at
 * most 4 instances (the latency-bound case, where the elimination's
 * column-serial pivot chain occupies one CU) run the Newton-Schulz
 * iteration X <- X (2I - M X) from X = 2 D^-1 - D^-1 M D^-1 (D = diag(M))
 * on the FP64 GEMM instead.  An instance is accepted once
 * max|I - M X_k| <= 1e-9, which makes the next iterate's residual at most
 * (n 1e-9)^2 in norm; one whose residual stays above that after max_iter
 * products (1 .. 16; 0 keeps the current limit, initially 6) is reported
 * through info with SINV_INFO_NOT_CONVERGED set in the key, and its output
 * is undefined: repeat the call with SINV_INV_PIVOTED.  The iteration
 * converges whenever every inverted block is strictly column diagonally
 * dominant (sum_{i != j} |m_ij| < |m_jj|, which every block RGF, BCR, PSR
 * and PDIV invert inherits from a matrix with that property; see
 * sinv_column_dominance), at a rate set by the dominance ratio.  Returns
 * the previous method, or -1 for an unknown method. */
#define SINV_INV_PIVOTED 0
#define SINV_INV_NEWTON 1
#define SINV_INFO_NOT_CONVERGED (1ULL << 62)
int sinv_inverse_method(int method, int max_iter);

/* Column dominance ratio of block-tridiagonal instances described as for
 * sinv_rgf_batched with a uniform block size (diag, upper, lower, strides,
 * batch, dscale, oscale, shift; bsizes must be NULL and the output blocks
 * are not used): rho[z] =
 * max_j sum_{i != j} |a_ij| / |a_jj| over the whole matrix of instance z
 * (device array of batch doubles; +inf where a diagonal entry is zero).
 * rho < 1 means strictly column diagonally dominant, so every Schur
 * complement the selected-inversion methods invert is too, with a ratio no
 * larger (no counterpart in the reference; it decides when
 * SINV_INV_NEWTON applies). */
int sinv_column_dominance(const sinv_rgf_desc* d, double* rho, void* stream);

/* ------------------------------------------------------------------------
 * Batched complex GEMM (kernel K2):
 *   C[z] = alpha * op(A[z]) op(B[z]) + beta * C0[z] (+ shift[z] * I)
 * Replaces the ndarray `@` products of the hot path (rgf.py:48,53,60,62;
 * bcr_serial.py:124-169, 293-304, 354-383).  precision = SINV_PREC_FP64
 * (FP64 MFMA) or SINV_PREC_MIXED (operands rounded to complex64, FP32 MFMA,
 * FP32 accumulation).  C0 may be NULL (beta ignored) or alias C.
 * ------------------------------------------------------------------------ */
int sinv_zgemm_batched(int M, int N, int K, int opA, int opB, sinv_c128 alpha,
                       const void* A, int lda, long long strideA,
                       const void* B, int ldb, long long strideB,
                       sinv_c128 beta, const void* C0, int ldc0, long long strideC0,
                       void* C, int ldc, long long strideC,
                       const void* shift, int batch, int precision, void* stream);

/* ------------------------------------------------------------------------
 * Complex128 GEMM from int8 digit planes (no counterpart in the reference,
 * whose `@` is FP64 BLAS; it stands in for the same products, e.g.
 * rgf.py:48,53,60,62).  Same contract as sinv_zgemm_batched.  Each row of
 * op(A) and column of op(B) is scaled by a power of two and split into
 * `slices` (5..7) signed 7-bit digits per real part (re, im, re + im); digit
 * products are summed exactly in int32 on the int8 matrix cores and pairs
 * with digit index sum >= slices are dropped, so one real product errs by at
 * most about 2^(1 - 7 slices) K max|row of op(A)| max|column of op(B)|.
 * K <= 16384.  The workspace (size from sinv_zgemm_int8x_workspace_bytes with
 * the same strides; stride 0 = one operand shared by the batch, sliced once)
 * receives the digit planes; nothing is allocated.
 * ------------------------------------------------------------------------ */
size_t sinv_zgemm_int8x_workspace_bytes(int M, int N, int K, long long strideA,
                                        long long strideB, int batch, int slices);
int sinv_zgemm_int8x_batched(int M, int N, int K, int opA, int opB, sinv_c128 alpha,
                             const void* A, int lda, long long strideA,
                             const void* B, int ldb, long long strideB,
                             sinv_c128 beta, const void* C0, int ldc0, long long strideC0,
                             void* C, int ldc, long long strideC,
                             const void* shift, int batch, int slices,
                             void* work, size_t work_bytes, void* stream);

/* Product kernel of the int8x GEMM (no counterpart in the reference): 0 =
 * two LDS buffers, one workgroup per CU; 1 = one LDS buffer, two workgroups
 * per CU, 80-byte padded LDS rows; 2 = the same with swizzled 64-byte rows
 * (default); -1 restores the default.  All give identical bits (the digit
 * sums are exact and the fold order is the same).  Process-wide. */
int sinv_int8x_force_variant(int variant);

/* ------------------------------------------------------------------------
 * Dense <-> block-tridiagonal conversion on the device.
 * Replaces matrices_utils.py:115-151 (convertDenseToBlkTridiag) and
 * matrices_utils.py:155-189 (convertBlkTridiagToDense) for matrices resident
 * in HBM; complex128, row-major, `batch` matrices at stride strideA.
 *  - sinv_dense_to_blocks: n must be a multiple of b (the reference's ragged
 *    last block raises numpy's broadcast error); writes diag (batch,N,b,b),
 *    upper/lower (batch,N-1,b,b), contiguous.
 *  - sinv_blocks_to_dense: writes the whole (N b) x (N b) matrix: band blocks
 *    from the stacked arrays, zeros elsewhere.
 *  - sinv_scatter_blocks: copies `count` b x b blocks (stride strideS between
 *    batch entries, contiguous within) to the block positions (rows[k],
 *    cols[k]) given in DEVICE int arrays; positions outside the nblocks x
 *    nblocks grid are skipped.
 * ------------------------------------------------------------------------ */
int sinv_dense_to_blocks(const void* A, int n, int lda, long long strideA, int b, int batch,
                         void* diag, void* upper, void* lower, void* stream);
int sinv_blocks_to_dense(const void* diag, const void* upper, const void* lower, int nblocks,
                         int b, int batch, void* A, int lda, long long strideA, void* stream);
int sinv_scatter_blocks(const void* blocks, long long strideS, int count, const int* rows,
                        const int* cols, int nblocks, int b, int batch, void* A, int lda,
                        long long strideA, void* stream);

/* ------------------------------------------------------------------------
 * GEMM tile configurations and tuned shapes (no counterpart in the reference:
 * numpy's BLAS picks its own blocking).  Every complex GEMM the library issues
 * (sinv_zgemm_batched and the ones inside the RGF, BCR, PSR, adjoint and
 * inverse drivers) runs one of sinv_gemm_config_count() tile configurations;
 * sinv_gemm_config_tile gives its workgroup tile (bm x bn), and
 * sinv_gemm_config_ksplit the number of k-slices its workgroup sums (1: one
 * sequential pass over k).  Configs with the same k-split accumulate every
 * output in the same order and give identical bits; configs with different
 * k-splits agree to rounding (-SINV_E_INVAL for a bad id).  The config is
 * chosen per launch: a forced config if set (sinv_gemm_force_config, -1
 * clears), else the tuned entry of the exact shape (M, N, K, batch, opA, opB,
 * C0 present, precision) if one was set with sinv_gemm_tune_set, else a fill
 * heuristic.  sinv_gemm_select returns the config a launch of that shape
 * would use.  The table is process-wide and unsynchronised: set it before
 * launching the work it applies to, not concurrently with launches.  The
 * autotuner that fills it is sinv_amd.autotune.
 * ------------------------------------------------------------------------ */
int sinv_gemm_config_count(void);
int sinv_gemm_config_tile(int cfg, int* bm, int* bn);
int sinv_gemm_config_ksplit(int cfg);
int sinv_gemm_force_config(int cfg);
int sinv_gemm_tune_set(int M, int N, int K, long long batch, int opA, int opB, int has_c0,
                       int precision, int cfg);
int sinv_gemm_tune_clear(void);
int sinv_gemm_select(int M, int N, int K, long long batch, int opA, int opB, int has_c0,
                     int precision);

/* ------------------------------------------------------------------------
 * Launch profiler (no counterpart in the reference, which has no timing code).
 * Between sinv_prof_begin and sinv_prof_end every kernel launch of the GEMM
 * and leaf-inverse kernels is bracketed by hipEvents on its own stream.
 * sinv_prof_end synchronises those events and returns, per kind
 * (0 = FP64 GEMM, 1 = FP32 GEMM, 2 = leaf inverse, 3 = other, 4 = int8x
 * product kernel, 5 = tall-panel elimination of the blocked inverse,
 * 6 = Newton-Schulz products, 7 = int8x digit split), the summed kernel time
 * in ms, the launch count, the algorithmic flops (8 M N K per complex GEMM,
 * int8x product included; 8 n^3 per inverse) and, when `bytes` is not NULL,
 * the algorithmic bytes (16 B per complex element of each operand read and
 * each result written once: M K + K N + M N, + M N more with a C0 term; 2 n^2
 * per in-place inverse; int8x: the split reads each distinct operand copy once
 * and writes 3 S B of digits per padded element, the product reads those
 * digits once and writes C).  Process-wide state: use from one thread, around a
 * region with at most `max_launches` launches.
 * ------------------------------------------------------------------------ */
int sinv_prof_begin(int max_launches);
int sinv_prof_end(double* ms, long long* launches, double* flops, double* bytes, int nkinds);
/* Per-launch records of the last region closed by sinv_prof_end, in launch
 * order: kind, 5 dims (GEMM: M, N, K, batch, opA | opB << 2 | (C0 != NULL) << 4;
 * zeros for other kinds) and the event-bracketed time in ms.  Copies at most
 * `cap` records and returns how many there are. */
int sinv_prof_records(int cap, int* kind, long long* dims, float* ms);

#ifdef __cplusplus
}
#endif

#endif /* SINV_H_ */
