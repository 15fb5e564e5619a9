// PSR (parallel Schur reduction) selected inversion on one GPU: the
// partition-local work of src/sinv/algorithms/psr/psr_seqsolve.py run as
// batched kernels, partitions playing the part of the reference's ranks.
//
//  * divide_matrix (psr_utils.py:61-99): m = N // P blocks per partition, the
//    first partition takes the remainder.
//  * reduction (psr_seqsolve.py:127-309): the top partition eliminates
//    downwards into its last row, the bottom one upwards into its first row,
//    and every central partition eliminates its interior rows into its last
//    row while carrying the fill-in coupling to its first row.  Central
//    partitions have the same shape, so each elimination step is ONE batched
//    launch over all of them (and over all instances).
//  * the 2(P-1) boundary rows form a block-tridiagonal reduced system
//    (aggregate_reduced_system :313-383); its selected inverse comes from
//    this library's RGF (the reference inverts it densely, :489 -- only its
//    selected blocks are used by the production).
//  * production (psr_seqsolve.py:693-898), batched over central partitions.
//
// The only data exchanged between partitions is the reduced system: on P
// GPUs this is the single all-gather of boundary rows (one rank's share of
// this driver's work is one partition).
#include <vector>

#include "sinv_core.hpp"

using namespace sinv;

namespace {

// info <- min(info, reduced-system info): keeps the first singular pivot of
// either stage
__global__ void merge_info_kernel(unsigned long long* dst, const unsigned long long* src) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    const unsigned long long v = *src;
    if (v < *dst) *dst = v;
  }
}

bool psr_desc_ok(const sinv_psr_desc* d) {
  if (d == nullptr || d->nblocks <= 0 || d->blocksize <= 0 || d->batch <= 0) return false;
  if (d->npart < 2) return false;
  if (d->first_rows != 0) {
    // explicit split: two partitions of at least one row each (RGF-2sided)
    if (d->npart != 2 || d->first_rows < 1 || d->first_rows >= d->nblocks) return false;
  } else if (d->nblocks < 3 * d->npart) {
    return false;
  }
  if (d->diag == nullptr || d->upper == nullptr || d->lower == nullptr) return false;
  if (d->g_diag == nullptr || d->g_upper == nullptr || d->g_lower == nullptr) return false;
  if (d->diag_stride < 0 || d->upper_stride < 0 || d->lower_stride < 0) return false;
  return true;
}

struct PsrPlan {
  int N = 0, b = 0, P = 0, m = 0, n0 = 0, R = 0;
  std::vector<int> red;  // boundary rows of the reduced system
  // per-instance arrays (block offsets), each N blocks unless noted
  long long Ad = 0, invA = 0, Lf = 0, Uf = 0, Ls = 0, Us = 0, Afc = 0, Afr = 0, Gsr = 0, Gsc = 0;
  long long Rd = 0, Ru = 0, Rl = 0;  // reduced system (R, R-1, R-1)
  long long per_inst = 0;
  size_t inv_elems = 0;       // inverse workspace of the central partitions
  size_t inv_side_elems = 0;  // each of the top and bottom partitions
  size_t rgf_bytes = 0;
  size_t bytes = 0;
};

PsrPlan psr_plan(int N, int b, int P, int batch, int first_rows) {
  PsrPlan Q;
  Q.N = N; Q.b = b; Q.P = P;
  Q.m = N / P;
  Q.n0 = first_rows > 0 ? first_rows : N - Q.m * (P - 1);
  Q.red.push_back(Q.n0 - 1);
  for (int p = 1; p <= P - 2; ++p) {
    const int s = Q.n0 + (p - 1) * Q.m;
    Q.red.push_back(s);
    Q.red.push_back(s + Q.m - 1);
  }
  Q.red.push_back(Q.n0 + (P - 2) * Q.m);
  Q.R = (int)Q.red.size();
  long long o = 0;
  auto take = [&](long long n) { const long long r = o; o += n; return r; };
  Q.Ad = take(N); Q.invA = take(N); Q.Lf = take(N); Q.Uf = take(N); Q.Ls = take(N);
  Q.Us = take(N); Q.Afc = take(N); Q.Afr = take(N); Q.Gsr = take(N); Q.Gsc = take(N);
  Q.Rd = take(Q.R); Q.Ru = take(Q.R - 1); Q.Rl = take(Q.R - 1);
  Q.per_inst = o;
  const int maxc = P - 2 > 1 ? P - 2 : 1;
  const bool newton = inverse_method() == SINV_INV_NEWTON;
  Q.inv_elems = inverse_workspace_elems(b, batch * maxc, newton);
  Q.inv_side_elems = inverse_workspace_elems(b, batch, newton);
  sinv_rgf_desc rd{};
  rd.nblocks = Q.R; rd.blocksize = b; rd.batch = batch; rd.precision = SINV_PREC_FP64;
  rd.save_off_diag = 1;
  rd.diag = rd.upper = rd.lower = (const void*)16;
  rd.g_diag = rd.g_upper = rd.g_lower = (void*)16;
  rd.dscale = sinv_c128{1.0, 0.0};
  rd.oscale = sinv_c128{1.0, 0.0};
  Q.rgf_bytes = sinv_rgf_workspace_bytes(&rd);
  const size_t b2 = (size_t)b * b;
  const size_t red_out = (size_t)(3 * Q.R - 2) * b2 * batch;
  Q.bytes = ((size_t)Q.per_inst * b2 * batch + red_out + Q.inv_elems + 2 * Q.inv_side_elems + 1) *
                sizeof(cplx) +
            Q.rgf_bytes + 1024;
  return Q;
}

}  // namespace

extern "C" size_t sinv_psr_workspace_bytes(const sinv_psr_desc* d) {
  if (!psr_desc_ok(d)) return 0;
  return psr_plan(d->nblocks, d->blocksize, d->npart, d->batch, d->first_rows).bytes;
}

extern "C" int sinv_psr_batched(const sinv_psr_desc* d, void* work, size_t work_bytes,
                                void* stream) {
  if (!psr_desc_ok(d)) {
    set_error("sinv_psr_batched: invalid descriptor (need npart >= 2 and nblocks >= 3 npart, "
              "or npart == 2 with 1 <= first_rows < nblocks)");
    return SINV_E_INVAL;
  }
  const int N = d->nblocks, b = d->blocksize, nb = d->batch, P = d->npart;
  const PsrPlan Q = psr_plan(N, b, P, nb, d->first_rows);
  if (work == nullptr || work_bytes < Q.bytes) {
    set_error("sinv_psr_batched: workspace too small (%zu < %zu)", work_bytes, Q.bytes);
    return SINV_E_INVAL;
  }
  hipStream_t s = (hipStream_t)stream;
  const long long b2 = (long long)b * b;
  cplx* base = (cplx*)(((uintptr_t)work + 255) & ~(uintptr_t)255);
  const long long WI = Q.per_inst * b2;
  cplx* W = base;
  cplx* GRd = W + WI * nb;                         // reduced outputs, instance-major
  cplx* GRu = GRd + (long long)Q.R * b2 * nb;
  cplx* GRl = GRu + (long long)(Q.R - 1) * b2 * nb;
  cplx* winv = GRl + (long long)(Q.R - 1) * b2 * nb;
  cplx* winv_top = winv + Q.inv_elems;
  cplx* winv_bot = winv_top + Q.inv_side_elems;
  unsigned long long* rinfo = (unsigned long long*)(winv_bot + Q.inv_side_elems);
  void* wrgf = (void*)(((uintptr_t)(rinfo + 1) + 255) & ~(uintptr_t)255);

  const cplx one{1.0, 0.0}, mone{-1.0, 0.0}, zero{0.0, 0.0};
  const cplx osc{d->oscale.re, d->oscale.im};
  const cplx mosc{-osc.re, -osc.im};
  cplx* Gd = (cplx*)d->g_diag;
  cplx* Gu = (cplx*)d->g_upper;
  cplx* Gl = (cplx*)d->g_lower;
  const long long GdI = (long long)N * b2, GoI = (long long)(N - 1) * b2;

  struct Arr {
    const cplx* p;
    long long inst;
  };
  auto warr = [&](long long blk) { return Arr{W + blk * b2, WI}; };
  const Arr Ad = warr(Q.Ad), invA = warr(Q.invA), Lf = warr(Q.Lf), Uf = warr(Q.Uf);
  const Arr Ls = warr(Q.Ls), Us = warr(Q.Us), Afc = warr(Q.Afc), Afr = warr(Q.Afr);
  const Arr Gsr = warr(Q.Gsr), Gsc = warr(Q.Gsc);
  const Arr Up{(const cplx*)d->upper, d->upper_stride}, Lo{(const cplx*)d->lower, d->lower_stride};
  const Arr GD{Gd, GdI}, GU{Gu, GoI}, GL{Gl, GoI};

  // The top, bottom and central partitions are independent between the
  // exchanges: each runs on its own stream (cur) with its own inverse
  // workspace (wcur), as the reference's ranks run concurrently.
  hipStream_t cur = s;
  cplx* wcur = winv;
  const bool newton = inverse_method() == SINV_INV_NEWTON;
  // C[k] (k < count; block index c0 + k*cs) = alpha A[a0 + k*as] B[b0 + k*bs] (+ C0[...])
  auto gemm = [&](int count, Arr A, long long a0, long long as, Arr B, long long b0,
                  long long bs, Arr C, long long c0, long long cs, cplx alpha, bool acc,
                  Arr C0 = Arr{nullptr, 0}, long long c00 = 0, long long c0s = 0) -> hipError_t {
    if (count <= 0) return hipSuccess;
    GemmArgs g{};
    g.M = b; g.N = b; g.K = b; g.opA = OP_N; g.opB = OP_N;
    g.A = A.p + a0 * b2; g.lda = b; g.sA_in = as * b2; g.sA_out = A.inst;
    g.B = B.p + b0 * b2; g.ldb = b; g.sB_in = bs * b2; g.sB_out = B.inst;
    g.C = const_cast<cplx*>(C.p) + c0 * b2; g.ldc = b; g.sC_in = cs * b2; g.sC_out = C.inst;
    if (acc) {
      const Arr& S = C0.p ? C0 : C;
      const long long s0 = C0.p ? c00 : c0, ss = C0.p ? c0s : cs;
      g.C0 = S.p + s0 * b2; g.ldc0 = b; g.sC0_in = ss * b2; g.sC0_out = S.inst;
      g.beta = one;
    } else {
      g.C0 = nullptr; g.beta = zero;
    }
    g.alpha = alpha;
    g.shift = nullptr;
    g.inner = count; g.outer = nb;
    return launch_gemm(g, Prec::F64, cur);
  };
  auto copy = [&](int count, Arr S, long long s0, long long ss, Arr C, long long c0, long long cs,
                  cplx a) {
    if (count <= 0) return hipSuccess;
    return launch_axpby_shift(b, b, a, S.p + s0 * b2, b, ss * b2, S.inst,
                              const_cast<cplx*>(C.p) + c0 * b2, b, cs * b2, C.inst, nullptr,
                              count, nb, cur);
  };
  // invA[r + k*st] = inv(Ad[r + k*st])
  auto invert = [&](int count, long long r, long long st) -> hipError_t {
    SINV_HIP_RET(copy(count, Ad, r, st, invA, r, st, one));
    return launch_inverse(W + (Q.invA + r) * b2, b, b, st * b2, WI, count, nb, wcur, d->info,
                          (int)r, cur, newton);
  };

  if (d->info != nullptr) SINV_HIP_TRY(launch_fill_u64(d->info, 1, SINV_INFO_OK, s));
  // Ad = dscale D + shift I
  {
    const cplx ds{d->dscale.re, d->dscale.im};
    SINV_HIP_TRY(launch_axpby_shift(b, b, ds, (const cplx*)d->diag, b, b2, d->diag_stride,
                                    W + Q.Ad * b2, b, b2, WI, (const cplx*)d->shift, N, nb, s));
  }
  const int m = Q.m, n0 = Q.n0, C = P - 2;
  const int s1 = n0;                       // first central partition
  const int sB = n0 + (P - 2) * m;         // bottom partition
  const int eB = N - 1;

  // ---------------------------------------------------------------- reduction
  StreamFork red_fork(s, 2);
  SINV_HIP_TRY(red_fork.status());
  // top partition, downwards (psr_seqsolve.py:157-167)
  cur = red_fork.side(0);
  wcur = winv_top;
  for (int i = 1; i < n0; ++i) {
    SINV_HIP_TRY(invert(1, i - 1, 0));
    SINV_HIP_TRY(gemm(1, Lo, i - 1, 0, invA, i - 1, 0, Lf, i, 0, osc, false));
    SINV_HIP_TRY(gemm(1, invA, i - 1, 0, Up, i - 1, 0, Uf, i, 0, osc, false));
    SINV_HIP_TRY(gemm(1, Lf, i, 0, Up, i - 1, 0, Ad, i, 0, mosc, true));
  }
  // bottom partition, upwards (psr_seqsolve.py:203-216)
  cur = red_fork.side(1);
  wcur = winv_bot;
  for (int i = eB - 1; i >= sB; --i) {
    SINV_HIP_TRY(invert(1, i + 1, 0));
    SINV_HIP_TRY(gemm(1, Up, i, 0, invA, i + 1, 0, Lf, i, 0, osc, false));
    SINV_HIP_TRY(gemm(1, invA, i + 1, 0, Lo, i, 0, Uf, i, 0, osc, false));
    SINV_HIP_TRY(gemm(1, Lf, i, 0, Lo, i, 0, Ad, i, 0, mosc, true));
  }
  // central partitions, all at once (psr_seqsolve.py:260-307); local row t, r = t - 1
  cur = s;
  wcur = winv;
  for (int t = 2; t < m; ++t) {
    const long long i = s1 + t, r = s1 + t - 1, sr = s1;
    SINV_HIP_TRY(invert(C, r, m));
    // L_{i,r} = A_{i,r} inv ; U_{r,i} = inv A_{r,i}
    SINV_HIP_TRY(gemm(C, Lo, r, m, invA, r, m, Lf, i, m, osc, false));
    SINV_HIP_TRY(gemm(C, invA, r, m, Up, r, m, Uf, i, m, osc, false));
    // L_{s,r} = A_{s,r} inv ; U_{r,s} = inv A_{r,s}  (original couplings at t = 2, fill after)
    if (t == 2) {
      SINV_HIP_TRY(gemm(C, Up, sr, m, invA, r, m, Ls, r, m, osc, false));
      SINV_HIP_TRY(gemm(C, invA, r, m, Lo, sr, m, Us, r, m, osc, false));
    } else {
      SINV_HIP_TRY(gemm(C, Afr, r, m, invA, r, m, Ls, r, m, one, false));
      SINV_HIP_TRY(gemm(C, invA, r, m, Afc, r, m, Us, r, m, one, false));
    }
    // A_ii -= L_{i,r} A_{r,i}
    SINV_HIP_TRY(gemm(C, Lf, i, m, Up, r, m, Ad, i, m, mosc, true));
    // A_ss -= L_{s,r} A_{r,s};  A_{i,s} = -L_{i,r} A_{r,s}  (A_{r,s}: original at t = 2)
    if (t == 2) {
      SINV_HIP_TRY(gemm(C, Ls, r, m, Lo, sr, m, Ad, sr, m, mosc, true));
      SINV_HIP_TRY(gemm(C, Lf, i, m, Lo, sr, m, Afc, i, m, mosc, false));
    } else {
      SINV_HIP_TRY(gemm(C, Ls, r, m, Afc, r, m, Ad, sr, m, mone, true));
      SINV_HIP_TRY(gemm(C, Lf, i, m, Afc, r, m, Afc, i, m, mone, false));
    }
    // A_{s,i} = -L_{s,r} A_{r,i}
    SINV_HIP_TRY(gemm(C, Ls, r, m, Up, r, m, Afr, i, m, mosc, false));
  }

  SINV_HIP_TRY(red_fork.join());

  // ---------------------------------------------------------------- reduced system
  const Arr Rd = warr(Q.Rd), Ru = warr(Q.Ru), Rl = warr(Q.Rl);
  for (int k = 0; k < Q.R; ++k) SINV_HIP_TRY(copy(1, Ad, Q.red[k], 0, Rd, k, 0, one));
  for (int k = 0; k + 1 < Q.R; ++k) {
    const int r = Q.red[k];
    if (k % 2 == 0) {  // e_p -- s_{p+1}: original couplings
      SINV_HIP_TRY(copy(1, Up, r, 0, Ru, k, 0, osc));
      SINV_HIP_TRY(copy(1, Lo, r, 0, Rl, k, 0, osc));
    } else {           // s_p -- e_p of a central partition: fill-in couplings
      const int e = Q.red[k + 1];
      SINV_HIP_TRY(copy(1, Afr, e, 0, Ru, k, 0, one));
      SINV_HIP_TRY(copy(1, Afc, e, 0, Rl, k, 0, one));
    }
  }
  {
    sinv_rgf_desc rd{};
    rd.nblocks = Q.R; rd.blocksize = b; rd.bsizes = nullptr; rd.batch = nb;
    rd.sym_mat = 0; rd.save_off_diag = 1; rd.precision = SINV_PREC_FP64;
    rd.diag = Rd.p; rd.diag_stride = WI;
    rd.upper = Ru.p; rd.upper_stride = WI;
    rd.lower = Rl.p; rd.lower_stride = WI;
    rd.dscale = sinv_c128{1.0, 0.0};
    rd.oscale = sinv_c128{1.0, 0.0};
    rd.shift = nullptr;
    rd.g_diag = GRd; rd.g_upper = GRu; rd.g_lower = GRl;
    rd.info = rinfo;
    const int rc = sinv_rgf_batched(&rd, wrgf, Q.rgf_bytes, stream);
    if (rc != SINV_OK) return rc;
    if (d->info != nullptr) {
      hipLaunchKernelGGL(merge_info_kernel, dim3(1), dim3(64), 0, s, d->info, rinfo);
      SINV_HIP_TRY(hipGetLastError());
    }
  }
  const Arr GRD{GRd, (long long)Q.R * b2}, GRU{GRu, (long long)(Q.R - 1) * b2},
      GRL{GRl, (long long)(Q.R - 1) * b2};
  for (int k = 0; k < Q.R; ++k) SINV_HIP_TRY(copy(1, GRD, k, 0, GD, Q.red[k], 0, one));
  for (int k = 0; k + 1 < Q.R; ++k) {
    const int r = Q.red[k];
    if (k % 2 == 0) {
      SINV_HIP_TRY(copy(1, GRU, k, 0, GU, r, 0, one));
      SINV_HIP_TRY(copy(1, GRL, k, 0, GL, r, 0, one));
    } else {  // G_{s,e}, G_{e,s} of a central partition, kept at row e of Gsr / Gsc
      const int e = Q.red[k + 1];
      SINV_HIP_TRY(copy(1, GRU, k, 0, Gsr, e, 0, one));
      SINV_HIP_TRY(copy(1, GRL, k, 0, Gsc, e, 0, one));
    }
  }

  // ---------------------------------------------------------------- production
  StreamFork prod_fork(s, 2);
  SINV_HIP_TRY(prod_fork.status());
  // top partition, upwards (psr_seqsolve.py:731-740)
  cur = prod_fork.side(0);
  for (int i = n0 - 1; i >= 1; --i) {
    SINV_HIP_TRY(gemm(1, GD, i, 0, Lf, i, 0, GL, i - 1, 0, mone, false));
    SINV_HIP_TRY(gemm(1, Uf, i, 0, GD, i, 0, GU, i - 1, 0, mone, false));
    SINV_HIP_TRY(gemm(1, Uf, i, 0, GL, i - 1, 0, GD, i - 1, 0, mone, true, invA, i - 1, 0));
  }
  // bottom partition, downwards (psr_seqsolve.py:777-786)
  cur = prod_fork.side(1);
  for (int i = sB; i < eB; ++i) {
    SINV_HIP_TRY(gemm(1, GD, i, 0, Lf, i, 0, GU, i, 0, mone, false));
    SINV_HIP_TRY(gemm(1, Uf, i, 0, GD, i, 0, GL, i, 0, mone, false));
    SINV_HIP_TRY(gemm(1, Uf, i, 0, GU, i, 0, GD, i + 1, 0, mone, true, invA, i + 1, 0));
  }
  // central partitions (psr_seqsolve.py:823-898); s = s1 + q m, e = s + m - 1
  cur = s;
  if (C > 0) {
    const long long sr = s1, e = s1 + m - 1;
    // G_{e,e-1} = -G_{e,s} L_{s,e-1} - G_ee L_{e,e-1}
    SINV_HIP_TRY(gemm(C, Gsc, e, m, Ls, e - 1, m, GL, e - 1, m, mone, false));
    SINV_HIP_TRY(gemm(C, GD, e, m, Lf, e, m, GL, e - 1, m, mone, true));
    // G_{e-1,e} = -U_{e-1,e} G_ee - U_{e-1,s} G_{s,e}
    SINV_HIP_TRY(gemm(C, Uf, e, m, GD, e, m, GU, e - 1, m, mone, false));
    SINV_HIP_TRY(gemm(C, Us, e - 1, m, Gsr, e, m, GU, e - 1, m, mone, true));
    for (long long i = e - 1; i > sr; --i) {
      // G_{s,i} = -G_ss L_{s,i} - G_{s,i+1} L_{i+1,i}
      SINV_HIP_TRY(gemm(C, GD, sr, m, Ls, i, m, Gsr, i, m, mone, false));
      SINV_HIP_TRY(gemm(C, Gsr, i + 1, m, Lf, i + 1, m, Gsr, i, m, mone, true));
      // G_{i,s} = -U_{i,i+1} G_{i+1,s} - U_{i,s} G_ss
      SINV_HIP_TRY(gemm(C, Uf, i + 1, m, Gsc, i + 1, m, Gsc, i, m, mone, false));
      SINV_HIP_TRY(gemm(C, Us, i, m, GD, sr, m, Gsc, i, m, mone, true));
    }
    for (long long i = e - 1; i > sr + 1; --i) {
      // G_ii = inv(A_ii) - U_{i,s} G_{s,i} - U_{i,i+1} G_{i+1,i}
      SINV_HIP_TRY(gemm(C, Us, i, m, Gsr, i, m, GD, i, m, mone, true, invA, i, m));
      SINV_HIP_TRY(gemm(C, Uf, i + 1, m, GL, i, m, GD, i, m, mone, true));
      // G_{i-1,i} = -U_{i-1,s} G_{s,i} - U_{i-1,i} G_ii
      SINV_HIP_TRY(gemm(C, Us, i - 1, m, Gsr, i, m, GU, i - 1, m, mone, false));
      SINV_HIP_TRY(gemm(C, Uf, i, m, GD, i, m, GU, i - 1, m, mone, true));
      // G_{i,i-1} = -G_{i,s} L_{s,i-1} - G_ii L_{i,i-1}
      SINV_HIP_TRY(gemm(C, Gsc, i, m, Ls, i - 1, m, GL, i - 1, m, mone, false));
      SINV_HIP_TRY(gemm(C, GD, i, m, Lf, i, m, GL, i - 1, m, mone, true));
    }
    // G_{s+1,s+1} = inv(A) - U_{s+1,s} G_{s,s+1} - U_{s+1,s+2} G_{s+2,s+1}
    SINV_HIP_TRY(gemm(C, Us, sr + 1, m, Gsr, sr + 1, m, GD, sr + 1, m, mone, true, invA, sr + 1, m));
    SINV_HIP_TRY(gemm(C, Uf, sr + 2, m, GL, sr + 1, m, GD, sr + 1, m, mone, true));
    // G_{s,s+1} and G_{s+1,s} are the first entries of the G_{s,i} / G_{i,s} chains
    SINV_HIP_TRY(copy(C, Gsr, sr + 1, m, GU, sr, m, one));
    SINV_HIP_TRY(copy(C, Gsc, sr + 1, m, GL, sr, m, one));
  }
  SINV_HIP_TRY(prod_fork.join());
  return SINV_OK;
}
