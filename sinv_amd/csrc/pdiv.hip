// PDIV (P-Division) selected inversion on one GPU: the partition-local work of
// src/sinv/algorithms/pdiv/pdiv_localmap.py run as batched kernels over
// partitions (the reference's ranks) and instances.
//
//  * divide_matrix / partition_subdomain (pdiv_utils.py:84-170): m = N // P
//    blocks per partition, the first takes the remainder; the bridges are the
//    couplings A[e_p, e_p + 1] = upper[e_p] and A[e_p + 1, e_p] = lower[e_p].
//  * invert_partition (pdiv_localmap.py:94-115): each partition K_p is
//    assembled dense on the device and inverted, Phi_p = K_p^-1, batched over
//    the equal-size partitions 1..P-1 (one launch sequence) and partition 0.
//  * The reference merges partitions pairwise with Sherman-Morrison updates
//    kept as per-rank "maps" (update_maps :188-246).  With every partition on
//    one device the same Woodbury algebra is applied directly: the rest of the
//    chain couples into partition p only through its first block f and last
//    block l, with self-energies
//      SL_p = lower[s_p - 1] gL_{p-1} upper[s_p - 1],
//      SR_p = upper[e_p] gR_{p+1} lower[e_p],
//    where gL_p (gR_p) is the last (first) diagonal block of the inverse of
//    partitions 0..p (p..P-1) alone, obtained by one sweep each over the
//    partitions from Phi's corners:
//      gL_p = Phi_ll + Phi_lf (I - SL_p Phi_ff)^-1 SL_p Phi_fl,
//      gR_p = Phi_ff + Phi_fl (I - SR_p Phi_ll)^-1 SR_p Phi_lf.
//    Then (Woodbury, E = [e_f, e_l], Sigma_p = diag(SL_p, SR_p))
//      X_pp = Phi_p + Phi_p E C_p E^T Phi_p,  C_p = (I - Sigma_p E^T Phi_p E)^-1 Sigma_p,
//    which is produce_matrix_elements (:1194-1241) with the maps M9..M12 as
//    the blocks of C_p, produced for all tridiagonal blocks of all partitions
//    as batched GEMMs.
//  * bridges (produce_bridges :1291-1436), s = e_p + 1: splitting the chain
//    after e_p, the inverse's block identities give two exact forms,
//      X[s, e] = -gR_{p+1} lower[e] X[e, e] = -X[s, s] lower[e] gL_p,
//      X[e, s] = -X[e, e] upper[e] gR_{p+1} = -gL_p upper[e] X[s, s],
//    whose rounding errors scale with the sub-chain inverse they use: a
//    nearly singular half of the chain (a large |gR| or |gL|) spoils one of
//    them.  Both are formed and, per bridge and instance, the pair using the
//    smaller of |gL_p|_F, |gR_{p+1}|_F is kept.
#include <vector>

#include "sinv_core.hpp"

using namespace sinv;

namespace {

bool pdiv_desc_ok(const sinv_pdiv_desc* d) {
  if (d == nullptr || d->nblocks <= 0 || d->blocksize <= 0 || d->batch <= 0) return false;
  const int P = d->npart;
  if (P < 1 || (P & (P - 1)) != 0 || d->nblocks < P) return false;  // pdiv_utils.py:16-80
  if (d->diag == nullptr || d->g_diag == nullptr) return false;
  if (d->nblocks > 1 && (d->upper == nullptr || d->lower == nullptr || d->g_upper == nullptr ||
                         d->g_lower == nullptr))
    return false;
  if (d->diag_stride < 0 || d->upper_stride < 0 || d->lower_stride < 0) return false;
  return true;
}

struct PdivPlan {
  int N = 0, b = 0, P = 0, m = 0, n0 = 0;
  std::vector<int> start, size;
  std::vector<long long> phi;  // offset of Phi_p (dense (size_p b)^2)
  std::vector<long long> tf;   // offset of T_p = [T_f; T_l] (2b x size_p b)
  long long SL = 0, SR = 0, gL = 0, gR = 0, Fc = 0, Y = 0, Cm = 0, tmp = 0;
  long long per_inst = 0;
  size_t inv_elems = 0;
  size_t bytes = 0;
};

PdivPlan pdiv_plan(int N, int b, int P, int batch) {
  PdivPlan Q;
  Q.N = N; Q.b = b; Q.P = P;
  Q.m = N / P;
  Q.n0 = N - Q.m * (P - 1);
  long long o = 0;
  const long long b2 = (long long)b * b;
  for (int p = 0; p < P; ++p) {
    Q.start.push_back(p == 0 ? 0 : Q.n0 + (p - 1) * Q.m);
    Q.size.push_back(p == 0 ? Q.n0 : Q.m);
  }
  for (int p = 0; p < P; ++p) {
    Q.phi.push_back(o);
    o += (long long)Q.size[p] * b * Q.size[p] * b;
  }
  for (int p = 0; p < P; ++p) {
    Q.tf.push_back(o);
    o += 2LL * b * Q.size[p] * b;
  }
  Q.SL = o; o += P * b2;
  Q.SR = o; o += P * b2;
  Q.gL = o; o += P * b2;
  Q.gR = o; o += P * b2;
  Q.Fc = o; o += 4LL * P * b2;
  Q.Y = o; o += 4LL * P * b2;
  Q.Cm = o; o += 4LL * P * b2;
  Q.tmp = o; o += 3 * b2;
  Q.per_inst = o;
  size_t w = inverse_workspace_elems(Q.n0 * b, batch);
  auto mx = [&](size_t v) { if (v > w) w = v; };
  if (P > 1) mx(inverse_workspace_elems(Q.m * b, batch * (P - 1)));
  mx(inverse_workspace_elems(2 * b, batch * P));
  mx(inverse_workspace_elems(b, batch));
  Q.inv_elems = w;
  Q.bytes = ((size_t)Q.per_inst * batch + Q.inv_elems + (size_t)batch) * sizeof(cplx) + 1024;
  return Q;
}

// dense partition K_p (block rows start.. start+n-1) of the block-tridiagonal
// input, grid: x over (n b)^2 elements, y over (partition in group, instance)
__global__ void assemble_kernel(const cplx* D, long long sD, const cplx* U, long long sU,
                                const cplx* L, long long sL, int b, int n, int start0, int pstep,
                                int nparts, cplx* Phi, long long sPhi_p, long long sPhi_z) {
  const long long nb = (long long)n * b;
  const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nb * nb) return;
  const int p = blockIdx.y % nparts, z = blockIdx.y / nparts;
  const int row = (int)(e / nb), col = (int)(e - (long long)row * nb);
  const int br = row / b, bc = col / b, r = row - br * b, c = col - bc * b;
  const int s = start0 + p * pstep;
  const long long b2 = (long long)b * b;
  double2 v = double2{0.0, 0.0};
  if (br == bc)
    v = *reinterpret_cast<const double2*>(D + z * sD + (s + br) * b2 + (long long)r * b + c);
  else if (bc == br + 1)
    v = *reinterpret_cast<const double2*>(U + z * sU + (s + br) * b2 + (long long)r * b + c);
  else if (br == bc + 1)
    v = *reinterpret_cast<const double2*>(L + z * sL + (s + bc) * b2 + (long long)r * b + c);
  *reinterpret_cast<double2*>(Phi + z * sPhi_z + p * sPhi_p + e) = v;
}

// Per (bridge p, instance z): keep the bridge pair formed with the smaller-norm
// sub-chain block.  cand holds 4 arrays of nbr blocks: lower via gR, lower via
// gL, upper via gR, upper via gL.
__global__ __launch_bounds__(256) void bridge_select_kernel(const cplx* gL, const cplx* gR1,
                                                            long long sgI, const cplx* cand,
                                                            long long scI, int nbr, int b,
                                                            cplx* Gl, cplx* Gu, long long se,
                                                            long long GoI) {
  __shared__ double rl[256], rr[256];
  const int p = blockIdx.x, z = blockIdx.y;
  const long long b2 = (long long)b * b;
  const cplx* L = gL + z * sgI + p * b2;
  const cplx* R = gR1 + z * sgI + p * b2;
  double nl = 0.0, nr = 0.0;
  for (long long i = threadIdx.x; i < b2; i += blockDim.x) {
    nl += L[i].re * L[i].re + L[i].im * L[i].im;
    nr += R[i].re * R[i].re + R[i].im * R[i].im;
  }
  rl[threadIdx.x] = nl;
  rr[threadIdx.x] = nr;
  __syncthreads();
  for (int o = blockDim.x / 2; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) {
      rl[threadIdx.x] += rl[threadIdx.x + o];
      rr[threadIdx.x] += rr[threadIdx.x + o];
    }
    __syncthreads();
  }
  const bool use_left = rl[0] < rr[0];
  const cplx* base = cand + z * scI + p * b2;
  const cplx* lo = base + (use_left ? 1 : 0) * nbr * b2;
  const cplx* up = base + (use_left ? 3 : 2) * nbr * b2;
  cplx* gl = Gl + z * GoI + p * se;
  cplx* gu = Gu + z * GoI + p * se;
  for (long long i = threadIdx.x; i < b2; i += blockDim.x) {
    *reinterpret_cast<double2*>(gl + i) = *reinterpret_cast<const double2*>(lo + i);
    *reinterpret_cast<double2*>(gu + i) = *reinterpret_cast<const double2*>(up + i);
  }
}

}  // namespace

extern "C" size_t sinv_pdiv_workspace_bytes(const sinv_pdiv_desc* d) {
  if (!pdiv_desc_ok(d)) return 0;
  return pdiv_plan(d->nblocks, d->blocksize, d->npart, d->batch).bytes;
}

extern "C" int sinv_pdiv_batched(const sinv_pdiv_desc* d, void* work, size_t work_bytes,
                                 void* stream) {
  if (!pdiv_desc_ok(d)) {
    set_error("sinv_pdiv_batched: invalid descriptor (npart must be a power of 2 and at most "
              "nblocks)");
    return SINV_E_INVAL;
  }
  const int N = d->nblocks, b = d->blocksize, nb = d->batch, P = d->npart;
  const PdivPlan Q = pdiv_plan(N, b, P, nb);
  if (work == nullptr || work_bytes < Q.bytes) {
    set_error("sinv_pdiv_batched: workspace too small (%zu < %zu)", work_bytes, Q.bytes);
    return SINV_E_INVAL;
  }
  hipStream_t s = (hipStream_t)stream;
  const long long b2 = (long long)b * b;
  cplx* base = (cplx*)(((uintptr_t)work + 255) & ~(uintptr_t)255);
  const long long WI = Q.per_inst;
  cplx* invw = base + WI * nb;
  cplx* ones = invw + Q.inv_elems;
  const cplx one{1.0, 0.0}, mone{-1.0, 0.0}, zero{0.0, 0.0};
  auto W = [&](long long off) { return base + off; };

  const cplx* D = (const cplx*)d->diag;
  const cplx* Uu = (const cplx*)d->upper;
  const cplx* Ll = (const cplx*)d->lower;
  const long long sD = d->diag_stride, sU = d->upper_stride, sL = d->lower_stride;
  cplx* Gd = (cplx*)d->g_diag;
  cplx* Gu = (cplx*)d->g_upper;
  cplx* Gl = (cplx*)d->g_lower;
  const long long GdI = (long long)N * b2, GoI = (long long)(N > 1 ? N - 1 : 0) * b2;

  if (d->info != nullptr) SINV_HIP_TRY(launch_fill_u64(d->info, 1, SINV_INFO_OK, s));
  SINV_HIP_TRY(launch_fill(ones, nb, one, s));

  // C = alpha op(A) B (+ beta C0) (+ I), batched inner x outer
  struct Op {
    const cplx* p;
    int ld;
    long long si, so;
  };
  auto gemm = [&](int M, int Nn, int K, Op A, int opA, Op B, cplx* C, int ldc, long long sci,
                  long long sco, cplx alpha, const cplx* C0, cplx beta, bool plus_identity,
                  int inner, int outer) {
    GemmArgs g{};
    g.M = M; g.N = Nn; g.K = K; g.opA = opA; g.opB = OP_N;
    g.A = A.p; g.lda = A.ld; g.sA_in = A.si; g.sA_out = A.so;
    g.B = B.p; g.ldb = B.ld; g.sB_in = B.si; g.sB_out = B.so;
    g.C = C; g.ldc = ldc; g.sC_in = sci; g.sC_out = sco;
    g.C0 = C0; g.ldc0 = ldc; g.sC0_in = sci; g.sC0_out = sco;
    g.alpha = alpha;
    g.beta = C0 != nullptr ? beta : zero;
    g.shift = plus_identity ? ones : nullptr;
    g.inner = inner; g.outer = outer;
    return launch_gemm(g, Prec::F64, s);
  };
  // b x b block copy (batched inner x outer): C = A
  auto copy = [&](const cplx* A, int lda, long long sai, long long sao, cplx* C, int ldc,
                  long long sci, long long sco, int inner, int outer) {
    return launch_axpby_shift(b, b, one, A, lda, sai, sao, C, ldc, sci, sco, nullptr, inner,
                              outer, s);
  };

  // ---- 1. dense partitions and their inverses (pdiv_localmap.py:94-115)
  for (int grp = 0; grp < (P > 1 ? 2 : 1); ++grp) {
    const int p0 = grp == 0 ? 0 : 1;
    const int np = grp == 0 ? 1 : P - 1;
    const int n = Q.size[p0];
    const long long nbb = (long long)n * b;
    const long long elems = nbb * nbb;
    const long long gx = (elems + 255) / 256;
    if (gx > 0x7fffffffLL || (long long)np * nb > 65535) {
      set_error("sinv_pdiv_batched: partition too large for one launch");
      return SINV_E_INVAL;
    }
    hipLaunchKernelGGL(assemble_kernel, dim3((unsigned)gx, (unsigned)(np * nb)), dim3(256), 0, s,
                       D, sD, Uu, sU, Ll, sL, b, n, Q.start[p0], Q.m, np, W(Q.phi[p0]), elems, WI);
    SINV_HIP_TRY(hipGetLastError());
    SINV_HIP_TRY(launch_inverse(W(Q.phi[p0]), (int)nbb, (int)nbb, elems, WI, np, nb, invw,
                                d->info, 1 + grp, s));
  }
  auto phi = [&](int p, int br, int bc) {  // block (br, bc) of Phi_p
    const long long ld = (long long)Q.size[p] * b;
    return W(Q.phi[p]) + (long long)br * b * ld + (long long)bc * b;
  };
  auto ldp = [&](int p) { return Q.size[p] * b; };
  auto last = [&](int p) { return Q.size[p] - 1; };

  // ---- 2. boundary sweeps over partitions (b x b, batched over instances)
  cplx* T0 = W(Q.tmp);
  cplx* S = W(Q.tmp + b2);
  cplx* Wm = W(Q.tmp + 2 * b2);
  for (int z = 0; z < nb; ++z) {  // SL_0 = 0, SR_{P-1} = 0
    SINV_HIP_TRY(launch_fill(W(Q.SL) + z * WI, b2, zero, s));
    SINV_HIP_TRY(launch_fill(W(Q.SR) + z * WI + (long long)(P - 1) * b2, b2, zero, s));
  }
  if (P > 1) {
    // gL_0 = Phi_0[l, l]; gR_{P-1} = Phi_{P-1}[f, f]
    SINV_HIP_TRY(copy(phi(0, last(0), last(0)), ldp(0), 0, WI, W(Q.gL), b, 0, WI, 1, nb));
    SINV_HIP_TRY(copy(phi(P - 1, 0, 0), ldp(P - 1), 0, WI, W(Q.gR) + (long long)(P - 1) * b2, b,
                      0, WI, 1, nb));
  }
  for (int p = 1; p < P; ++p) {  // left: SL_p, gL_p
    const int e = Q.start[p] - 1;
    cplx* SLp = W(Q.SL) + (long long)p * b2;
    SINV_HIP_TRY(gemm(b, b, b, Op{Ll + (long long)e * b2, b, 0, sL}, OP_N,
                      Op{W(Q.gL) + (long long)(p - 1) * b2, b, 0, WI}, T0, b, 0, WI, one, nullptr,
                      zero, false, 1, nb));
    SINV_HIP_TRY(gemm(b, b, b, Op{T0, b, 0, WI}, OP_N, Op{Uu + (long long)e * b2, b, 0, sU}, SLp,
                      b, 0, WI, one, nullptr, zero, false, 1, nb));
    if (p == P - 1) break;  // gL_{P-1} is not needed
    // S = I - SL Phi_ff; W = S^-1 SL; gL = Phi_ll + Phi_lf (W Phi_fl)
    SINV_HIP_TRY(gemm(b, b, b, Op{SLp, b, 0, WI}, OP_N, Op{phi(p, 0, 0), ldp(p), 0, WI}, S, b, 0,
                      WI, mone, nullptr, zero, true, 1, nb));
    SINV_HIP_TRY(launch_inverse(S, b, b, 0, WI, 1, nb, invw, d->info, 3, s));
    SINV_HIP_TRY(gemm(b, b, b, Op{S, b, 0, WI}, OP_N, Op{SLp, b, 0, WI}, Wm, b, 0, WI, one,
                      nullptr, zero, false, 1, nb));
    SINV_HIP_TRY(gemm(b, b, b, Op{Wm, b, 0, WI}, OP_N, Op{phi(p, 0, last(p)), ldp(p), 0, WI}, T0,
                      b, 0, WI, one, nullptr, zero, false, 1, nb));
    cplx* gLp = W(Q.gL) + (long long)p * b2;
    SINV_HIP_TRY(copy(phi(p, last(p), last(p)), ldp(p), 0, WI, gLp, b, 0, WI, 1, nb));
    SINV_HIP_TRY(gemm(b, b, b, Op{phi(p, last(p), 0), ldp(p), 0, WI}, OP_N, Op{T0, b, 0, WI},
                      gLp, b, 0, WI, one, gLp, one, false, 1, nb));
  }
  for (int p = P - 2; p >= 0; --p) {  // right: SR_p, gR_p
    const int e = Q.start[p] + Q.size[p] - 1;
    cplx* SRp = W(Q.SR) + (long long)p * b2;
    SINV_HIP_TRY(gemm(b, b, b, Op{Uu + (long long)e * b2, b, 0, sU}, OP_N,
                      Op{W(Q.gR) + (long long)(p + 1) * b2, b, 0, WI}, T0, b, 0, WI, one, nullptr,
                      zero, false, 1, nb));
    SINV_HIP_TRY(gemm(b, b, b, Op{T0, b, 0, WI}, OP_N, Op{Ll + (long long)e * b2, b, 0, sL}, SRp,
                      b, 0, WI, one, nullptr, zero, false, 1, nb));
    if (p == 0) break;  // gR_0 is not needed
    SINV_HIP_TRY(gemm(b, b, b, Op{SRp, b, 0, WI}, OP_N,
                      Op{phi(p, last(p), last(p)), ldp(p), 0, WI}, S, b, 0, WI, mone, nullptr,
                      zero, true, 1, nb));
    SINV_HIP_TRY(launch_inverse(S, b, b, 0, WI, 1, nb, invw, d->info, 4, s));
    SINV_HIP_TRY(gemm(b, b, b, Op{S, b, 0, WI}, OP_N, Op{SRp, b, 0, WI}, Wm, b, 0, WI, one,
                      nullptr, zero, false, 1, nb));
    SINV_HIP_TRY(gemm(b, b, b, Op{Wm, b, 0, WI}, OP_N, Op{phi(p, last(p), 0), ldp(p), 0, WI}, T0,
                      b, 0, WI, one, nullptr, zero, false, 1, nb));
    cplx* gRp = W(Q.gR) + (long long)p * b2;
    SINV_HIP_TRY(copy(phi(p, 0, 0), ldp(p), 0, WI, gRp, b, 0, WI, 1, nb));
    SINV_HIP_TRY(gemm(b, b, b, Op{phi(p, 0, last(p)), ldp(p), 0, WI}, OP_N, Op{T0, b, 0, WI}, gRp,
                      b, 0, WI, one, gRp, one, false, 1, nb));
  }

  // ---- 3. C_p = (I - Sigma_p Fc_p)^-1 Sigma_p, Fc_p = [[Phi_ff, Phi_fl], [Phi_lf, Phi_ll]]
  const int b2i = 2 * b;
  const long long F2 = 4 * b2;  // one 2b x 2b matrix
  for (int grp = 0; grp < (P > 1 ? 2 : 1); ++grp) {
    const int p0 = grp == 0 ? 0 : 1;
    const int np = grp == 0 ? 1 : P - 1;
    const long long sp = grp == 0 ? 0 : (long long)Q.size[1] * b * Q.size[1] * b;  // Phi stride
    cplx* Fc = W(Q.Fc) + p0 * F2;
    const int L = last(p0);
    SINV_HIP_TRY(copy(phi(p0, 0, 0), ldp(p0), sp, WI, Fc, b2i, F2, WI, np, nb));
    SINV_HIP_TRY(copy(phi(p0, 0, L), ldp(p0), sp, WI, Fc + b, b2i, F2, WI, np, nb));
    SINV_HIP_TRY(copy(phi(p0, L, 0), ldp(p0), sp, WI, Fc + (long long)b * b2i, b2i, F2, WI, np,
                      nb));
    SINV_HIP_TRY(copy(phi(p0, L, L), ldp(p0), sp, WI, Fc + (long long)b * b2i + b, b2i, F2, WI,
                      np, nb));
  }
  {
    // Y = [[I - SL Fff, -SL Ffl], [-SR Flf, I - SR Fll]], batched over all P partitions
    cplx* Fc = W(Q.Fc);
    cplx* Y = W(Q.Y);
    const Op SL{W(Q.SL), b, b2, WI}, SR{W(Q.SR), b, b2, WI};
    SINV_HIP_TRY(gemm(b, b, b, SL, OP_N, Op{Fc, b2i, F2, WI}, Y, b2i, F2, WI, mone, nullptr, zero,
                      true, P, nb));
    SINV_HIP_TRY(gemm(b, b, b, SL, OP_N, Op{Fc + b, b2i, F2, WI}, Y + b, b2i, F2, WI, mone,
                      nullptr, zero, false, P, nb));
    SINV_HIP_TRY(gemm(b, b, b, SR, OP_N, Op{Fc + (long long)b * b2i, b2i, F2, WI},
                      Y + (long long)b * b2i, b2i, F2, WI, mone, nullptr, zero, false, P, nb));
    SINV_HIP_TRY(gemm(b, b, b, SR, OP_N, Op{Fc + (long long)b * b2i + b, b2i, F2, WI},
                      Y + (long long)b * b2i + b, b2i, F2, WI, mone, nullptr, zero, true, P, nb));
    SINV_HIP_TRY(launch_inverse(Y, b2i, b2i, F2, WI, P, nb, invw, d->info, 5, s));
    // C = Y^-1 Sigma: left column block times SL, right column block times SR
    cplx* Cm = W(Q.Cm);
    SINV_HIP_TRY(gemm(b2i, b, b, Op{Y, b2i, F2, WI}, OP_N, SL, Cm, b2i, F2, WI, one, nullptr,
                      zero, false, P, nb));
    SINV_HIP_TRY(gemm(b2i, b, b, Op{Y + b, b2i, F2, WI}, OP_N, SR, Cm + b, b2i, F2, WI, one,
                      nullptr, zero, false, P, nb));
  }

  // ---- 4. production (produce_matrix_elements :1194-1241), per partition group
  for (int grp = 0; grp < (P > 1 ? 2 : 1); ++grp) {
    const int p0 = grp == 0 ? 0 : 1;
    const int np = grp == 0 ? 1 : P - 1;
    const int n = Q.size[p0];
    const int ld = n * b;
    const long long sp = grp == 0 ? 0 : (long long)ld * ld;   // Phi stride between partitions
    const long long st = grp == 0 ? 0 : 2LL * b * ld;         // T stride
    const long long sg = grp == 0 ? 0 : (long long)Q.m * b2;  // output stride (blocks)
    const int L = n - 1;
    cplx* Cm = W(Q.Cm) + p0 * F2;
    cplx* Tf = W(Q.tf[p0]);
    cplx* Tl = Tf + (long long)b * ld;
    // T_f = C_ff Phi[f, :] + C_fl Phi[l, :];  T_l = C_lf Phi[f, :] + C_ll Phi[l, :]
    for (int h = 0; h < 2; ++h) {
      cplx* T = h == 0 ? Tf : Tl;
      const cplx* Crow = Cm + (long long)h * b * b2i;
      SINV_HIP_TRY(gemm(b, ld, b, Op{Crow, b2i, F2, WI}, OP_N, Op{phi(p0, 0, 0), ld, sp, WI}, T,
                        ld, st, WI, one, nullptr, zero, false, np, nb));
      SINV_HIP_TRY(gemm(b, ld, b, Op{Crow + b, b2i, F2, WI}, OP_N, Op{phi(p0, L, 0), ld, sp, WI},
                        T, ld, st, WI, one, T, one, false, np, nb));
    }
    // X[r, c] = Phi[r, c] + Phi[r, f] T_f[:, c] + Phi[r, l] T_l[:, c]; one launch per kind
    // and instance (the block index and the partition are the two batch levels)
    const long long rowstep = (long long)b * ld;
    for (int z = 0; z < nb; ++z) {
      const long long zp = z * WI;
      auto produce = [&](int count, int r0, int c0, cplx* out, long long out_sz) {
        if (count <= 0) return hipSuccess;
        const long long sa = rowstep;       // next row block
        const long long sbs = (long long)b;  // next column block of T
        const cplx* Pf = phi(p0, r0, 0) + zp;
        const cplx* Pl = phi(p0, r0, L) + zp;
        const cplx* P0 = phi(p0, r0, c0) + zp;
        // out = Phi[r,c] + Phi[r,f] T_f[:,c]
        GemmArgs g{};
        g.M = b; g.N = b; g.K = b; g.opA = OP_N; g.opB = OP_N;
        g.A = Pf; g.lda = ld; g.sA_in = sa; g.sA_out = sp;
        g.B = Tf + zp + (long long)c0 * b; g.ldb = ld; g.sB_in = sbs; g.sB_out = st;
        g.C0 = P0; g.ldc0 = ld; g.sC0_in = rowstep + b; g.sC0_out = sp;
        g.C = out; g.ldc = b; g.sC_in = b2; g.sC_out = out_sz;
        g.alpha = one; g.beta = one; g.shift = nullptr;
        g.inner = count; g.outer = np;
        hipError_t err = launch_gemm(g, Prec::F64, s);
        if (err != hipSuccess) return err;
        // out += Phi[r,l] T_l[:,c]
        g.A = Pl;
        g.B = Tl + zp + (long long)c0 * b;
        g.C0 = out; g.ldc0 = b; g.sC0_in = b2; g.sC0_out = out_sz;
        return launch_gemm(g, Prec::F64, s);
      };
      cplx* gd = Gd + z * GdI + (long long)Q.start[p0] * b2;
      SINV_HIP_TRY(produce(n, 0, 0, gd, sg));
      if (N > 1) {
        cplx* gu = Gu + z * GoI + (long long)Q.start[p0] * b2;
        cplx* gl = Gl + z * GoI + (long long)Q.start[p0] * b2;
        SINV_HIP_TRY(produce(n - 1, 0, 1, gu, sg));
        SINV_HIP_TRY(produce(n - 1, 1, 0, gl, sg));
      }
    }
  }

  // ---- 5. bridges (both exact forms, then the better-conditioned pair)
  if (P > 1) {
    const int e0 = Q.n0 - 1;
    const int nbr = P - 1;
    const long long se = (long long)Q.m * b2;  // e_{p+1} - e_p = m blocks
    const Op gR1{W(Q.gR) + b2, b, b2, WI};     // gR_{p+1}
    const Op gL0{W(Q.gL), b, b2, WI};          // gL_p
    const Op Lb{Ll + (long long)e0 * b2, b, se, sL};
    const Op Ub{Uu + (long long)e0 * b2, b, se, sU};
    const Op Xee{Gd + (long long)e0 * b2, b, se, GdI};
    const Op Xss{Gd + (long long)(e0 + 1) * b2, b, se, GdI};
    cplx* cand = W(Q.Y);  // 4 (P - 1) blocks per instance
    cplx* T1 = W(Q.Fc);   // P - 1 blocks per instance
    auto cb = [&](int k) { return cand + (long long)k * nbr * b2; };
    // lower via gR: -(gR lower) X[e, e]
    SINV_HIP_TRY(gemm(b, b, b, gR1, OP_N, Lb, T1, b, b2, WI, one, nullptr, zero, false, nbr, nb));
    SINV_HIP_TRY(gemm(b, b, b, Op{T1, b, b2, WI}, OP_N, Xee, cb(0), b, b2, WI, mone, nullptr, zero,
                      false, nbr, nb));
    // lower via gL: -(X[s, s] lower) gL
    SINV_HIP_TRY(gemm(b, b, b, Xss, OP_N, Lb, T1, b, b2, WI, one, nullptr, zero, false, nbr, nb));
    SINV_HIP_TRY(gemm(b, b, b, Op{T1, b, b2, WI}, OP_N, gL0, cb(1), b, b2, WI, mone, nullptr, zero,
                      false, nbr, nb));
    // upper via gR: -(X[e, e] upper) gR
    SINV_HIP_TRY(gemm(b, b, b, Xee, OP_N, Ub, T1, b, b2, WI, one, nullptr, zero, false, nbr, nb));
    SINV_HIP_TRY(gemm(b, b, b, Op{T1, b, b2, WI}, OP_N, gR1, cb(2), b, b2, WI, mone, nullptr, zero,
                      false, nbr, nb));
    // upper via gL: -(gL upper) X[s, s]
    SINV_HIP_TRY(gemm(b, b, b, gL0, OP_N, Ub, T1, b, b2, WI, one, nullptr, zero, false, nbr, nb));
    SINV_HIP_TRY(gemm(b, b, b, Op{T1, b, b2, WI}, OP_N, Xss, cb(3), b, b2, WI, mone, nullptr, zero,
                      false, nbr, nb));
    if ((long long)nb > 65535) {
      set_error("sinv_pdiv_batched: batch too large");
      return SINV_E_INVAL;
    }
    hipLaunchKernelGGL(bridge_select_kernel, dim3((unsigned)nbr, (unsigned)nb), dim3(256), 0, s,
                       W(Q.gL), W(Q.gR) + b2, WI, cand, WI, nbr, b, Gl + (long long)e0 * b2,
                       Gu + (long long)e0 * b2, se, GoI);
    SINV_HIP_TRY(hipGetLastError());
  }
  return SINV_OK;
}
