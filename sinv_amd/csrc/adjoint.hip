// Adjoint (reverse-mode gradient) of the RGF selected inversion, batched.
//
// The reference has no adjoint; this is the gradient of the selected inverse
// computed by rgf.py:40-62.  For a real loss L of G = A^-1, with the complex
// gradient convention grad z = dL/dRe z + i dL/dIm z,
//
//     grad_A = -G^H grad_G G^H                      (restricted to the pattern).
//
// With B = A^H and F = B^-1 = G^H this is the forward-mode derivative of
// B -> B^-1 in direction grad_G:  d(F) = -F grad_G F.  grad_G lives on the
// block-tridiagonal pattern, so it is a tangent of a block-tridiagonal B, and
// the selected blocks of -F grad_G F are the tangents of RGF's selected
// outputs on B.  The adjoint is therefore one RGF sweep over B carrying
// tangents (dual numbers): every product C = X Y also forms dC = dX Y + X dY,
// every inverse g = M^-1 forms dg = -g dM g.  B's blocks are A's blocks read
// conjugate-transposed by the GEMM loader (B_ii = A_ii^H, B_{i,i+1} =
// A_{i+1,i}^H), so nothing is copied except the diagonal B_ii.
//
// Memory: the backward sweep needs the forward's left-connected g_i and dg_i.
// With checkpoint_every = k they are stored only at the end of every k-block
// segment and recomputed one segment at a time during the backward sweep
// (2 N/k + 2 k blocks per instance instead of 2 N).
#include <vector>

#include "sinv_core.hpp"

using namespace sinv;

namespace {

bool adj_desc_ok(const sinv_rgf_adjoint_desc* d) {
  if (d == nullptr || d->nblocks <= 0 || d->blocksize <= 0 || d->batch <= 0) return false;
  if (d->precision != SINV_PREC_FP64 && d->precision != SINV_PREC_MIXED) return false;
  if (d->diag == nullptr || d->gbar_diag == nullptr || d->abar_diag == nullptr) return false;
  if (d->nblocks > 1 && (d->upper == nullptr || d->lower == nullptr || d->gbar_upper == nullptr ||
                         d->gbar_lower == nullptr || d->abar_upper == nullptr ||
                         d->abar_lower == nullptr))
    return false;
  if (d->diag_stride < 0 || d->upper_stride < 0 || d->lower_stride < 0 ||
      d->gbar_diag_stride < 0 || d->gbar_upper_stride < 0 || d->gbar_lower_stride < 0)
    return false;
  return true;
}

struct AdjPlan {
  int k = 1, nseg = 1;
  long long blocks = 0;  // per instance
  size_t inv_elems = 0;
  size_t bytes = 0;
};

constexpr int kTemps = 12;

AdjPlan adj_plan(int N, int b, int batch, int every) {
  AdjPlan P;
  int k = every;
  if (k <= 0 || k >= N) k = N;
  if (k < 2) k = 2;  // the forward step reads slot i-1 while writing slot i
  P.k = k;
  P.nseg = (N + k - 1) / k;
  P.blocks = kTemps + 2LL * k + 2LL * (P.nseg > 1 ? P.nseg - 1 : 0);
  P.inv_elems = inverse_workspace_elems(b, batch);
  P.bytes = ((size_t)P.blocks * b * b * batch + P.inv_elems) * sizeof(cplx) + 512;
  return P;
}

}  // namespace

extern "C" size_t sinv_rgf_adjoint_workspace_bytes(const sinv_rgf_adjoint_desc* d) {
  if (!adj_desc_ok(d)) return 0;
  return adj_plan(d->nblocks, d->blocksize, d->batch, d->checkpoint_every).bytes;
}

extern "C" int sinv_rgf_adjoint_batched(const sinv_rgf_adjoint_desc* d, void* work,
                                        size_t work_bytes, void* stream) {
  if (!adj_desc_ok(d)) {
    set_error("sinv_rgf_adjoint_batched: invalid descriptor");
    return SINV_E_INVAL;
  }
  const int N = d->nblocks, b = d->blocksize, nb = d->batch;
  const AdjPlan P = adj_plan(N, b, nb, d->checkpoint_every);
  if (work == nullptr || work_bytes < P.bytes) {
    set_error("sinv_rgf_adjoint_batched: workspace too small (%zu < %zu)", work_bytes, P.bytes);
    return SINV_E_INVAL;
  }
  hipStream_t s = (hipStream_t)stream;
  const Prec prec = d->precision == SINV_PREC_FP64 ? Prec::F64 : Prec::F32;
  const long long b2 = (long long)b * b;
  cplx* base = (cplx*)(((uintptr_t)work + 255) & ~(uintptr_t)255);
  const long long WI = P.blocks * b2;  // instance stride in the workspace
  cplx* winv = base + WI * nb;
  auto wblk = [&](long long k) { return base + k * b2; };
  // temporaries
  cplx* Bd = wblk(0);
  cplx* Y = wblk(1);
  cplx* Qf = wblk(2);
  cplx* dM = wblk(3);
  cplx* Wt = wblk(4);
  cplx* Pb = wblk(5);
  cplx* Xn = wblk(6);
  cplx* Qb = wblk(7);
  cplx* Rb = wblk(8);
  cplx* Sb = wblk(9);
  cplx* Gpp[2] = {wblk(10), wblk(11)};
  auto slot_g = [&](int i) { return wblk(kTemps + (i % P.k)); };
  auto slot_dg = [&](int i) { return wblk(kTemps + P.k + (i % P.k)); };
  auto ck_g = [&](int seg) { return wblk(kTemps + 2LL * P.k + 2LL * (seg - 1)); };
  auto ck_dg = [&](int seg) { return wblk(kTemps + 2LL * P.k + 2LL * (seg - 1) + 1); };

  const cplx* D = (const cplx*)d->diag;
  const cplx* U = (const cplx*)d->upper;
  const cplx* Lo = (const cplx*)d->lower;
  const cplx* Td = (const cplx*)d->gbar_diag;
  const cplx* Tu = (const cplx*)d->gbar_upper;
  const cplx* Tl = (const cplx*)d->gbar_lower;
  cplx* Ad = (cplx*)d->abar_diag;
  cplx* Au = (cplx*)d->abar_upper;
  cplx* Al = (cplx*)d->abar_lower;
  const long long AdI = (long long)N * b2, AoI = (long long)(N > 1 ? N - 1 : 0) * b2;
  const cplx one{1.0, 0.0}, mone{-1.0, 0.0}, zero{0.0, 0.0};
  const cplx cds{d->dscale.re, -d->dscale.im};  // conj(dscale)
  const cplx cos_{d->oscale.re, -d->oscale.im};  // conj(oscale)
  const cplx mcos{-cos_.re, -cos_.im};

  // operand descriptor: pointer of instance 0, instance stride, op
  struct Opd {
    const cplx* p;
    long long inst;
    int op;
  };
  auto ws = [&](const cplx* p) { return Opd{p, WI, OP_N}; };
  auto gemm = [&](Opd A, Opd B, cplx* C, long long cinst, cplx alpha, const cplx* C0,
                  long long c0inst) -> hipError_t {
    GemmArgs g{};
    g.M = b; g.N = b; g.K = b; g.opA = A.op; g.opB = B.op;
    g.A = A.p; g.lda = b; g.sA_in = 0; g.sA_out = A.inst;
    g.B = B.p; g.ldb = b; g.sB_in = 0; g.sB_out = B.inst;
    g.C = C; g.ldc = b; g.sC_in = 0; g.sC_out = cinst;
    g.C0 = C0; g.ldc0 = b; g.sC0_in = 0; g.sC0_out = c0inst;
    g.alpha = alpha;
    g.beta = C0 ? one : zero;
    g.shift = nullptr;
    g.inner = 1; g.outer = nb;
    return launch_gemm(g, prec, s);
  };
  // C = alpha A B (+ C when acc)
  auto mm = [&](Opd A, Opd B, cplx* C, long long cinst, cplx alpha, bool acc) {
    return gemm(A, B, C, cinst, alpha, acc ? C : nullptr, cinst);
  };
  auto copy = [&](const cplx* src, long long sinst, cplx* dst, long long dinst) {
    return launch_axpby_shift(b, b, one, src, b, 0, sinst, dst, b, 0, dinst, nullptr, 1, nb, s);
  };
  // blocks of A / grad_G
  auto Ublk = [&](int i) { return Opd{U + i * b2, d->upper_stride, OP_C}; };   // -> B_{i+1,i}
  auto Lblk = [&](int i) { return Opd{Lo + i * b2, d->lower_stride, OP_C}; };  // -> B_{i,i+1}
  auto Tdb = [&](int i) { return Opd{Td + i * b2, d->gbar_diag_stride, OP_N}; };
  auto Tub = [&](int i) { return Opd{Tu + i * b2, d->gbar_upper_stride, OP_N}; };
  auto Tlb = [&](int i) { return Opd{Tl + i * b2, d->gbar_lower_stride, OP_N}; };

  // forward step i of RGF on B with tangents: (g_prev, dg_prev) -> (g_out, dg_out)
  auto fwd_step = [&](int i, cplx* gp, cplx* dgp, cplx* go, cplx* dgo) -> hipError_t {
    // B_ii = conj(dscale) D_ii^H + conj(shift) I
    cplx* bdst = i == 0 ? go : Bd;
    SINV_HIP_RET(launch_transpose(b, b, OP_C, D + i * b2, b, 0, d->diag_stride, bdst, b, 0, WI,
                                  1, nb, s, cds, (const cplx*)d->shift, true));
    if (i == 0) {
      SINV_HIP_RET(launch_inverse(go, b, b, 0, WI, 1, nb, winv, d->info, 0, s));
      SINV_HIP_RET(mm(ws(go), Tdb(0), Wt, WI, one, false));
      return mm(ws(Wt), ws(go), dgo, WI, mone, false);
    }
    // Y = B_{i,i-1} g_{i-1};  M = B_ii - Y B_{i-1,i}
    SINV_HIP_RET(mm(Ublk(i - 1), ws(gp), Y, WI, cos_, false));
    SINV_HIP_RET(gemm(ws(Y), Lblk(i - 1), go, WI, mcos, Bd, WI));
    // dM = T_ii - (T_{i,i-1} g_{i-1} + B_{i,i-1} dg_{i-1}) B_{i-1,i} - Y T_{i-1,i}
    SINV_HIP_RET(mm(Tlb(i - 1), ws(gp), Qf, WI, one, false));
    SINV_HIP_RET(mm(Ublk(i - 1), ws(dgp), Qf, WI, cos_, true));
    SINV_HIP_RET(gemm(ws(Qf), Lblk(i - 1), dM, WI, mcos, Td + i * b2, d->gbar_diag_stride));
    SINV_HIP_RET(mm(ws(Y), Tub(i - 1), dM, WI, mone, true));
    // g_i = M^-1, dg_i = -g_i dM g_i
    SINV_HIP_RET(launch_inverse(go, b, b, 0, WI, 1, nb, winv, d->info, i, s));
    SINV_HIP_RET(mm(ws(go), ws(dM), Wt, WI, one, false));
    return mm(ws(Wt), ws(go), dgo, WI, mone, false);
  };

  if (d->info != nullptr) SINV_HIP_TRY(launch_fill_u64(d->info, 1, SINV_INFO_OK, s));

  // ---- forward sweep, keeping the last segment and the segment checkpoints
  for (int i = 0; i < N; ++i) {
    cplx* gp = i > 0 ? slot_g(i - 1) : nullptr;
    cplx* dgp = i > 0 ? slot_dg(i - 1) : nullptr;
    SINV_HIP_TRY(fwd_step(i, gp, dgp, slot_g(i), slot_dg(i)));
    if (i % P.k == P.k - 1 && i + 1 < N) {
      const int seg = (i + 1) / P.k;
      SINV_HIP_TRY(copy(slot_g(i), WI, ck_g(seg), WI));
      SINV_HIP_TRY(copy(slot_dg(i), WI, ck_dg(seg), WI));
    }
  }

  // ---- backward sweep (rgf.py:51-62 on B, with tangents)
  int cur = 0;
  SINV_HIP_TRY(copy(slot_g(N - 1), WI, Gpp[0], WI));               // G_{N-1,N-1}(B)
  SINV_HIP_TRY(copy(slot_dg(N - 1), WI, Ad + (N - 1) * b2, AdI));  // its tangent
  for (int seg = P.nseg - 1; seg >= 0; --seg) {
    const int i0 = seg * P.k, i1 = (seg + 1) * P.k < N ? (seg + 1) * P.k : N;
    if (seg != P.nseg - 1) {
      for (int i = i0; i < i1; ++i) {
        cplx* gp = i == 0 ? nullptr : (i == i0 ? ck_g(seg) : slot_g(i - 1));
        cplx* dgp = i == 0 ? nullptr : (i == i0 ? ck_dg(seg) : slot_dg(i - 1));
        SINV_HIP_TRY(fwd_step(i, gp, dgp, slot_g(i), slot_dg(i)));
      }
    }
    for (int i = (i1 - 1 < N - 2 ? i1 - 1 : N - 2); i >= i0; --i) {
      cplx* g = slot_g(i);
      cplx* dg = slot_dg(i);
      cplx* Gn = Gpp[cur];
      cplx* Gi = Gpp[cur ^ 1];
      const cplx* dGn = Ad + (i + 1) * b2;
      cplx* Ali = Al + i * b2;
      cplx* Aui = Au + i * b2;
      cplx* Adi = Ad + i * b2;
      // primal: P = Gn B_{i+1,i}; Xn = -P g; Q = g B_{i,i+1}; G_ii = g - Q Xn
      SINV_HIP_TRY(mm(ws(Gn), Ublk(i), Pb, WI, cos_, false));
      SINV_HIP_TRY(mm(ws(Pb), ws(g), Xn, WI, mone, false));
      SINV_HIP_TRY(mm(ws(g), Lblk(i), Qb, WI, cos_, false));
      SINV_HIP_TRY(gemm(ws(Qb), ws(Xn), Gi, WI, mone, g, WI));
      // tangent of G_{i+1,i}: -(dGn B_{i+1,i} + Gn T_{i+1,i}) g - P dg
      SINV_HIP_TRY(mm(Opd{dGn, AdI, OP_N}, Ublk(i), Rb, WI, cos_, false));
      SINV_HIP_TRY(mm(ws(Gn), Tlb(i), Rb, WI, one, true));
      SINV_HIP_TRY(mm(ws(Rb), ws(g), Ali, AoI, mone, false));
      SINV_HIP_TRY(mm(ws(Pb), ws(dg), Ali, AoI, mone, true));
      // tangent of G_{i,i+1}: -(dg B_{i,i+1} + g T_{i,i+1}) Gn - Q dGn
      SINV_HIP_TRY(mm(ws(dg), Lblk(i), Sb, WI, cos_, false));
      SINV_HIP_TRY(mm(ws(g), Tub(i), Sb, WI, one, true));
      SINV_HIP_TRY(mm(ws(Sb), ws(Gn), Aui, AoI, mone, false));
      SINV_HIP_TRY(mm(ws(Qb), Opd{dGn, AdI, OP_N}, Aui, AoI, mone, true));
      // tangent of G_ii: dg - S Xn - Q dXn
      SINV_HIP_TRY(gemm(ws(Sb), ws(Xn), Adi, AdI, mone, dg, WI));
      SINV_HIP_TRY(mm(ws(Qb), Opd{Ali, AoI, OP_N}, Adi, AdI, mone, true));
      cur ^= 1;
    }
  }
  return SINV_OK;
}
