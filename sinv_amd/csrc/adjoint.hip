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
//
// The forward step's Y_i = B_{i,i-1} g_{i-1} and its tangent
// Qf_i = T_{i,i-1} g_{i-1} + B_{i,i-1} dg_{i-1} are kept too (k + 1 slots each,
// indexed modulo k + 1, so a segment's backward steps find those of its own
// steps and of the next segment's first step), which makes the backward
// step's G_{i+1,i} = -G_{i+1,i+1} Y_{i+1} one product and its tangent
// -dG_{i+1,i+1} Y_{i+1} - G_{i+1,i+1} Qf_{i+1} two: 11 products per
// backward step instead of 14.
#include <vector>

#include "sinv_core.hpp"

using namespace sinv;

namespace {

bool adj_desc_ok(const sinv_rgf_adjoint_desc* d) {
  if (d == nullptr || d->nblocks <= 0 || d->batch <= 0) return false;
  if (d->bsizes == nullptr && d->blocksize <= 0) return false;
  if (d->precision != SINV_PREC_FP64 && d->precision != SINV_PREC_MIXED &&
      d->precision != SINV_PREC_INT8X)
    return false;
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
  size_t i8x_bytes = 0;  // digit planes of the sliced-int8 GEMMs (SINV_PREC_INT8X)
  size_t bytes = 0;
};

constexpr int kTemps = 8;

AdjPlan adj_plan(const BlockLayout& L, int batch, int every, int precision) {
  AdjPlan P;
  const int N = L.N;
  int k = every;
  if (k <= 0 || k >= N) k = N;
  if (k < 2) k = 2;  // the forward step reads slot i-1 while writing slot i
  P.k = k;
  P.nseg = (N + k - 1) / k;
  P.blocks = kTemps + 2LL * k + 2LL * (k + 1) + 2LL * (P.nseg > 1 ? P.nseg - 1 : 0);
  size_t inv = 0;
  for (int i = 0; i < N; ++i) {
    const size_t e = inverse_workspace_elems(L.b[i], batch);
    if (e > inv) inv = e;
  }
  P.inv_elems = inv;
  if (precision == SINV_PREC_INT8X) {
    // largest product: bmax^3 with both operands distinct per instance
    GemmArgs g{};
    g.M = g.N = g.K = L.bmax;
    g.sA_out = g.sB_out = 1;
    g.inner = 1;
    g.outer = batch;
    P.i8x_bytes = gemm_i8x_workspace_bytes(g, kI8xDefaultSlices);
  }
  P.bytes = ((size_t)P.blocks * L.bmax * L.bmax * batch + P.inv_elems) * sizeof(cplx) +
            P.i8x_bytes + 512;
  return P;
}

}  // namespace

extern "C" size_t sinv_rgf_adjoint_workspace_bytes(const sinv_rgf_adjoint_desc* d) {
  if (!adj_desc_ok(d)) return 0;
  const BlockLayout L = make_block_layout(d->nblocks, d->blocksize, d->bsizes);
  if (!L.ok) return 0;
  return adj_plan(L, d->batch, d->checkpoint_every, d->precision).bytes;
}

extern "C" int sinv_rgf_adjoint_batched(const sinv_rgf_adjoint_desc* d, void* work,
                                        size_t work_bytes, void* stream) {
  if (!adj_desc_ok(d)) {
    set_error("sinv_rgf_adjoint_batched: invalid descriptor");
    return SINV_E_INVAL;
  }
  const BlockLayout Ly = make_block_layout(d->nblocks, d->blocksize, d->bsizes);
  if (!Ly.ok) {
    set_error("sinv_rgf_adjoint_batched: invalid block sizes");
    return SINV_E_INVAL;
  }
  const int N = Ly.N, nb = d->batch;
  const AdjPlan P = adj_plan(Ly, nb, d->checkpoint_every, d->precision);
  if (work == nullptr || work_bytes < P.bytes) {
    set_error("sinv_rgf_adjoint_batched: workspace too small (%zu < %zu)", work_bytes, P.bytes);
    return SINV_E_INVAL;
  }
  hipStream_t s = (hipStream_t)stream;
  const Prec prec = d->precision == SINV_PREC_MIXED ? Prec::F32 : Prec::F64;
  const bool i8x = d->precision == SINV_PREC_INT8X;
  const long long b2 = (long long)Ly.bmax * Ly.bmax;  // slot size
  cplx* base = (cplx*)(((uintptr_t)work + 255) & ~(uintptr_t)255);
  const long long WI = P.blocks * b2;  // instance stride in the workspace
  cplx* winv = base + WI * nb;
  void* wi8x = winv + P.inv_elems;
  auto wblk = [&](long long k) { return base + k * b2; };
  // temporaries
  cplx* Bd = wblk(0);
  cplx* dM = wblk(1);
  cplx* Wt = wblk(2);
  cplx* Xn = wblk(3);
  cplx* Qb = wblk(4);
  cplx* Sb = wblk(5);
  cplx* Gpp[2] = {wblk(6), wblk(7)};
  auto slot_g = [&](int i) { return wblk(kTemps + (i % P.k)); };
  auto slot_dg = [&](int i) { return wblk(kTemps + P.k + (i % P.k)); };
  auto slot_y = [&](int i) { return wblk(kTemps + 2LL * P.k + (i % (P.k + 1))); };
  auto slot_qf = [&](int i) { return wblk(kTemps + 3LL * P.k + 1 + (i % (P.k + 1))); };
  const long long ck0 = kTemps + 4LL * P.k + 2;
  auto ck_g = [&](int seg) { return wblk(ck0 + 2LL * (seg - 1)); };
  auto ck_dg = [&](int seg) { return wblk(ck0 + 2LL * (seg - 1) + 1); };

  const cplx* D = (const cplx*)d->diag;
  const cplx* U = (const cplx*)d->upper;
  const cplx* Lo = (const cplx*)d->lower;
  const cplx* Td = (const cplx*)d->gbar_diag;
  const cplx* Tu = (const cplx*)d->gbar_upper;
  const cplx* Tl = (const cplx*)d->gbar_lower;
  cplx* Ad = (cplx*)d->abar_diag;
  cplx* Au = (cplx*)d->abar_upper;
  cplx* Al = (cplx*)d->abar_lower;
  const long long AdI = Ly.dsize, AoI = Ly.osize;
  const cplx one{1.0, 0.0}, mone{-1.0, 0.0}, zero{0.0, 0.0};
  const cplx cds{d->dscale.re, -d->dscale.im};  // conj(dscale)
  const cplx cos_{d->oscale.re, -d->oscale.im};  // conj(oscale)
  const cplx mcos{-cos_.re, -cos_.im};
  auto bs = [&](int i) { return Ly.b[i]; };

  // operand: pointer of instance 0, instance stride, op, leading dimension
  struct Opd {
    const cplx* p;
    long long inst;
    int op;
    int ld;
  };
  auto ws = [&](const cplx* p, int ld) { return Opd{p, WI, OP_N, ld}; };
  // C (M x N, ldc) = alpha op(A) op(B) (+ C0 when given), contraction length K
  // reuse_a (int8x): op(A) is the previous product's op(A), whose digit planes
  // are still in the workspace
  auto gemm = [&](int M, int Nn, int K, Opd A, Opd B, cplx* C, long long cinst, int ldc,
                  cplx alpha, const cplx* C0, long long c0inst,
                  bool reuse_a = false) -> hipError_t {
    GemmArgs g{};
    g.M = M; g.N = Nn; g.K = K; g.opA = A.op; g.opB = B.op;
    g.A = A.p; g.lda = A.ld; g.sA_in = 0; g.sA_out = A.inst;
    g.B = B.p; g.ldb = B.ld; g.sB_in = 0; g.sB_out = B.inst;
    g.C = C; g.ldc = ldc; g.sC_in = 0; g.sC_out = cinst;
    g.C0 = C0; g.ldc0 = ldc; g.sC0_in = 0; g.sC0_out = c0inst;
    g.alpha = alpha;
    g.beta = C0 ? one : zero;
    g.shift = nullptr;
    g.inner = 1; g.outer = nb;
    if (i8x) return launch_gemm_i8x(g, kI8xDefaultSlices, wi8x, P.i8x_bytes, s, reuse_a);
    return launch_gemm(g, prec, s);
  };
  auto mm = [&](int M, int Nn, int K, Opd A, Opd B, cplx* C, long long cinst, int ldc, cplx alpha,
                bool acc, bool reuse_a = false) {
    return gemm(M, Nn, K, A, B, C, cinst, ldc, alpha, acc ? C : nullptr, cinst, reuse_a);
  };
  auto copy = [&](int r, int c, const cplx* src, long long sinst, cplx* dst, long long dinst) {
    return launch_axpby_shift(r, c, one, src, c, 0, sinst, dst, c, 0, dinst, nullptr, 1, nb, s);
  };
  // blocks of A (read conjugate-transposed as blocks of B = A^H) and of grad_G
  //   U[i] = A_{i,i+1} (b_i x b_{i+1});  U[i]^H = B_{i+1,i}
  //   Lo[i] = A_{i+1,i} (b_{i+1} x b_i); Lo[i]^H = B_{i,i+1}
  auto Ublk = [&](int i) { return Opd{U + Ly.ooff[i], d->upper_stride, OP_C, bs(i + 1)}; };
  auto Lblk = [&](int i) { return Opd{Lo + Ly.ooff[i], d->lower_stride, OP_C, bs(i)}; };
  auto Tdb = [&](int i) { return Opd{Td + Ly.doff[i], d->gbar_diag_stride, OP_N, bs(i)}; };
  auto Tub = [&](int i) { return Opd{Tu + Ly.ooff[i], d->gbar_upper_stride, OP_N, bs(i + 1)}; };
  auto Tlb = [&](int i) { return Opd{Tl + Ly.ooff[i], d->gbar_lower_stride, OP_N, bs(i)}; };

  // forward step i of RGF on B with tangents: (g_prev, dg_prev) -> (g_out, dg_out)
  auto fwd_step = [&](int i, cplx* gp, cplx* dgp, cplx* go, cplx* dgo) -> hipError_t {
    const int bi = bs(i);
    // blocks above the leaf size form M in Bd and invert it into g (the
    // elimination works in Bd); smaller ones invert in place
    const bool oop = bi > kLeaf;
    auto invert = [&](int tag) -> hipError_t {
      if (oop)
        return launch_inverse_to(Bd, bi, bi, 0, WI, 1, nb, go, bi, 0, WI, winv, d->info, tag, s);
      return launch_inverse(go, bi, bi, 0, WI, 1, nb, winv, d->info, tag, s);
    };
    // B_ii = conj(dscale) D_ii^H + conj(shift) I
    cplx* bdst = i == 0 && !oop ? go : Bd;
    SINV_HIP_RET(launch_transpose(bi, bi, OP_C, D + Ly.doff[i], bi, 0, d->diag_stride, bdst, bi,
                                  0, WI, 1, nb, s, cds, (const cplx*)d->shift, true));
    if (i == 0) {
      SINV_HIP_RET(invert(0));
      SINV_HIP_RET(mm(bi, bi, bi, ws(go, bi), Tdb(0), Wt, WI, bi, one, false));
      return mm(bi, bi, bi, ws(Wt, bi), ws(go, bi), dgo, WI, bi, mone, false);
    }
    const int bp = bs(i - 1);
    cplx* Y = slot_y(i);
    cplx* Qf = slot_qf(i);
    // Y = B_{i,i-1} g_{i-1}  (b_i x b_{i-1});  Qf = B_{i,i-1} dg_{i-1} + T_{i,i-1} g_{i-1}
    // (products with the same left operand back to back: in int8x mode its
    // digit planes are split once)
    SINV_HIP_RET(mm(bi, bp, bp, Ublk(i - 1), ws(gp, bp), Y, WI, bp, cos_, false));
    SINV_HIP_RET(mm(bi, bp, bp, Ublk(i - 1), ws(dgp, bp), Qf, WI, bp, cos_, false, true));
    SINV_HIP_RET(mm(bi, bp, bp, Tlb(i - 1), ws(gp, bp), Qf, WI, bp, one, true));
    // M = B_ii - Y B_{i-1,i};  dM = T_ii - Y T_{i-1,i} - Qf B_{i-1,i}
    SINV_HIP_RET(
        gemm(bi, bi, bp, ws(Y, bp), Lblk(i - 1), oop ? Bd : go, WI, bi, mcos, Bd, WI));
    SINV_HIP_RET(gemm(bi, bi, bp, ws(Y, bp), Tub(i - 1), dM, WI, bi, mone, Td + Ly.doff[i],
                      d->gbar_diag_stride, true));
    SINV_HIP_RET(mm(bi, bi, bp, ws(Qf, bp), Lblk(i - 1), dM, WI, bi, mcos, true));
    // g_i = M^-1, dg_i = -g_i dM g_i
    SINV_HIP_RET(invert(i));
    SINV_HIP_RET(mm(bi, bi, bi, ws(go, bi), ws(dM, bi), Wt, WI, bi, one, false));
    return mm(bi, bi, bi, ws(Wt, bi), ws(go, bi), dgo, WI, bi, mone, false);
  };

  if (d->info != nullptr) SINV_HIP_TRY(launch_fill_u64(d->info, 1, SINV_INFO_OK, s));

  // ---- forward sweep, keeping the last segment and the segment checkpoints
  for (int i = 0; i < N; ++i) {
    cplx* gp = i > 0 ? slot_g(i - 1) : nullptr;
    cplx* dgp = i > 0 ? slot_dg(i - 1) : nullptr;
    SINV_HIP_TRY(fwd_step(i, gp, dgp, slot_g(i), slot_dg(i)));
    if (i % P.k == P.k - 1 && i + 1 < N) {
      const int seg = (i + 1) / P.k;
      SINV_HIP_TRY(copy(bs(i), bs(i), slot_g(i), WI, ck_g(seg), WI));
      SINV_HIP_TRY(copy(bs(i), bs(i), slot_dg(i), WI, ck_dg(seg), WI));
    }
  }

  // ---- backward sweep (rgf.py:51-62 on B, with tangents)
  int cur = 0;
  const int bl = bs(N - 1);
  SINV_HIP_TRY(copy(bl, bl, slot_g(N - 1), WI, Gpp[0], WI));                  // G_{N-1,N-1}(B)
  SINV_HIP_TRY(copy(bl, bl, slot_dg(N - 1), WI, Ad + Ly.doff[N - 1], AdI));  // its tangent
  cplx* Gout = (cplx*)d->g_diag;  // optional G_ii = (G_ii of B)^H
  if (Gout != nullptr)
    SINV_HIP_TRY(launch_transpose(bl, bl, OP_C, Gpp[0], bl, 0, WI, Gout + Ly.doff[N - 1], bl, 0,
                                  AdI, 1, nb, s));
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
      const int bi = bs(i), bj = bs(i + 1);
      cplx* g = slot_g(i);
      cplx* dg = slot_dg(i);
      cplx* Gn = Gpp[cur];
      cplx* Gi = Gpp[cur ^ 1];
      const Opd dGn{Ad + Ly.doff[i + 1], AdI, OP_N, bj};
      cplx* Ali = Al + Ly.ooff[i];  // b_{i+1} x b_i
      cplx* Aui = Au + Ly.ooff[i];  // b_i x b_{i+1}
      cplx* Adi = Ad + Ly.doff[i];  // b_i x b_i
      // the forward step i+1's Y = B_{i+1,i} g and Qf = T_{i+1,i} g + B_{i+1,i} dg
      const Opd Yn = ws(slot_y(i + 1), bi), Qn = ws(slot_qf(i + 1), bi);
      // Products with the same left operand run back to back (in int8x mode
      // its digit planes are split once: Gn, g, Q and S each once).
      // Xn = -Gn Y;  tangent of G_{i+1,i} = -Gn B_{i+1,i} g:  -Gn Qf - dGn Y
      SINV_HIP_TRY(mm(bj, bi, bj, ws(Gn, bj), Yn, Xn, WI, bi, mone, false));
      SINV_HIP_TRY(mm(bj, bi, bj, ws(Gn, bj), Qn, Ali, AoI, bi, mone, false, true));
      SINV_HIP_TRY(mm(bj, bi, bj, dGn, Yn, Ali, AoI, bi, mone, true));
      // Q = g B_{i,i+1};  S = g T_{i,i+1} + dg B_{i,i+1}
      SINV_HIP_TRY(mm(bi, bj, bi, ws(g, bi), Lblk(i), Qb, WI, bj, cos_, false));
      SINV_HIP_TRY(mm(bi, bj, bi, ws(g, bi), Tub(i), Sb, WI, bj, one, false, true));
      SINV_HIP_TRY(mm(bi, bj, bi, ws(dg, bi), Lblk(i), Sb, WI, bj, cos_, true));
      // G_ii = g - Q Xn;  tangent of G_{i,i+1}: -Q dGn - S Gn;
      // tangent of G_ii: dg - Q dXn - S Xn
      SINV_HIP_TRY(gemm(bi, bi, bj, ws(Qb, bj), ws(Xn, bi), Gi, WI, bi, mone, g, WI));
      SINV_HIP_TRY(mm(bi, bj, bj, ws(Qb, bj), dGn, Aui, AoI, bj, mone, false, true));
      SINV_HIP_TRY(gemm(bi, bi, bj, ws(Qb, bj), Opd{Ali, AoI, OP_N, bi}, Adi, AdI, bi, mone, dg,
                        WI, true));
      SINV_HIP_TRY(mm(bi, bj, bj, ws(Sb, bj), ws(Gn, bj), Aui, AoI, bj, mone, true));
      SINV_HIP_TRY(mm(bi, bi, bj, ws(Sb, bj), ws(Xn, bi), Adi, AdI, bi, mone, true, true));
      if (Gout != nullptr)
        SINV_HIP_TRY(launch_transpose(bi, bi, OP_C, Gi, bi, 0, WI, Gout + Ly.doff[i], bi, 0, AdI,
                                      1, nb, s));
      cur ^= 1;
    }
  }
  return SINV_OK;
}
