// BCR-serial selected inversion driver, batched over instances.
//
// Follows src/sinv/algorithms/bcr/bcr_serial.py:19-466 (reduce :66-169,
// reduce_bcr :173-216, invert_block :220-249, corner_produce :253-304,
// center_produce :308-383, produce_bcr :387-466) on block-compressed storage:
//
//  * Level l (s = 2^l) works on the active rows r_m = m*s + s - 1 < N,
//    m = 0..n_l-1.  Even m are eliminated, odd m are kept; the reduced system
//    of a level is block tridiagonal over its active rows, with couplings
//    Up_l[c] = A[r_c, r_{c+1}], Lo_l[c] = A[r_{c+1}, r_c].
//  * The identity padding to 2^p - 1 blocks (bcr_utils.py:35-71) is skipped:
//    padded rows have zero coupling, so every term they contribute to a real
//    row is an exact zero.
//  * inv(A_tt) is computed once per eliminated row and reused by the
//    production (the reference inverts it up to three times).
//  * Every operation of a level is one batched launch over (rows x
//    instances): rows of a level are independent, so each level is a handful
//    of wide GEMMs -- the log-depth parallelism RGF lacks.
//
// Per coupling c of level l, one end is eliminated (t) and the other kept (j):
//    c even: t = c,   j = c+1:  U_c = inv(A_tt) Up_l[c],  L_c = Lo_l[c] inv(A_tt)
//    c odd:  t = c+1, j = c:    U_c = inv(A_tt) Lo_l[c],  L_c = Up_l[c] inv(A_tt)
// so U_c = U[t,j] and L_c = L[j,t] of the reference.
#include <vector>

#include "sinv_core.hpp"

namespace sinv {

namespace {

struct Level {
  int s = 1, n = 0;          // stride, active rows
  long long up = 0, lo = 0;  // couplings (n-1 blocks each), level >= 1 in workspace
  long long U = 0, L = 0;    // factors per coupling (n-1 blocks each)
  long long ainv = 0;        // inv(A_tt) of eliminated rows (ceil(n/2) blocks)
  long long gup = 0, glo = 0;  // produced G couplings of level >= 1 ((n-1) each)
};

struct Plan {
  int N = 0, b = 0, height = 0, root = 0;
  std::vector<Level> lv;     // levels 0..height (height has n == 1)
  long long dwork = 0;       // N diagonal blocks
  long long per_inst = 0;    // workspace blocks per instance
  long long fill_blocks = 0; // G blocks of levels >= 1 (per instance)
  size_t inv_elems = 0;
  size_t bytes = 0;
};

Plan make_plan(int N, int b, int batch) {
  Plan P;
  P.N = N;
  P.b = b;
  // padded size 2^p - 1 >= N (bcr_utils.py:35-38); height = floor(log2(2^p - 1)) = p - 1
  int p = 0;
  while ((1LL << p) - 1 < N) ++p;
  P.height = p - 1;
  P.root = (1 << P.height) - 1;
  long long off = 0;
  const long long b2 = (long long)b * b;
  P.dwork = off;
  off += N;
  P.lv.resize(P.height + 1);
  int maxelim = 0;
  for (int l = 0; l <= P.height; ++l) {
    Level& L = P.lv[l];
    L.s = 1 << l;
    L.n = N / L.s;  // active rows r = m*s + s - 1 < N
    const int nc = L.n > 1 ? L.n - 1 : 0;
    if (l >= 1) {
      L.up = off; off += nc;
      L.lo = off; off += nc;
      L.gup = off; off += nc;
      L.glo = off; off += nc;
      P.fill_blocks += 2LL * nc;
    }
    if (l < P.height) {
      L.U = off; off += nc;
      L.L = off; off += nc;
      L.ainv = off; off += (L.n + 1) / 2;
      if ((L.n + 1) / 2 > maxelim) maxelim = (L.n + 1) / 2;
    }
  }
  P.per_inst = off;
  const bool newton = inverse_method() == SINV_INV_NEWTON;
  P.inv_elems = inverse_workspace_elems(b, batch * (maxelim > 0 ? maxelim : 1), newton);
  // root inverse runs with batch instances only
  const size_t root_inv = inverse_workspace_elems(b, batch, newton);
  if (root_inv > P.inv_elems) P.inv_elems = root_inv;
  P.bytes = ((size_t)P.per_inst * (size_t)b2 * (size_t)batch + P.inv_elems) * sizeof(cplx) + 512;
  return P;
}

}  // namespace

}  // namespace sinv

using namespace sinv;

namespace {

bool bcr_desc_ok(const sinv_bcr_desc* d) {
  if (d == nullptr || d->nblocks <= 0 || d->blocksize <= 0 || d->batch <= 0) return false;
  if (d->diag == nullptr || d->g_diag == nullptr) return false;
  if (d->nblocks > 1 && (d->upper == nullptr || d->lower == nullptr || d->g_upper == nullptr ||
                         d->g_lower == nullptr))
    return false;
  if (d->diag_stride < 0 || d->upper_stride < 0 || d->lower_stride < 0) return false;
  return true;
}

}  // namespace

extern "C" size_t sinv_bcr_workspace_bytes(const sinv_bcr_desc* d) {
  if (!bcr_desc_ok(d)) return 0;
  return make_plan(d->nblocks, d->blocksize, d->batch).bytes;
}

extern "C" long long sinv_bcr_fill_blocks(int nblocks) {
  if (nblocks <= 0) return 0;
  return make_plan(nblocks, 1, 1).fill_blocks;
}

namespace {

// The BCR driver.  npart > 1 splits every level operation into one launch per
// partition: partition p owns the padded block rows [p * stride, (p + 1) *
// stride) (the last one the rest; bcr_utils.py:133-184 with stride =
// (2^p - 1) / npart), and issues the operations anchored on its rows --
// reducing its kept rows (their factors U / L, the A_jj update and the next
// level's fill-in couplings; bcr_parallel.py:227-256), inverting its
// eliminated rows, and producing its rows (bcr_parallel.py:1004-1015).  Each
// operation family of a level is issued for every partition before the next
// family, as the ranks exchange halo rows between them; the arithmetic and
// its order per block are those of npart = 1, so results are bitwise equal.
int bcr_run(const sinv_bcr_desc* d, int npart, void* work, size_t work_bytes, void* stream) {
  const int N = d->nblocks, b = d->blocksize, nb = d->batch;
  const Plan P = make_plan(N, b, nb);
  if (work == nullptr || work_bytes < P.bytes) {
    set_error("sinv_bcr_batched: workspace too small (%zu < %zu)", work_bytes, P.bytes);
    return SINV_E_INVAL;
  }
  const long long npad = (1LL << (P.height + 1)) - 1;  // padded block rows
  if (npart < 1 || npart > npad) {
    set_error("sinv_bcr_parallel_batched: npart = %d must be in [1, %lld]", npart, npad);
    return SINV_E_INVAL;
  }
  const long long pstride = npad / npart;
  auto owner = [&](long long row) {
    const long long o = row / pstride;
    return (int)(o < npart ? o : npart - 1);
  };
  hipStream_t s = (hipStream_t)stream;
  const long long b2 = (long long)b * b;
  cplx* base = (cplx*)(((uintptr_t)work + 255) & ~(uintptr_t)255);
  cplx* W = base;                                // per-instance blocks, instance stride per_inst
  cplx* winv = base + P.per_inst * b2 * nb;      // inverse recursion scratch
  const long long WI = P.per_inst * b2;          // instance stride of W
  const bool newton = inverse_method() == SINV_INV_NEWTON;  // as in make_plan

  const cplx one{1.0, 0.0}, mone{-1.0, 0.0}, zero{0.0, 0.0};
  const cplx osc{d->oscale.re, d->oscale.im};
  const cplx mosc{-osc.re, -osc.im};
  cplx* Gd = (cplx*)d->g_diag;
  cplx* Gu = (cplx*)d->g_upper;
  cplx* Gl = (cplx*)d->g_lower;
  cplx* Gf = (cplx*)d->g_fill;  // optional: G couplings of levels >= 1
  const long long GdI = (long long)N * b2, GoI = (long long)(N > 1 ? N - 1 : 0) * b2;
  const long long GfI = P.fill_blocks * b2;

  // block-array view: element pointer of block k of instance 0 and the instance stride
  struct Arr {
    const cplx* p;
    long long inst;
  };
  auto warr = [&](long long blk) { return Arr{W + blk * b2, WI}; };
  // C[k] (count rows, block steps) = alpha A[k] B[k] + beta C0[k]
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
    return launch_gemm(g, Prec::F64, s);
  };
  // Index range [0, count) of an operation family split into the runs owned by
  // one partition (anchor(k) = block row of operation k, increasing in k);
  // issue(k0, k1) launches one run.
  auto split = [&](int count, auto anchor, auto issue) -> hipError_t {
    int k0 = 0;
    while (k0 < count) {
      const int o = owner(anchor(k0));
      int k1 = k0 + 1;
      while (k1 < count && owner(anchor(k1)) == o) ++k1;
      SINV_HIP_RET(issue(k0, k1));
      k0 = k1;
    }
    return hipSuccess;
  };
  // gemm over index k of one family, run by run: operand offsets advance by
  // their block steps
  auto pgemm = [&](auto anchor, int count, Arr A, long long a0, long long as, Arr B, long long b0,
                   long long bs, Arr C, long long c0, long long cs, cplx alpha, bool acc,
                   Arr C0 = Arr{nullptr, 0}, long long c00 = 0, long long c0s = 0) {
    return split(count, anchor, [&](int k0, int k1) {
      return gemm(k1 - k0, A, a0 + k0 * as, as, B, b0 + k0 * bs, bs, C, c0 + k0 * cs, cs, alpha,
                  acc, C0, c00 + k0 * c0s, c0s);
    });
  };

  if (d->info != nullptr) SINV_HIP_TRY(launch_fill_u64(d->info, 1, SINV_INFO_OK, s));

  // Dwork = dscale * D + shift I     (bcr_serial.py works on a padded copy, bcr_utils.py:66-69)
  {
    const cplx ds{d->dscale.re, d->dscale.im};
    SINV_HIP_TRY(launch_axpby_shift(b, b, ds, (const cplx*)d->diag, b, b2, d->diag_stride,
                                    W + P.dwork * b2, b, b2, WI, (const cplx*)d->shift, N, nb,
                                    s));
  }
  const Arr Dw = warr(P.dwork);

  // coupling arrays of each level (level 0 reads the caller's blocks, scaled by oscale)
  auto up_arr = [&](int l) {
    return l == 0 ? Arr{(const cplx*)d->upper, d->upper_stride} : warr(P.lv[l].up);
  };
  auto lo_arr = [&](int l) {
    return l == 0 ? Arr{(const cplx*)d->lower, d->lower_stride} : warr(P.lv[l].lo);
  };

  // ---------------------------------------------------------------- reduction
  for (int l = 0; l < P.height; ++l) {
    const Level& Lv = P.lv[l];
    const int n = Lv.n, sl = Lv.s;
    const int nelim = (n + 1) / 2, nkeep = n / 2;
    const cplx cs = l == 0 ? osc : one, mcs = l == 0 ? mosc : mone;
    const Arr Up = up_arr(l), Lo = lo_arr(l);
    const Arr U = warr(Lv.U), Lf = warr(Lv.L), Ai = warr(Lv.ainv);
    // block row of active index m at this level
    auto rowm = [=](long long m) { return m * sl + sl - 1; };
    auto elim = [=](int q) { return rowm(2LL * q); };      // eliminated row m = 2q
    auto kept = [=](int q) { return rowm(2LL * q + 1); };  // kept row m = 2q+1
    // inv(A_tt) for eliminated rows r = 2q*s + s - 1  (bcr_serial.py:121,148);
    // tag = row of the first eliminated block of the run; instances report z
    SINV_HIP_TRY(split(nelim, elim, [&](int q0, int q1) -> hipError_t {
      SINV_HIP_RET(launch_axpby_shift(b, b, one, Dw.p + (sl - 1 + 2LL * sl * q0) * b2, b,
                                      2LL * sl * b2, WI, W + (Lv.ainv + q0) * b2, b, b2, WI,
                                      nullptr, q1 - q0, nb, s));
      return launch_inverse(W + (Lv.ainv + q0) * b2, b, b, b2, WI, q1 - q0, nb, winv, d->info,
                            (int)elim(q0), s, newton);
    }));
    const int ce = (n - 1 + 1) / 2;  // even couplings c = 2q < n-1
    const int co = (n - 1) / 2;      // odd couplings c = 2q+1 < n-1
    // c even: U_c = Ainv[q] Up[c], L_c = Lo[c] Ainv[q]      (bcr_serial.py:124-128)
    //   (computed by the partition of the kept row c + 1)
    SINV_HIP_TRY(pgemm(kept, ce, Ai, 0, 1, Up, 0, 2, U, 0, 2, cs, false));
    SINV_HIP_TRY(pgemm(kept, ce, Lo, 0, 2, Ai, 0, 1, Lf, 0, 2, cs, false));
    // c odd: U_c = Ainv[q+1] Lo[c], L_c = Up[c] Ainv[q+1]   (bcr_serial.py:150-154)
    //   (kept row c)
    SINV_HIP_TRY(pgemm(kept, co, Ai, 1, 1, Lo, 1, 2, U, 1, 2, cs, false));
    SINV_HIP_TRY(pgemm(kept, co, Up, 1, 2, Ai, 1, 1, Lf, 1, 2, cs, false));
    // kept rows m = 2q+1 (r = 2(q+1)s - 1): A_jj -= L_{m-1} Up[m-1]   (bcr_serial.py:130-133)
    SINV_HIP_TRY(pgemm(kept, nkeep, Lf, 0, 2, Up, 0, 2,
                       Arr{Dw.p + (long long)(2 * sl - 1) * b2, WI}, 0, 2LL * sl, mcs, true));
    //   then A_jj -= L_m Lo[m] for m <= n-2                           (bcr_serial.py:156-159)
    SINV_HIP_TRY(pgemm(kept, co, Lf, 1, 2, Lo, 1, 2, Arr{Dw.p + (long long)(2 * sl - 1) * b2, WI},
                       0, 2LL * sl, mcs, true));
    // next level's couplings (fill-in, bcr_serial.py:137-143, 163-169)
    const Level& Nx = P.lv[l + 1];
    const int ncn = Nx.n > 1 ? Nx.n - 1 : 0;
    //   Up_{l+1}[c'] = -L_{2c'+1} Up[2c'+2]      (kept row 2c'+1)
    SINV_HIP_TRY(pgemm(kept, ncn, Lf, 1, 2, Up, 2, 2, warr(Nx.up), 0, 1, mcs, false));
    //   Lo_{l+1}[c'] = -L_{2c'+2} Lo[2c'+1]      (kept row 2c'+3)
    SINV_HIP_TRY(pgemm([=](int q) { return kept(q + 1); }, ncn, Lf, 2, 2, Lo, 1, 2, warr(Nx.lo),
                       0, 1, mcs, false));
  }

  // ---------------------------------------------------------------- root (bcr_serial.py:248-249)
  SINV_HIP_TRY(launch_axpby_shift(b, b, one, Dw.p + (long long)P.root * b2, b, 0, WI,
                                  Gd + (long long)P.root * b2, b, 0, GdI, nullptr, 1, nb, s));
  SINV_HIP_TRY(launch_inverse(Gd + (long long)P.root * b2, b, b, 0, GdI, 1, nb, winv, d->info,
                              P.root, s, newton));

  // ---------------------------------------------------------------- production
  const Arr GD{Gd, GdI};
  // G couplings of level l >= 1 live in the workspace; copied to g_fill at the end
  auto GUP = [&](int l) { return l == 0 ? Arr{Gu, GoI} : warr(P.lv[l].gup); };
  auto GLO = [&](int l) { return l == 0 ? Arr{Gl, GoI} : warr(P.lv[l].glo); };

  for (int l = P.height - 1; l >= 0; --l) {
    const Level& Lv = P.lv[l];
    const int n = Lv.n, sl = Lv.s;
    const Arr U = warr(Lv.U), Lf = warr(Lv.L), Ai = warr(Lv.ainv);
    const Arr Gup = GUP(l), Glo = GLO(l), Gup1 = GUP(l + 1), Glo1 = GLO(l + 1);
    // Gd row of active index m at this level: r = m*s + s - 1
    auto gd = [&](long long m) { return (m * sl + sl - 1); };

    // -- corner from below, m = 0 (t = r_0, f = r_1)        (bcr_serial.py:293-304)
    {
      // G[f,t] = Glo[0] = -G_ff L_0
      SINV_HIP_TRY(gemm(1, GD, gd(1), 0, Lf, 0, 0, Glo, 0, 0, mone, false));
      // G[t,f] = Gup[0] = -U_0 G_ff
      SINV_HIP_TRY(gemm(1, U, 0, 0, GD, gd(1), 0, Gup, 0, 0, mone, false));
      // G[t,t] = Ainv[0] - Gup[0] L_0
      SINV_HIP_TRY(gemm(1, Gup, 0, 0, Lf, 0, 0, GD, gd(0), 0, mone, true, Ai, 0, 0));
    }
    // -- centers m = 2q, q = 1..nc (a = m-1, c = m+1)        (bcr_serial.py:354-383)
    //   (produced by the partition of row t = r_m, m = 2k + 2 for run index k)
    const int ncen = (n - 2) / 2;  // even m in [2, n-2]
    auto tcen = [=](int k) { return gd(2LL * k + 2); };
    if (ncen > 0) {
      // G[a,t] = Gup[m-1] = -G_aa L_{m-1} - G_ac L_m      with G_ac = Gup1[q-1]
      SINV_HIP_TRY(pgemm(tcen, ncen, GD, gd(1), 2LL * sl, Lf, 1, 2, Gup, 1, 2, mone, false));
      SINV_HIP_TRY(pgemm(tcen, ncen, Gup1, 0, 1, Lf, 2, 2, Gup, 1, 2, mone, true));
      // G[c,t] = Glo[m] = -G_ca L_{m-1} - G_cc L_m        with G_ca = Glo1[q-1]
      SINV_HIP_TRY(pgemm(tcen, ncen, Glo1, 0, 1, Lf, 1, 2, Glo, 2, 2, mone, false));
      SINV_HIP_TRY(pgemm(tcen, ncen, GD, gd(3), 2LL * sl, Lf, 2, 2, Glo, 2, 2, mone, true));
      // G[t,a] = Glo[m-1] = -U_{m-1} G_aa - U_m G_ca
      SINV_HIP_TRY(pgemm(tcen, ncen, U, 1, 2, GD, gd(1), 2LL * sl, Glo, 1, 2, mone, false));
      SINV_HIP_TRY(pgemm(tcen, ncen, U, 2, 2, Glo1, 0, 1, Glo, 1, 2, mone, true));
      // G[t,c] = Gup[m] = -U_{m-1} G_ac - U_m G_cc
      SINV_HIP_TRY(pgemm(tcen, ncen, U, 1, 2, Gup1, 0, 1, Gup, 2, 2, mone, false));
      SINV_HIP_TRY(pgemm(tcen, ncen, U, 2, 2, GD, gd(3), 2LL * sl, Gup, 2, 2, mone, true));
      // G[t,t] = Ainv[q] - Glo[m-1] L_{m-1} - Gup[m] L_m
      SINV_HIP_TRY(pgemm(tcen, ncen, Glo, 1, 2, Lf, 1, 2, GD, gd(2), 2LL * sl, mone, true, Ai, 1,
                         1));
      SINV_HIP_TRY(pgemm(tcen, ncen, Gup, 2, 2, Lf, 2, 2, GD, gd(2), 2LL * sl, mone, true));
    }
    // -- corner from above: last eliminated row m = n-1 when n is odd and n >= 3
    if (n % 2 == 1 && n >= 3) {
      const long long m = n - 1;
      // G[a,t] = Gup[m-1] = -G_aa L_{m-1}
      SINV_HIP_TRY(gemm(1, GD, gd(m - 1), 0, Lf, m - 1, 0, Gup, m - 1, 0, mone, false));
      // G[t,a] = Glo[m-1] = -U_{m-1} G_aa
      SINV_HIP_TRY(gemm(1, U, m - 1, 0, GD, gd(m - 1), 0, Glo, m - 1, 0, mone, false));
      // G[t,t] = Ainv[m/2] - Glo[m-1] L_{m-1}
      SINV_HIP_TRY(gemm(1, Glo, m - 1, 0, Lf, m - 1, 0, GD, gd(m), 0, mone, true, Ai, m / 2, 0));
    }
  }

  // optional: hand out the produced couplings of levels >= 1 (the stride-2^l blocks of the
  // reference's dense output), instance-major, level by level, Gup_l then Glo_l
  if (Gf != nullptr && P.fill_blocks > 0) {
    long long o = 0;
    for (int l = 1; l <= P.height; ++l) {
      const int nc = P.lv[l].n > 1 ? P.lv[l].n - 1 : 0;
      if (nc == 0) continue;
      SINV_HIP_TRY(launch_axpby_shift(b, b, one, W + P.lv[l].gup * b2, b, b2, WI, Gf + o * b2, b,
                                      b2, GfI, nullptr, nc, nb, s));
      o += nc;
      SINV_HIP_TRY(launch_axpby_shift(b, b, one, W + P.lv[l].glo * b2, b, b2, WI, Gf + o * b2, b,
                                      b2, GfI, nullptr, nc, nb, s));
      o += nc;
    }
  }
  return SINV_OK;
}

}  // namespace

extern "C" int sinv_bcr_batched(const sinv_bcr_desc* d, void* work, size_t work_bytes,
                                void* stream) {
  if (!bcr_desc_ok(d)) {
    set_error("sinv_bcr_batched: invalid descriptor");
    return SINV_E_INVAL;
  }
  return bcr_run(d, 1, work, work_bytes, stream);
}

extern "C" int sinv_bcr_parallel_batched(const sinv_bcr_desc* d, int npart, void* work,
                                         size_t work_bytes, void* stream) {
  if (!bcr_desc_ok(d)) {
    set_error("sinv_bcr_parallel_batched: invalid descriptor");
    return SINV_E_INVAL;
  }
  return bcr_run(d, npart, work, work_bytes, stream);
}
