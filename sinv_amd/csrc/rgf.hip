// RGF selected inversion driver (host side of the hot path), batched over
// instances.  Follows the recurrences of src/sinv/algorithms/rgf/rgf.py:40-62:
//
//   forward  (rgf.py:43-48):  g_0 = A_00^-1
//                             g_i = (A_ii - A_{i,i-1} g_{i-1} A_{i-1,i})^-1
//   backward (rgf.py:50-62):  X_i      = G_{i+1,i+1} A_{i+1,i} g_i
//                             G_{i+1,i} = -X_i
//                             G_{i,i+1} = G_{i+1,i}^T (sym) | -g_i A_{i,i+1} G_{i+1,i+1}
//                             G_ii      = g_i + g_i A_{i,i+1} X_i
//
// The forward sweep's T_{i+1} = A_{i+1,i} g_i is kept (in the G_{i,i+1}
// output slot, unused until the backward step that overwrites it), so X_i is
// one product, G_{i+1,i+1} T_{i+1}: six block products per block instead of
// seven (five with sym_mat).  The reference evaluates (G_{i+1,i+1} A_{i+1,i}) g_i;
// the other association agrees to rounding.  Without saved off-diagonals
// there is no slot, and X_i is formed as (G_{i+1,i+1} A_{i+1,i}) g_i.
//
// Precision SINV_PREC_MIXED_REFINED runs the forward Schur products in
// complex64 and refines each g~_i = inv(M~_i) once in FP64 against M_i formed
// in FP64: R = M_i g~_i - I, g_i = g~_i - g~_i R; the backward sweep is FP64.
//
// Every step is one batched launch over all instances (energies).  The
// product g_i A_{i,i+1} is formed once and shared by G_{i,i+1} and G_ii (the
// reference forms it twice), and G_ii is accumulated in place over g_i.  A(E)
// is never materialised: A_ii = dscale D_ii + shift_z I is produced by the
// Schur GEMM's epilogue, and oscale folds into the GEMM alphas.
#include <cstdarg>
#include <cstdio>
#include <mutex>
#include <vector>

#include "sinv_core.hpp"

namespace sinv {

namespace {

thread_local char g_err[512] = "";

using Layout = BlockLayout;

Layout make_layout(int N, int blocksize, const int* bsizes) {
  return make_block_layout(N, blocksize, bsizes);
}

struct WorkPlan {
  size_t tmp_elems = 0;  // per-instance temporaries P, Q, X (each bmax^2)
  size_t inv_elems = 0;  // inverse workspace for the whole batch
  size_t i8x_bytes = 0;  // digit planes of the sliced-int8 GEMMs (SINV_PREC_INT8X)
  size_t total_bytes = 0;
};

WorkPlan plan_work(const Layout& L, int batch, int precision) {
  WorkPlan w;
  const size_t b2 = (size_t)L.bmax * L.bmax;
  w.tmp_elems = 3 * b2 * (size_t)batch;
  size_t inv = 0;
  const bool newton = inverse_method() == SINV_INV_NEWTON;
  for (int i = 0; i < L.N; ++i) {
    const size_t e = inverse_workspace_elems(L.b[i], batch, newton);
    if (e > inv) inv = e;
  }
  w.inv_elems = inv;
  if (precision == SINV_PREC_INT8X) {
    // largest sweep product: bmax^3 with both operands distinct per instance
    GemmArgs g{};
    g.M = g.N = g.K = L.bmax;
    g.sA_out = g.sB_out = 1;
    g.inner = 1;
    g.outer = batch;
    w.i8x_bytes = gemm_i8x_workspace_bytes(g, kI8xDefaultSlices);
  }
  // -1 per instance: the shift of the refinement residual M g~ - I
  w.total_bytes = (w.tmp_elems + w.inv_elems + (size_t)batch) * sizeof(cplx) + w.i8x_bytes + 256;
  return w;
}

bool desc_valid(const sinv_rgf_desc* d) {
  if (d == nullptr) return false;
  if (d->nblocks <= 0 || d->batch <= 0) return false;
  if (d->bsizes == nullptr && d->blocksize <= 0) return false;
  if (d->precision != SINV_PREC_FP64 && d->precision != SINV_PREC_MIXED &&
      d->precision != SINV_PREC_INT8X && d->precision != SINV_PREC_MIXED_REFINED)
    return false;
  if (d->diag == nullptr || d->g_diag == nullptr) return false;
  if (d->nblocks > 1 && (d->upper == nullptr || d->lower == nullptr)) return false;
  if (d->nblocks > 1 && d->save_off_diag && (d->g_upper == nullptr || d->g_lower == nullptr))
    return false;
  if (d->diag_stride < 0 || d->upper_stride < 0 || d->lower_stride < 0) return false;
  return true;
}

}  // namespace

void set_error(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
}

const char* last_error() { return g_err; }

}  // namespace sinv

using namespace sinv;

extern "C" int sinv_version(void) { return 100; }

extern "C" const char* sinv_last_error(void) { return sinv::last_error(); }

extern "C" size_t sinv_rgf_workspace_bytes(const sinv_rgf_desc* d) {
  if (!desc_valid(d)) return 0;
  const Layout L = make_layout(d->nblocks, d->blocksize, d->bsizes);
  if (!L.ok) return 0;
  return plan_work(L, d->batch, d->precision).total_bytes;
}

extern "C" int sinv_rgf_batched(const sinv_rgf_desc* d, void* work, size_t work_bytes,
                                void* stream) {
  if (!desc_valid(d)) {
    set_error("sinv_rgf_batched: invalid descriptor");
    return SINV_E_INVAL;
  }
  const Layout L = make_layout(d->nblocks, d->blocksize, d->bsizes);
  if (!L.ok) {
    set_error("sinv_rgf_batched: invalid block sizes");
    return SINV_E_INVAL;
  }
  const WorkPlan wp = plan_work(L, d->batch, d->precision);
  if (work == nullptr || work_bytes < wp.total_bytes) {
    set_error("sinv_rgf_batched: workspace too small (%zu < %zu)", work_bytes, wp.total_bytes);
    return SINV_E_INVAL;
  }
  hipStream_t s = (hipStream_t)stream;
  const int N = L.N, nb = d->batch;
  const bool refine = d->precision == SINV_PREC_MIXED_REFINED;
  // forward Schur products; the backward sweep is FP64 when refining
  const Prec prec = d->precision == SINV_PREC_MIXED || refine ? Prec::F32 : Prec::F64;
  const Prec bprec = refine ? Prec::F64 : prec;
  const bool i8x = d->precision == SINV_PREC_INT8X;
  const long long b2 = (long long)L.bmax * L.bmax;

  const cplx* D = (const cplx*)d->diag;
  const cplx* U = (const cplx*)d->upper;
  const cplx* Lo = (const cplx*)d->lower;
  cplx* Gd = (cplx*)d->g_diag;
  cplx* Gu = (cplx*)d->g_upper;
  cplx* Gl = (cplx*)d->g_lower;
  const cplx* shift = (const cplx*)d->shift;
  const cplx dscale{d->dscale.re, d->dscale.im};
  const cplx oscale{d->oscale.re, d->oscale.im};
  const cplx osq{oscale.re * oscale.re - oscale.im * oscale.im, 2.0 * oscale.re * oscale.im};
  const cplx one{1.0, 0.0}, mone{-1.0, 0.0}, zero{0.0, 0.0};
  const cplx mosq{-osq.re, -osq.im};

  cplx* base = (cplx*)(((uintptr_t)work + 255) & ~(uintptr_t)255);
  cplx* P = base;                     // bmax^2 per instance
  cplx* Q = P + b2 * nb;
  cplx* Xw = Q + b2 * nb;
  cplx* winv = Xw + b2 * nb;
  cplx* mones = winv + wp.inv_elems;  // nb entries of -1 (refinement only)
  void* wi8x = mones + nb;

  // one GEMM launch over all instances: instance z is the outer batch index
  auto gemm = [&](int M, int Nn, int K, const cplx* A, int lda, long long sA, const cplx* B,
                  int ldb, long long sB, cplx* C, int ldc, long long sC, cplx alpha,
                  const cplx* C0, int ldc0, long long sC0, cplx beta, const cplx* sh,
                  Prec p, bool reuse_a = false) -> hipError_t {
    GemmArgs g{};
    g.M = M; g.N = Nn; g.K = K; g.opA = OP_N; g.opB = OP_N;
    g.A = A; g.lda = lda; g.sA_in = 0; g.sA_out = sA;
    g.B = B; g.ldb = ldb; g.sB_in = 0; g.sB_out = sB;
    g.C0 = C0; g.ldc0 = ldc0; g.sC0_in = 0; g.sC0_out = sC0;
    g.C = C; g.ldc = ldc; g.sC_in = 0; g.sC_out = sC;
    g.alpha = alpha; g.beta = beta; g.shift = sh;
    g.inner = 1; g.outer = nb;
    if (i8x) return launch_gemm_i8x(g, kI8xDefaultSlices, wi8x, wp.i8x_bytes, s, reuse_a);
    return launch_gemm(g, p, s);
  };

  if (d->info != nullptr) SINV_HIP_TRY(launch_fill_u64(d->info, 1, SINV_INFO_OK, s));
  if (refine) SINV_HIP_TRY(launch_fill(mones, nb, cplx{-1.0, 0.0}, s));

  // ---- forward sweep (rgf.py:44-48).  Blocks above the leaf size form M_i in
  // the temporary Q and invert it into the g_i slot (the elimination works in
  // Q, so no copy of M_i is needed); smaller ones invert in place.
  auto m_dst = [&](int i, cplx*& dst, long long& sdst) {
    const bool oop = L.b[i] > kLeaf;
    dst = oop ? Q : Gd + L.doff[i];
    sdst = oop ? b2 : L.dsize;
  };
  const bool newton = inverse_method() == SINV_INV_NEWTON;
  auto invert_m = [&](int i) -> hipError_t {
    const int bi = L.b[i];
    if (bi > kLeaf)
      return launch_inverse_to(Q, bi, bi, 0, b2, 1, nb, Gd + L.doff[i], bi, 0, L.dsize, winv,
                               d->info, i, s, newton);
    return launch_inverse(Gd + L.doff[i], bi, bi, 0, L.dsize, 1, nb, winv, d->info, i, s, newton);
  };
  // T_i = lower[i-1] g_{i-1}: in the G_{i-1,i} output slot (same element
  // count, leading dimension b_{i-1}) when off-diagonals are saved
  const bool keep_t = d->save_off_diag != 0 && N > 1;
  auto t_dst = [&](int i, cplx*& dst, long long& sdst) {
    dst = keep_t ? Gu + L.ooff[i - 1] : P;
    sdst = keep_t ? L.osize : b2;
  };
  {
    const int b0 = L.b[0];
    cplx* dst;
    long long sdst;
    m_dst(0, dst, sdst);
    SINV_HIP_TRY(launch_axpby_shift(b0, b0, dscale, D + L.doff[0], b0, 0, d->diag_stride, dst, b0,
                                    0, sdst, shift, 1, nb, s));
    SINV_HIP_TRY(invert_m(0));
  }
  for (int i = 1; i < N; ++i) {
    const int bi = L.b[i], bp = L.b[i - 1];
    cplx* T;
    long long sT;
    t_dst(i, T, sT);
    // T = lower[i-1] g_{i-1}              (b_i x b_{i-1})
    SINV_HIP_TRY(gemm(bi, bp, bp, Lo + L.ooff[i - 1], bp, d->lower_stride, Gd + L.doff[i - 1], bp,
                      L.dsize, T, bp, sT, one, nullptr, 0, 0, zero, nullptr, prec));
    // M = dscale D_ii + shift I - oscale^2 T upper[i-1]
    cplx* dst;
    long long sdst;
    m_dst(i, dst, sdst);
    SINV_HIP_TRY(gemm(bi, bi, bp, T, bp, sT, U + L.ooff[i - 1], bi, d->upper_stride, dst, bi, sdst,
                      mosq, D + L.doff[i], bi, d->diag_stride, dscale, shift, prec));
    SINV_HIP_TRY(invert_m(i));
    if (refine) {
      // one FP64 refinement step of g~_i against M_i formed in FP64 (in Q,
      // free once the inverse has consumed it)
      // (T in FP64 replaces the complex64 one, so the backward sweep reads FP64)
      cplx* gi = Gd + L.doff[i];
      SINV_HIP_TRY(gemm(bi, bp, bp, Lo + L.ooff[i - 1], bp, d->lower_stride, Gd + L.doff[i - 1],
                        bp, L.dsize, T, bp, sT, one, nullptr, 0, 0, zero, nullptr, Prec::F64));
      SINV_HIP_TRY(gemm(bi, bi, bp, T, bp, sT, U + L.ooff[i - 1], bi, d->upper_stride, Q, bi, b2,
                        mosq, D + L.doff[i], bi, d->diag_stride, dscale, shift, Prec::F64));
      // R = M_i g~_i - I                     (in Xw)
      SINV_HIP_TRY(gemm(bi, bi, bi, Q, bi, b2, gi, bi, L.dsize, Xw, bi, b2, one, nullptr, 0, 0,
                        zero, mones, Prec::F64));
      // g_i = g~_i - g~_i R                  (in Q, then back to the g_i slot)
      SINV_HIP_TRY(gemm(bi, bi, bi, gi, bi, L.dsize, Xw, bi, b2, Q, bi, b2, mone, gi, bi, L.dsize,
                        one, nullptr, Prec::F64));
      SINV_HIP_TRY(launch_axpby_shift(bi, bi, one, Q, bi, 0, b2, gi, bi, 0, L.dsize, nullptr, 1,
                                      nb, s));
    }
  }

  // ---- backward sweep (rgf.py:51-62)
  const bool save = d->save_off_diag != 0;
  for (int i = N - 2; i >= 0; --i) {
    const int bi = L.b[i], bj = L.b[i + 1];
    cplx* Xn = save ? Gl + L.ooff[i] : Xw;  // holds G_{i+1,i} = -X_i (b_{i+1} x b_i)
    const long long sX = save ? L.osize : b2;
    if (keep_t) {
      // G_{i+1,i} = -oscale G_{i+1,i+1} T_{i+1}   (T_{i+1} in the G_{i,i+1} slot)
      const cplx mos{-oscale.re, -oscale.im};
      SINV_HIP_TRY(gemm(bj, bi, bj, Gd + L.doff[i + 1], bj, L.dsize, Gu + L.ooff[i], bi, L.osize,
                        Xn, bi, sX, mos, nullptr, 0, 0, zero, nullptr, bprec));
    } else {
      // P = oscale G_{i+1,i+1} lower[i]     (b_{i+1} x b_i)
      SINV_HIP_TRY(gemm(bj, bi, bj, Gd + L.doff[i + 1], bj, L.dsize, Lo + L.ooff[i], bi,
                        d->lower_stride, P, bi, b2, oscale, nullptr, 0, 0, zero, nullptr, bprec));
      // G_{i+1,i} = -P g_i
      SINV_HIP_TRY(gemm(bj, bi, bi, P, bi, b2, Gd + L.doff[i], bi, L.dsize, Xn, bi, sX, mone,
                        nullptr, 0, 0, zero, nullptr, bprec));
    }
    // Q = oscale g_i upper[i]              (b_i x b_{i+1})
    SINV_HIP_TRY(gemm(bi, bj, bi, Gd + L.doff[i], bi, L.dsize, U + L.ooff[i], bj,
                      d->upper_stride, Q, bj, b2, oscale, nullptr, 0, 0, zero, nullptr, bprec));
    if (save) {
      if (d->sym_mat) {
        SINV_HIP_TRY(launch_transpose(bi, bj, OP_T, Xn, bi, 0, sX, Gu + L.ooff[i], bj, 0,
                                      L.osize, 1, nb, s));
      } else {
        // G_{i,i+1} = -Q G_{i+1,i+1}
        SINV_HIP_TRY(gemm(bi, bj, bj, Q, bj, b2, Gd + L.doff[i + 1], bj, L.dsize,
                          Gu + L.ooff[i], bj, L.osize, mone, nullptr, 0, 0, zero, nullptr,
                          bprec));
      }
    }
    // G_ii = g_i + Q X_i = g_i - Q G_{i+1,i}   (in place over g_i); in int8x
    // mode Q's digits are those of the G_{i,i+1} product just before
    const bool q_digits_ready = save && !d->sym_mat;
    SINV_HIP_TRY(gemm(bi, bi, bj, Q, bj, b2, Xn, bi, sX, Gd + L.doff[i], bi, L.dsize, mone,
                      Gd + L.doff[i], bi, L.dsize, one, nullptr, bprec, q_digits_ready));
  }
  if (!save && N > 1) {
    // rgf.py:41,55: unsaved off-diagonal blocks of G stay NaN
    const cplx nan{__builtin_nan(""), __builtin_nan("")};
    if (Gu) SINV_HIP_TRY(launch_fill(Gu, L.osize * nb, nan, s));
    if (Gl) SINV_HIP_TRY(launch_fill(Gl, L.osize * nb, nan, s));
  }
  return SINV_OK;
}

// ---------------------------------------------------------------------------
// Allocation-free RGF: the workspace is the file-scope device array below,
// one copy per GPU (loaded with the code object, held until the process
// ends).  Host code never passes the array's name to a kernel or a copy: its
// device address is looked up once per device with hipGetSymbolAddress (with
// that device current) and only that pointer is used.  Every call on a GPU
// shares that GPU's array; nothing in the RGF path reads workspace it has not
// written in the same call, so no clearing is needed, and calls that overlap
// on one GPU race.
static __device__ __align__(256) unsigned char scratch[SINV_NOALLOC_SCRATCH_BYTES];

namespace {

constexpr int kMaxScratchDevices = 64;
void* g_scratch_ptr[kMaxScratchDevices] = {};
std::mutex g_scratch_mu;

hipError_t scratch_ptr(void** out) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  if (dev < 0 || dev >= kMaxScratchDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_ptr[dev] == nullptr) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return hipSuccess;
}

}  // namespace

extern "C" int sinv_rgf_batched_noalloc(const sinv_rgf_desc* d, void* stream) {
  if (!desc_valid(d)) {
    set_error("sinv_rgf_batched_noalloc: invalid descriptor");
    return SINV_E_INVAL;
  }
  const size_t need = sinv_rgf_workspace_bytes(d);
  if (need == 0) {
    set_error("sinv_rgf_batched_noalloc: invalid block sizes");
    return SINV_E_INVAL;
  }
  if (need > (size_t)SINV_NOALLOC_SCRATCH_BYTES) {
    set_error("sinv_rgf_batched_noalloc: this call needs %zu bytes of workspace, more than the "
              "%zu the library holds per GPU; split the batch or use sinv_rgf_batched",
              need, (size_t)SINV_NOALLOC_SCRATCH_BYTES);
    return SINV_E_INVAL;
  }
  void* p = nullptr;
  SINV_HIP_TRY(scratch_ptr(&p));
  return sinv_rgf_batched(d, p, (size_t)SINV_NOALLOC_SCRATCH_BYTES, stream);
}
