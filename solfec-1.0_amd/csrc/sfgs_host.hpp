// sfgs_host.hpp -- host-side helpers shared by the HIP translation units:
// error macros, device buffer management and the argument blocks of the sweep
// and merit kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <string>
#include <algorithm>
#include "sfgs_ctx.hpp"
#include "sfgs_sweep.hpp"

#define HIPCHK(call) do { hipError_t e_ = (call); if (e_ != hipSuccess) { \
    ctx->err = std::string(#call) + ": " + hipGetErrorString(e_); return SFGS_E_HIP; } } while (0)
#define CHK(call) do { int r_ = (call); if (r_) return r_; } while (0)

namespace sfgs {

template <typename T>
static int ensure(sfgs_ctx *ctx, DBuf<T> &b, size_t n)
{
  if (n == 0) n = 1;
  if (b.cap < n)
  {
    if (b.p) { (void)hipStreamSynchronize(ctx->stream); (void)hipFree(b.p); }
    b.p = nullptr; b.cap = 0;
    size_t cap = n + n/4 + 16;
    hipError_t e = hipMalloc(&b.p, cap * sizeof(T));
    if (e != hipSuccess)
    {
      b.p = nullptr;
      ctx->err = std::string("hipMalloc: ") + hipGetErrorString(e);
      return SFGS_E_HIP;
    }
    b.cap = cap;
  }
  b.n = n;
  return 0;
}

template <typename T>
static void release(DBuf<T> &b) { if (b.p) (void)hipFree(b.p); b.p = nullptr; b.cap = b.n = 0; }

template <typename T>
static int upload(sfgs_ctx *ctx, DBuf<T> &b, const T *h, size_t n)
{
  CHK(ensure(ctx, b, n));
  if (n && h) HIPCHK(hipMemcpyAsync(b.p, h, n*sizeof(T), hipMemcpyHostToDevice, ctx->stream));
  return 0;
}

static inline int nblk(long long n, int t = TPB) { return (int)((n + t - 1) / t); }

static inline int nblk_ll(long long n) { return (int)((n + TPB - 1) / TPB); }


template <typename T>
static SweepArgs<T> sweep_args(sfgs_ctx *ctx, const SFGS_PARAM *par, const T *W, const T *Wd, const T *B,
                               const T *rho, const T *fri, const T *coh, const T *shift, T *U, T *Um, double diageps)
{
  SweepArgs<T> a;
  a.row0 = ctx->slice_row0.p; a.nrows = ctx->slice_nrows.p; a.rowlen = ctx->rowlen.p; a.off = ctx->slice_off.p;
  a.slen = ctx->slice_len.p;
  a.W = W; a.wcol = ctx->wcol.p; a.Wd = Wd; a.A = ctx->A.p; a.B = B; a.rho = rho;
  a.rec = sizeof(T) == 4 ? reinterpret_cast<const float4 *>(ctx->rec32.p) : nullptr;
  a.fri = fri; a.coh = coh; a.shift = shift; a.sep = ctx->sep.p;
  a.P = nullptr; a.Q = nullptr; a.U = U; a.Um = Um; a.Uin = U; a.n = ctx->nc;
  a.diageps = (T)diageps; a.diagmaxiter = par->diagmaxiter; a.diagsolver = par->diagsolver;
  a.failure = par->failure; a.store_U = 1; a.store_Um = 1; a.partials = ctx->partials.p; a.failkey = ctx->failkey.p; a.failcount = ctx->failcount.p;
  a.ctrl = ctx->ctrl.p;
  return a;
}

[[maybe_unused]] static MeritArgs merit_args(sfgs_ctx *ctx, int update_U, double *conmerit)
{
  MeritArgs m;
  m.row0 = ctx->slice_row0.p; m.nrows = ctx->slice_nrows.p; m.rowlen = ctx->rowlen.p; m.off = ctx->slice_off.p;
  m.W = ctx->W.p; m.wcol = ctx->wcol.p; m.Wd = ctx->Wd.p; m.A = ctx->A.p; m.B = ctx->B.p; m.fri = ctx->pfri.p;
  m.coh = ctx->pcoh.p; m.shift = ctx->shift.p; m.R = ctx->R.p; m.U = ctx->U.p; m.n = ctx->nc; m.update_U = update_U;
  m.mpartials = ctx->mpartials.p; m.conmerit = conmerit;
  return m;
}

// the sweeps of SFGS_Solve as one persistent dataflow launch (sfgs_flow.hip);
// the final R and U are left in X0 and sa.U, hc.final_buf = hc.final_U = 0
int run_flow(sfgs_ctx *ctx, const SweepArgs<double> &sa, double *X0, double *X1, int maxiter, int nomerit,
             double eps, double meritval, double *hist, Ctrl &hc, float *ms);
int run_flow(sfgs_ctx *ctx, const SweepArgs<float> &sa, float *X0, float *X1, int maxiter, int nomerit,
             double eps, double meritval, double *hist, Ctrl &hc, float *ms);

// the sweeps of SFGS_Solve as one persistent launch with a device-wide barrier
// per colour (sfgs_bar.hip); the final state is named by hc.final_buf / final_U
// as after the per-colour launches
int run_bar(sfgs_ctx *ctx, const SweepArgs<double> &sa, double *X0, double *X1, int maxiter, int nomerit,
            double eps, double meritval, int reverse, double *hist, Ctrl &hc, float *ms);
int run_bar(sfgs_ctx *ctx, const SweepArgs<float> &sa, float *X0, float *X1, int maxiter, int nomerit,
            double eps, double meritval, int reverse, double *hist, Ctrl &hc, float *ms);

// FP32 copies of the sweep data (sfgs_kernels.hip)
int build_f32_mirror(sfgs_ctx *ctx);
// re-run the assembly of the current bodies and contacts (after a layout change)
int reassemble(sfgs_ctx *ctx);
// body-space operator of NEWTON_SOLVER with locdyn OFF: z = W x formed as
// H (h M^-1 H' x) through the bodies (x at x + stride p for row p); mode 0:
// y = beta y + alpha (X z + Y x) + alpha delta x (x 3 per row), mode 1:
// U = B + z (U, B in SoA planes)
int bodyspace_product(sfgs_ctx *ctx, const double *x, int stride, int mode, const double *X, const double *Y,
                      double alpha, double beta, double delta, double *y);

} // namespace sfgs
