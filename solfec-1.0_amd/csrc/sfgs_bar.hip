// sfgs_bar.hip -- the Gauss-Seidel sweeps of one system in one persistent
// launch with a device-wide barrier after every colour (schedule 3 of
// SFGS_Solve).
//
// Per-colour launches pay, for every colour of every sweep, the ramp and drain
// of a kernel and the dependent loads of a fresh start; at the sizes of a
// settled packing most of a colour's time is that fixed part.  Here the
// workgroups stay resident: logical workgroup L (XCD-contiguous, as the
// per-colour launches map them) sweeps slices s0 + L*NW + wave, s0 + (L +
// nwg)*NW + wave, ... of each colour, and all workgroups meet at a barrier
// before the next colour.  The same slices go to the same wavefront in every
// sweep, so each row's own U, merit U and failure bookkeeping stay with one
// wavefront; the reactions, which other workgroups read, are written and read
// with device-scope relaxed atomics (coherent across the XCDs' L2s), and a
// wavefront waits for its own stores (s_waitcnt) before its workgroup arrives
// at the barrier, so no cache-wide fence is needed.  Iterates are those of the
// per-colour schedule bit for bit.
//
// The sweep's partial sums (error, merit numerator) are written per workgroup
// before the barrier that ends the last colour; every workgroup then sums all
// of them in the same order and takes the same termination decision (as
// k_finish, bgs.c:1016-1036), so no broadcast is needed.  Workgroup 0 records
// the histories and the control block.  Waits are bounded: a wait that exceeds
// the bound raises a flag that ends every wait, and the solve reports an error.
#include <hip/hip_runtime.h>
#include <cmath>
#include <cfloat>
#include <cstring>
#include <vector>
#include <algorithm>
#include "sfgs_ctx.hpp"
#include "sfgs_math.hpp"
#include "sfgs_sweep.hpp"
#include "sfgs_host.hpp"

using namespace sfgs;

namespace {

constexpr int BAR_NW = 4;   // wavefronts per workgroup

template <typename T>
struct BarArgs {
  SweepArgs<T> a;
  T *X0, *X1;               // R ping-pong: sweep k reads X[k & 1], writes X[(k + 1) & 1]
  const int *cslice;        // first slice of every colour, plus the end
  const int *colptr;        // first row of every colour, plus the end
  int ncol, n, nwg, maxsweeps, maxiter, nomerit, reverse;
  double eps, meritval, energy;
  double *part;             // 2 parities x nwg x 4: errup, errlo, merit numerator
  unsigned *bar;            // arrivals, monotone over the launch
  int *abort_all;
  Ctrl *ctrl;
  double *hist;
  unsigned spin_limit;
};

template <typename V>
__device__ __forceinline__ V cld(const V *p) { return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }
template <typename V>
__device__ __forceinline__ void cst(V *p, V v) { __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }

// this wavefront's earlier stores are complete before what follows
__device__ __forceinline__ void publish_fence()
{
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
  __builtin_amdgcn_s_waitcnt(0);
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
}

__device__ __forceinline__ int xcd_logical(int b, int nb)
{
  const int x = b & 7, k = b >> 3, q = nb >> 3, rm = nb & 7;
  return x*q + min(x, rm) + k;
}

// every workgroup arrives (after its wavefronts' stores are complete) and waits
// for all nwg of them; false once any wait of the launch timed out
__device__ __forceinline__ bool grid_wait(unsigned *bar, unsigned target, int *abort_all, unsigned spin_limit,
                                          int *flag)
{
  publish_fence();
  __syncthreads();
  if (threadIdx.x == 0)
  {
    int ab = 0;
    __hip_atomic_fetch_add(bar, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    unsigned spins = 0;
    while (cld(bar) < target)
    {
      if (++spins > spin_limit) { atomicExch(abort_all, 1); ab = 1; break; }
      if ((spins & 63u) == 0 && cld(abort_all)) { ab = 1; break; }
      __builtin_amdgcn_s_sleep(1);
    }
    *flag = ab;
  }
  __syncthreads();
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
  return *flag == 0;
}

template <typename T, int G>
__global__ __launch_bounds__(BAR_NW*64) void k_bar(BarArgs<T> b)
{
  __shared__ double red[3][BAR_NW];
  __shared__ int flag, sweepno, sdone;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int L = xcd_logical(blockIdx.x, b.nwg);
  const bool rec = blockIdx.x == 0;   // the recording workgroup
  unsigned target = 0;
  int done = 0, iters = 0, final_buf = 0, final_U = 0;
  double prev_error = 0.0;
  const double lo_energy = b.energy < DBL_EPSILON ? 1.0 : b.energy;
  bool alive = true;
  for (int k = 0; alive && k < b.maxsweeps; k ++)
  {
    if (threadIdx.x == 0) sweepno = k;
    __syncthreads();
    SweepArgs<T> a = b.a;
    a.P = (k & 1) ? b.X1 : b.X0;
    a.Q = (k & 1) ? b.X0 : b.X1;
    const int back = b.reverse && (k & 1);
    const bool mer = !b.nomerit && k > 0;
    double eu = 0.0, el = 0.0, up = 0.0;
    for (int ci = 0; alive && ci < b.ncol; ci ++)
    {
      const int c = back ? b.ncol - 1 - ci : ci;
      const int s0 = b.cslice[c], s1 = b.cslice[c + 1];
      const int lo = back ? b.colptr[c + 1] : 0, hi = back ? b.n : b.colptr[c];
      for (int s = s0 + L*BAR_NW + wave; s < s1; s += b.nwg*BAR_NW)
      {
        if (mer)
          sweep_slice<T, G, true, true>(a, s, true, lo, hi, back, &sweepno, a.failkey, a.failcount, &b.ctrl->aborted,
                                        eu, el, up);
        else
          sweep_slice<T, G, false, true>(a, s, true, lo, hi, back, &sweepno, a.failkey, a.failcount,
                                         &b.ctrl->aborted, eu, el, up);
      }
      if (ci == b.ncol - 1)
      {
        // this workgroup's share of the sweep, before the barrier that ends it
        eu = wave_sum(eu); el = wave_sum(el); up = wave_sum(up);
        if (lane == 0) { red[0][wave] = eu; red[1][wave] = el; red[2][wave] = up; }
        __syncthreads();
        if (threadIdx.x == 0)
        {
          double t0 = red[0][0], t1 = red[1][0], t2 = red[2][0];
          for (int w = 1; w < BAR_NW; w ++) { t0 += red[0][w]; t1 += red[1][w]; t2 += red[2][w]; }
          double *pt = b.part + 4*((size_t)b.nwg*(k & 1) + L);
          cst(pt, t0); cst(pt + 1, t1); cst(pt + 2, t2);
        }
      }
      target += (unsigned)b.nwg;
      alive = grid_wait(b.bar, target, b.abort_all, b.spin_limit, &flag);
    }
    if (!alive) break;
    // the termination test of sweep k, the same in every workgroup
    if (wave == 0)
    {
      const double *pt = b.part + 4*(size_t)b.nwg*(k & 1);
      double s0 = 0.0, s1 = 0.0, s2 = 0.0;
      for (int w = lane; w < b.nwg; w += 64)
      {
        s0 += cld(pt + 4*(size_t)w); s1 += cld(pt + 4*(size_t)w + 1); s2 += cld(pt + 4*(size_t)w + 2);
      }
      s0 = wave_sum(s0); s1 = wave_sum(s1); s2 = wave_sum(s2);
      if (lane == 0)
      {
        Ctrl *ctl = b.ctrl;
        const int aborted = cld(&ctl->aborted);
        const double error_k = sqrt(s0) / sqrt(s1 == 0.0 ? 1.0 : s1);
        if (b.nomerit)
        {
          iters = k + 1; final_buf = (k + 1) & 1; final_U = 0;
          if (rec) { b.hist[k] = error_k; b.hist[b.maxiter + k] = 0.0; ctl->error = error_k; ctl->merit = 0.0; }
          if (aborted || !(k + 1 < b.maxiter && error_k > b.eps)) done = 1;
        }
        else if (aborted) { iters = k + 1; final_buf = (k + 1) & 1; final_U = 0; done = 1; }
        else if (k == 0) prev_error = error_k;
        else
        {
          const double merit = (0.5*s2) / lo_energy, error = prev_error;
          iters = k;
          if (rec) { b.hist[k - 1] = error; b.hist[b.maxiter + k - 1] = merit; ctl->error = error; ctl->merit = merit; }
          if (!(k < b.maxiter && (error > b.eps || merit > b.meritval))) { final_buf = k & 1; final_U = 1; done = 1; }
          else prev_error = error_k;
        }
        if (rec) { ctl->sweeps = k + 1; ctl->iters = iters; }
        sdone = done;
      }
    }
    __syncthreads();
    if (sdone) break;
  }
  if (rec && threadIdx.x == 0)
  {
    if (!done) { final_buf = b.maxsweeps & 1; final_U = 0; }
    b.ctrl->final_buf = final_buf;
    b.ctrl->final_U = final_U;
    b.ctrl->done = done;
  }
}

template <typename T>
static int bar_cap(sfgs_ctx *ctx, int G, int *cap)
{
  int nb = 0, cus = 0;
  const void *kern = G == 2 ? (const void*)k_bar<T, 2> : (const void*)k_bar<T, 1>;
  HIPCHK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, kern, BAR_NW*64, 0));
  HIPCHK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, ctx->device));
  *cap = (nb * cus * 3) / 4;   // every workgroup must be resident at once
  return 0;
}

template <typename T>
static int run_bar_t(sfgs_ctx *ctx, const SweepArgs<T> &sa, T *X0, T *X1, int maxiter, int nomerit, double eps,
                     double meritval, int reverse, double *hist, Ctrl &hc, float *ms)
{
  hipStream_t st = ctx->stream;
  const int n = ctx->nc, ncol = ctx->ncolours, G = ctx->G_built;
  std::memset(&hc, 0, sizeof(hc));
  hc.eps = eps; hc.meritval = meritval; hc.maxiter = maxiter; hc.nomerit = nomerit; hc.energy = ctx->energy;
  if (G != 1 && G != 2) { ctx->err = "the barrier schedule supports 1 or 2 lanes per row"; return SFGS_E_STATE; }
  if (n == 0 || ctx->nslices == 0) { hc.done = 1; return 0; }
  int cap = 0, widest = 0;
  CHK(bar_cap<T>(ctx, G, &cap));
  for (int c = 0; c < ncol; c ++) widest = std::max(widest, ctx->h_cslice[c + 1] - ctx->h_cslice[c]);
  const int nwg = std::max(1, std::min(cap, (widest + BAR_NW - 1) / BAR_NW));
  const int maxsweeps = nomerit ? maxiter : maxiter + 1;
  if ((double)maxsweeps * ncol * nwg >= 4.0e9) { ctx->err = "too many barriers for one launch"; return SFGS_E_ARG; }
  // scratch: colour tables, partials (2 parities), counter and flag
  const size_t off_tab = 0, off_part = 64*((sizeof(int)*2*(size_t)(ncol + 1) + 63)/64);
  const size_t off_cnt = off_part + sizeof(double)*8*(size_t)nwg;
  const size_t bytes = off_cnt + 256;
  CHK(ensure(ctx, ctx->bar_mem, bytes));
  unsigned char *base = ctx->bar_mem.p;
  std::vector<int> tab(2*(size_t)(ncol + 1));
  for (int c = 0; c <= ncol; c ++) { tab[c] = ctx->h_cslice[c]; tab[(size_t)ncol + 1 + c] = ctx->h_colptr[c]; }
  HIPCHK(hipMemcpyAsync(base + off_tab, tab.data(), sizeof(int)*tab.size(), hipMemcpyHostToDevice, st));
  HIPCHK(hipMemsetAsync(base + off_cnt, 0, 256, st));
  HIPCHK(hipMemcpyAsync(ctx->ctrl.p, &hc, sizeof(Ctrl), hipMemcpyHostToDevice, st));
  BarArgs<T> b;
  b.a = sa;
  b.X0 = X0; b.X1 = X1;
  b.cslice = (const int*)(base + off_tab);
  b.colptr = (const int*)(base + off_tab) + ncol + 1;
  b.ncol = ncol; b.n = n; b.nwg = nwg; b.maxsweeps = maxsweeps; b.maxiter = maxiter; b.nomerit = nomerit;
  b.reverse = reverse;
  b.eps = eps; b.meritval = meritval; b.energy = ctx->energy;
  b.part = (double*)(base + off_part);
  b.bar = (unsigned*)(base + off_cnt);
  b.abort_all = (int*)(base + off_cnt + 64);
  b.ctrl = ctx->ctrl.p;
  b.hist = hist;
  b.spin_limit = 1u << 22;
  HIPCHK(hipEventRecord(ctx->ev2, st));
  auto k1 = k_bar<T, 1>;
  auto k2 = k_bar<T, 2>;
  if (G == 2) hipLaunchKernelGGL(k2, dim3(nwg), dim3(BAR_NW*64), 0, st, b);
  else hipLaunchKernelGGL(k1, dim3(nwg), dim3(BAR_NW*64), 0, st, b);
  HIPCHK(hipEventRecord(ctx->ev3, st));
  int timed_out = 0;
  HIPCHK(hipMemcpyAsync(&hc, ctx->ctrl.p, sizeof(Ctrl), hipMemcpyDeviceToHost, st));
  HIPCHK(hipMemcpyAsync(&timed_out, b.abort_all, sizeof(int), hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  HIPCHK(hipGetLastError());
  if (ms) HIPCHK(hipEventElapsedTime(ms, ctx->ev2, ctx->ev3));
  if (timed_out) { ctx->err = "a barrier of the barrier schedule timed out"; return SFGS_E_STATE; }
  return 0;
}

} // namespace

void sfgs_bar_release(sfgs_ctx *ctx)
{
  release(ctx->bar_mem);
}

int sfgs::run_bar(sfgs_ctx *ctx, const SweepArgs<double> &sa, double *X0, double *X1, int maxiter, int nomerit,
                  double eps, double meritval, int reverse, double *hist, Ctrl &hc, float *ms)
{
  return run_bar_t<double>(ctx, sa, X0, X1, maxiter, nomerit, eps, meritval, reverse, hist, hc, ms);
}

int sfgs::run_bar(sfgs_ctx *ctx, const SweepArgs<float> &sa, float *X0, float *X1, int maxiter, int nomerit,
                  double eps, double meritval, int reverse, double *hist, Ctrl &hc, float *ms)
{
  return run_bar_t<float>(ctx, sa, X0, X1, maxiter, nomerit, eps, meritval, reverse, hist, hc, ms);
}
