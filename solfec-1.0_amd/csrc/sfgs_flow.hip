// sfgs_flow.hip -- the Gauss-Seidel sweeps of one system as a dataflow over
// rows, in one persistent launch (the "flow" schedule of SFGS_Solve).
//
// The per-colour schedule ends every colour at a kernel boundary, and at the
// sizes of a settled packing each colour launch is a chain of dependent
// latencies (row data, gather, 3x3 Newton) with under one wavefront per SIMD.
// Here every wavefront owns a fixed set of slices (slice s -> wave s mod
// nwaves) and runs them in colour-major order, sweep after sweep.  A row waits
// only for its own neighbours: with ver[j] the number of sweeps row j has
// completed, row i starts sweep k once every lower-colour neighbour j (j < i in
// the colour-major order) has ver[j] >= k + 1 and every higher-colour
// neighbour has ver[j] >= k.  Each row then reads exactly the values the
// serial colour-major sweep reads, so the iterates are those of the
// per-colour schedule bit for bit; the error and merit sums are formed per
// wavefront and summed in a fixed order, so they are reproducible but may
// differ from the per-colour sums in the last bits.  Every dependency points to
// an earlier (sweep, row) pair and all wavefronts are resident, so the earliest
// unfinished row can always run.
//
// R after sweep k lives in buffer (k + 1) mod 3; the local-solver U and the
// merit U of sweep k in buffers k mod 2.  The termination test of sweep s
// (as k_finish: iteration s - 1 with the merit on, s without) is taken by the
// wavefront that completes sweep s last; a wavefront starts sweep m only when
// the test of sweep m - 2 is known, so the state a stopping test names is
// never overwritten.  Waits are bounded; a wait that exceeds the bound raises
// a flag that ends every wait, and the solve reports an error.
//
// Publication without cache-wide fences: everything one wavefront hands to
// another (reactions, versions, partial sums, the control block) is written
// and read with device-scope relaxed atomics, which the per-XCD L2s keep
// coherent; a writer waits for its own stores to complete (s_waitcnt) before
// it stores the version or counter that publishes them, and a reader issues
// its data loads only after it has seen that version.  A device-scope
// release / acquire fence would instead write back / invalidate the whole L2
// at every slice.
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

template <typename T>
struct FlowArgs {
  SweepArgs<T> a;
  T *Rb[3];             // R after sweep k in Rb[(k + 1) % 3]; the initial R in Rb[0]
  T *Ub[2];             // local-solver U of sweep k in Ub[k % 2]; the initial U in Ub[1]
  T *Umb[2];            // merit U formed in sweep k in Umb[k % 2]
  int *ver;             // sweeps completed, per row
  double *part;         // per sweep mod 3 and wavefront: errup, errlo, merit numerator
  unsigned *cnt;        // wavefronts done with the sweep, per sweep mod 3
  int *decided;         // termination tests taken: those of sweeps < *decided
  int *abort_all;
  int *sweepno;         // sweep of the latest test, for failure keys
  Ctrl *ctrl;
  double *hist;
  int nslices, nwaves, maxsweeps;
  unsigned spin_limit;
};

__device__ __forceinline__ int ld_relaxed(const int *p)
{
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

template <typename V>
__device__ __forceinline__ V cld(const V *p) { return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }
template <typename V>
__device__ __forceinline__ void cst(V *p, V v) { __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }

// this wavefront's earlier device-scope stores are complete before what follows
// (the compiler keeps memory operations on their side of it)
__device__ __forceinline__ void publish_fence()
{
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
  __builtin_amdgcn_s_waitcnt(0);
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
}

// the termination test after sweep k (bgs.c:1016-1036), as k_finish, on the
// sums of all wavefronts' partials; run by one wavefront
template <typename T>
__device__ void flow_decide(const FlowArgs<T> &f, int k)
{
  const int lane = threadIdx.x & 63;
  const double *pt = f.part + 3*(size_t)f.nwaves*(k % 3);
  double eu = 0.0, el = 0.0, up = 0.0;
  for (int w = lane; w < f.nwaves; w += 64)
  {
    eu += __hip_atomic_load(pt + 3*(size_t)w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    el += __hip_atomic_load(pt + 3*(size_t)w + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    up += __hip_atomic_load(pt + 3*(size_t)w + 2, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  eu = wave_sum(eu); el = wave_sum(el); up = wave_sum(up);
  if (lane != 0) return;
  // the fields the tests of other wavefronts change are accessed coherently;
  // nomerit, maxiter, eps, meritval and energy are set before the launch
  Ctrl *c = f.ctrl;
  if (cld(&c->done)) return;
  const double error_k = sqrt(eu) / sqrt(el == 0.0 ? 1.0 : el);
  cst(&c->sweeps, k + 1);
  const int aborted = ld_relaxed(&c->aborted);
  if (c->nomerit || aborted)
  {
    // R after sweep k, the local-solver U of sweep k
    if (c->nomerit)
    {
      f.hist[k] = error_k;
      f.hist[c->maxiter + k] = 0.0;
      cst(&c->error, error_k);
      cst(&c->merit, 0.0);
    }
    cst(&c->iters, k + 1);
    cst(&c->final_buf, (k + 1) % 3);
    cst(&c->final_U, k % 2);
    if (aborted || !(k + 1 < c->maxiter && error_k > c->eps)) cst(&c->done, 1);
    return;
  }
  if (k == 0) { cst(&c->prev_error, error_k); return; }
  const double lo = c->energy < DBL_EPSILON ? 1.0 : c->energy;
  const double merit = (0.5*up) / lo, error = cld(&c->prev_error);
  f.hist[k - 1] = error;
  f.hist[c->maxiter + k - 1] = merit;
  cst(&c->iters, k);
  cst(&c->error, error);
  cst(&c->merit, merit);
  if (!(k < c->maxiter && (error > c->eps || merit > c->meritval)))
  {
    // R after sweep k - 1, the merit U formed in sweep k
    cst(&c->final_buf, k % 3);
    cst(&c->final_U, 2 + k % 2);
    cst(&c->done, 1);
  }
  else cst(&c->prev_error, error_k);
}

template <typename T, int G>
__global__ __launch_bounds__(64) void k_flow(FlowArgs<T> f)
{
  const int W = blockIdx.x, lane = threadIdx.x, r = lane / G, g = lane % G;
  const SweepArgs<T> &a0 = f.a;
  for (int k = 0; k < f.maxsweeps; k ++)
  {
    // gate: the termination test of sweep k - 2 is known and did not stop
    if (k >= 2)
    {
      unsigned spins = 0;
      int ab = 0;
      while (ld_relaxed(f.decided) < k - 1)
      {
        if (++spins > f.spin_limit) { atomicExch(f.abort_all, 1); ab = 1; break; }
        if ((spins & 63u) == 0 && (ld_relaxed(f.abort_all) || ld_relaxed(&f.ctrl->done))) { ab = 1; break; }
        __builtin_amdgcn_s_sleep(1);
      }
      __atomic_signal_fence(__ATOMIC_SEQ_CST);
      if (ab || ld_relaxed(&f.ctrl->done)) return;
    }
    SweepArgs<T> a = a0;
    a.P = f.Rb[k % 3];
    a.Q = f.Rb[(k + 1) % 3];
    a.Uin = f.Ub[(k + 1) % 2];
    a.U = f.Ub[k % 2];
    a.Um = f.Umb[k % 2];
    const bool merit = !f.ctrl->nomerit && k > 0;
    double eu = 0.0, el = 0.0, up = 0.0;
    for (int s = W; s < f.nslices; s += f.nwaves)
    {
      // wait for this lane's neighbours: all their versions loaded at once, again
      // after a pause while any is behind
      const int row0 = a.row0[s];
      const bool act = r < a.nrows[s];
      int ok = 1;
      if (act)
      {
        const int p = row0 + r, len = a.rowlen[p];
        const int mine = len > g ? (len - g + G - 1) / G : 0;
        const int *cb = a.wcol + a.off[s] + lane;
        unsigned spins = 0;
        for (int m0 = 0; m0 < mine && ok; )
        {
          constexpr int CH = 8;
          int jj[CH], vv[CH];
          bool ready = true;
#pragma unroll
          for (int q = 0; q < CH; q ++) jj[q] = cb[(size_t)min(m0 + q, mine - 1)*SLICE];
#pragma unroll
          for (int q = 0; q < CH; q ++) vv[q] = ld_relaxed(f.ver + jj[q]);   // all in flight together
#pragma unroll
          for (int q = 0; q < CH; q ++) ready = ready && (m0 + q >= mine || vv[q] >= k + (jj[q] < p ? 1 : 0));
          if (ready) { m0 += CH; continue; }
          if (++spins > f.spin_limit) { atomicExch(f.abort_all, 1); ok = 0; break; }
          // a stopped solve leaves rows of later sweeps undone: stop waiting for them
          if ((spins & 63u) == 0 && (ld_relaxed(f.abort_all) || ld_relaxed(&f.ctrl->done))) { ok = 0; break; }
          __builtin_amdgcn_s_sleep(1);
        }
      }
      if (__any(!ok)) return;
      __atomic_signal_fence(__ATOMIC_SEQ_CST);
      // lower-colour neighbours (j < row0 <= p) are read from Q, the rest from P
      if (merit)
        sweep_slice<T, G, true, true>(a, s, true, 0, row0, 0, f.sweepno, a.failkey, a.failcount, &f.ctrl->aborted,
                                      eu, el, up);
      else
        sweep_slice<T, G, false, true>(a, s, true, 0, row0, 0, f.sweepno, a.failkey, a.failcount, &f.ctrl->aborted,
                                       eu, el, up);
      publish_fence();
      if (act && g == 0) cst(f.ver + row0 + r, k + 1);
    }
    // this wavefront's share of sweep k
    eu = wave_sum(eu); el = wave_sum(el); up = wave_sum(up);
    if (lane == 0)
    {
      double *pt = f.part + 3*((size_t)f.nwaves*(k % 3) + W);
      cst(pt, eu); cst(pt + 1, el); cst(pt + 2, up);
    }
    publish_fence();
    unsigned last = 0;
    if (lane == 0) last = atomicAdd(f.cnt + k % 3, 1u) == (unsigned)f.nwaves - 1 ? 1u : 0u;
    last = __shfl(last, 0, 64);
    if (last)
    {
      // tests are taken in sweep order: the one of sweep k - 1 may still be running
      unsigned spins = 0;
      while (ld_relaxed(f.decided) < k)
      {
        if (++spins > f.spin_limit) { atomicExch(f.abort_all, 1); return; }
        if ((spins & 63u) == 0 && ld_relaxed(f.abort_all)) return;
        __builtin_amdgcn_s_sleep(1);
      }
      __atomic_signal_fence(__ATOMIC_SEQ_CST);
      flow_decide(f, k);
      if (lane == 0)
      {
        cst(f.sweepno, k + 1);
        cst(f.cnt + k % 3, 0u);
      }
      publish_fence();
      if (lane == 0) cst(f.decided, k + 1);
    }
  }
}

template <typename T>
__global__ void k_flow_init(int n, const T *R, const T *U, T *R0, T *U1, int *ver)
{
  int p = blockIdx.x*blockDim.x + threadIdx.x;
  if (p >= n) return;
#pragma unroll
  for (int k = 0; k < 4; k ++) R0[4*(size_t)p + k] = R[4*(size_t)p + k];
#pragma unroll
  for (int k = 0; k < 3; k ++) U1[(size_t)k*n + p] = U[(size_t)k*n + p];
  ver[p] = 0;
}

// final R and U of the flow into the canonical buffers
template <typename T>
__global__ void k_flow_final(int n, const T *Rf, const T *Uf, T *R, T *U)
{
  int p = blockIdx.x*blockDim.x + threadIdx.x;
  if (p >= n) return;
#pragma unroll
  for (int k = 0; k < 4; k ++) R[4*(size_t)p + k] = Rf[4*(size_t)p + k];
#pragma unroll
  for (int k = 0; k < 3; k ++) U[(size_t)k*n + p] = Uf[(size_t)k*n + p];
}

template <typename T>
static int flow_cap(sfgs_ctx *ctx, int G, int *cap)
{
  int nb = 0, cus = 0;
  const void *kern = G == 2 ? (const void*)k_flow<T, 2> : (const void*)k_flow<T, 1>;
  HIPCHK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, kern, 64, 0));
  HIPCHK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, ctx->device));
  *cap = (nb * cus * 3) / 4;   // every wavefront must be resident at once
  return 0;
}

template <typename T>
static int run_flow_t(sfgs_ctx *ctx, const SweepArgs<T> &sa, T *X0, T *X1, int maxiter, int nomerit, double eps,
                      double meritval, double *hist, Ctrl &hc, float *ms)
{
  hipStream_t st = ctx->stream;
  const int n = ctx->nc, ns = ctx->nslices, G = ctx->G_built;
  std::memset(&hc, 0, sizeof(hc));
  if (G != 1 && G != 2) { ctx->err = "the flow schedule supports 1 or 2 lanes per row"; return SFGS_E_STATE; }
  if (n == 0 || ns == 0) { hc.done = 1; return 0; }
  int cap = 0;
  CHK(flow_cap<T>(ctx, G, &cap));
  const int nwaves = std::max(1, std::min(cap, ns));
  // device scratch: third R buffer, U and merit-U pairs, row versions, partials, counters
  const size_t bR = sizeof(T)*4*(size_t)n, bU = sizeof(T)*3*(size_t)n;
  const size_t off_ver = bR + 4*bU, off_part = off_ver + sizeof(int)*(size_t)n + 64;
  const size_t off_cnt = off_part + sizeof(double)*9*(size_t)nwaves;
  const size_t bytes = off_cnt + 256;
  CHK(ensure(ctx, ctx->flow_mem, bytes));
  unsigned char *base = ctx->flow_mem.p;
  FlowArgs<T> f;
  f.a = sa;
  f.Rb[0] = (T*)base; f.Rb[1] = X0; f.Rb[2] = X1;
  f.Ub[0] = (T*)(base + bR); f.Ub[1] = (T*)(base + bR + bU);
  f.Umb[0] = (T*)(base + bR + 2*bU); f.Umb[1] = (T*)(base + bR + 3*bU);
  f.ver = (int*)(base + off_ver);
  f.part = (double*)(base + (off_part & ~(size_t)7));
  f.cnt = (unsigned*)(base + off_cnt);
  f.decided = (int*)(base + off_cnt + 16);
  f.abort_all = (int*)(base + off_cnt + 20);
  f.sweepno = (int*)(base + off_cnt + 24);
  f.ctrl = ctx->ctrl.p;
  f.hist = hist;
  f.nslices = ns;
  f.nwaves = nwaves;
  f.maxsweeps = nomerit ? maxiter : maxiter + 1;
  f.spin_limit = 1u << 22;
  hc.eps = eps; hc.meritval = meritval; hc.maxiter = maxiter; hc.nomerit = nomerit; hc.energy = ctx->energy;
  HIPCHK(hipMemcpyAsync(ctx->ctrl.p, &hc, sizeof(Ctrl), hipMemcpyHostToDevice, st));
  HIPCHK(hipMemsetAsync(base + off_cnt, 0, 256, st));
  // initial state: R into Rb[0], U into Ub[1], versions 0
  hipLaunchKernelGGL(k_flow_init<T>, dim3(nblk(n)), dim3(TPB), 0, st, n, (const T*)X0, (const T*)sa.U, f.Rb[0],
                     f.Ub[1], f.ver);
  HIPCHK(hipEventRecord(ctx->ev2, st));
  auto k1 = k_flow<T, 1>;
  auto k2 = k_flow<T, 2>;
  if (G == 2) hipLaunchKernelGGL(k2, dim3(nwaves), dim3(64), 0, st, f);
  else hipLaunchKernelGGL(k1, dim3(nwaves), dim3(64), 0, st, f);
  HIPCHK(hipEventRecord(ctx->ev3, st));
  int timed_out = 0;
  HIPCHK(hipMemcpyAsync(&hc, ctx->ctrl.p, sizeof(Ctrl), hipMemcpyDeviceToHost, st));
  HIPCHK(hipMemcpyAsync(&timed_out, f.abort_all, sizeof(int), hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  HIPCHK(hipGetLastError());
  if (ms) HIPCHK(hipEventElapsedTime(ms, ctx->ev2, ctx->ev3));
  if (timed_out) { ctx->err = "a wait of the flow schedule timed out"; return SFGS_E_STATE; }
  if (!hc.done)
  {
    // maxsweeps ran without a stopping test (cannot happen: the last test stops)
    const int k = f.maxsweeps - 1;
    hc.final_buf = (k + 1) % 3;
    hc.final_U = k % 2;
  }
  const T *Rf = f.Rb[hc.final_buf];
  const T *Uf = hc.final_U >= 2 ? f.Umb[hc.final_U - 2] : f.Ub[hc.final_U];
  hipLaunchKernelGGL(k_flow_final<T>, dim3(nblk(n)), dim3(TPB), 0, st, n, Rf, Uf, X0, sa.U);
  HIPCHK(hipStreamSynchronize(st));
  HIPCHK(hipGetLastError());
  // the state is now in X0 and sa.U; final_U = 1 tells the caller that U is a
  // merit pass's, which it forms again from R as it does for the colour launches
  hc.final_buf = 0;
  hc.final_U = hc.final_U >= 2 ? 1 : 0;
  return 0;
}

} // namespace

int sfgs::run_flow(sfgs_ctx *ctx, const SweepArgs<double> &sa, double *X0, double *X1, int maxiter, int nomerit,
                   double eps, double meritval, double *hist, Ctrl &hc, float *ms)
{
  return run_flow_t<double>(ctx, sa, X0, X1, maxiter, nomerit, eps, meritval, hist, hc, ms);
}

int sfgs::run_flow(sfgs_ctx *ctx, const SweepArgs<float> &sa, float *X0, float *X1, int maxiter, int nomerit,
                   double eps, double meritval, double *hist, Ctrl &hc, float *ms)
{
  return run_flow_t<float>(ctx, sa, X0, X1, maxiter, nomerit, eps, meritval, hist, hc, ms);
}
