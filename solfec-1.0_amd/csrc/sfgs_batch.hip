// sfgs_batch.hip -- many independent contact systems ("scenes") solved by one
// persistent kernel launch.
//
// The scenes share one device store: rows are ordered by (scene, colour,
// -row length, i), so the rows of a scene form one contiguous range and each
// (scene, colour) group one contiguous set of SELL-64 slices.  Scene s owns
// workgroups [wg0, wg0 + nwg); they sweep the scene's colours in order, the
// waves of the scene's workgroups taking the slices of a colour round robin,
// and meet at a scene barrier after every colour.  A scene of one workgroup
// synchronises with __syncthreads only; larger scenes use a counter in global
// memory with agent-scope fences and a bounded wait (a wait that exceeds the
// bound raises an abort flag that ends every wait, so the grid always drains).
// Every workgroup of a scene forms the same sums in the same order, so all of
// them reach the same termination decision without a broadcast.
//
// Per scene the kernel runs what SFGS_Solve runs for one system in FP64: the
// free energy, Gauss-Seidel sweeps with the fused merit of the previous iterate
// and the termination test of bgs.c:1016-1036 (as k_finish).  MIXED solves run
// the per-colour schedule below, which has the two phases of the single-system
// MIXED solve.  The sweep itself is sweep_slice, shared with the per-colour
// kernels.
#include <hip/hip_runtime.h>
#include <cmath>
#include <cfloat>
#include <cstring>
#include <vector>
#include <algorithm>
#include <type_traits>
#include "sfgs_ctx.hpp"
#include "sfgs_math.hpp"
#include "sfgs_sweep.hpp"
#include "sfgs_host.hpp"

using namespace sfgs;

namespace {

struct BatchScene {
  int wg0, nwg;     // workgroups of the scene
  int row0, row1;   // its rows in the colour-major order
};

struct BatchOut {
  int iters, nref, sweeps, done, aborted, failcount, timeout;
  int final_buf, final_U;   // per-colour schedule: buffer of the final R (0, 1), final U (0 U, 1 Um)
  int iters0;               // per-colour MIXED: iterations of the FP32 phase
  int sweeps0;              // per-colour MIXED: sweeps of the FP32 phase
  unsigned long long failkey;
  double error, merit, energy, prev_error;
};

// scene layout shared by the sweeps of both precisions
struct BatchGeom {
  const int *gslice;       // first slice of every (scene, colour) group, plus the end
  const int *gptr;         // first row of every group, plus the end
  int ncol;
  int *abort_all;          // raised by a wait that exceeded spin_limit
  unsigned spin_limit;
};

template <typename T>
struct BatchArgs {
  SweepArgs<T> a;          // sweep data of the first phase (P, Q set per sweep)
  SweepArgs<double> d;     // FP64 sweep data (energy, refinement)
  const BatchScene *scenes;
  const int *wg_scene;     // scene of every workgroup
  BatchGeom geo;
  T *X0, *X1;              // first-phase R buffers
  double *R64, *R64b;      // FP64 R buffers (R64 is the canonical store)
  T *Ut, *Umt;             // first-phase U and merit-U
  double *U64, *Um64;      // FP64 U and merit-U
  BatchOut *out;
  double *wpart;           // 2 parities x 4 doubles per workgroup
  unsigned *bar;           // barrier counter per scene
  int maxiter, nomerit, reverse;
  double eps, meritval;
};

// all workgroups of a scene meet; false once any wait of the launch timed out
__device__ __forceinline__ bool scene_barrier(unsigned *bar, int nw, int *abort_all, unsigned spin_limit, int *flag)
{
  __syncthreads();
  if (threadIdx.x == 0)
  {
    int ab = 0;
    if (nw > 1)
    {
      __threadfence();
      const unsigned arrived = atomicAdd(bar, 1u);
      const unsigned target = (arrived / (unsigned)nw + 1u) * (unsigned)nw;
      unsigned spins = 0;
      while (__hip_atomic_load(bar, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target)
      {
        if (++spins > spin_limit) { atomicExch(abort_all, 1); ab = 1; break; }
        if ((spins & 255u) == 0 && __hip_atomic_load(abort_all, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
        {
          ab = 1;
          break;
        }
        __builtin_amdgcn_s_sleep(2);
      }
      __threadfence();
    }
    *flag = ab;
  }
  __syncthreads();
  return *flag == 0;
}

// FP64 free energy term B'AB of the rows of one slice (lane g = 0 of each row)
template <int G>
__device__ __forceinline__ double slice_energy(const SweepArgs<double> &d, int s)
{
  const int l = threadIdx.x & 63, r = l / G, g = l % G;
  if (g != 0 || r >= d.nrows[s]) return 0.0;
  const int p = d.row0[s] + r, n = d.n;
  const double B0 = d.B[p], B1 = d.B[(size_t)n + p], B2 = d.B[2*(size_t)n + p];
  double A[9];
#pragma unroll
  for (int e = 0; e < 9; e ++) A[e] = d.A[(size_t)e*n + p];
  return B0*(A[0]*B0 + A[3]*B1 + A[6]*B2) + B1*(A[1]*B0 + A[4]*B1 + A[7]*B2) + B2*(A[2]*B0 + A[5]*B1 + A[8]*B2);
}

// row p of slice s handled by this lane (g = 0), or -1
template <int G>
__device__ __forceinline__ int slice_row(const int *row0, const int *nrows, int s)
{
  const int l = threadIdx.x & 63, r = l / G, g = l % G;
  return (g == 0 && r < nrows[s]) ? row0[s] + r : -1;
}

// one sweep of a scene: its colours in order (reversed when back), a scene
// barrier after each colour
template <typename T, int G, int NW, bool MERIT>
__device__ __forceinline__ bool batch_sweep(const BatchGeom &geo, const SweepArgs<T> &a, const BatchScene &sc, int g0,
                                            int w, int nw, int back, int *sweepno, BatchOut *o, unsigned *bar,
                                            int *flag, double &eu, double &el, double &up)
{
  const int wave = threadIdx.x >> 6;
  for (int ci = 0; ci < geo.ncol; ci ++)
  {
    const int c = back ? geo.ncol - 1 - ci : ci, gi = g0 + c;
    const int s0 = geo.gslice[gi], s1 = geo.gslice[gi + 1];
    if (s0 == s1) continue;   // same decision in every workgroup of the scene
    const int lo = back ? geo.gptr[gi + 1] : sc.row0, hi = back ? sc.row1 : geo.gptr[gi];
    for (int s = s0 + w*NW + wave; s < s1; s += nw*NW)
      sweep_slice<T, G, MERIT>(a, s, true, lo, hi, back, sweepno, &o->failkey, &o->failcount, &o->aborted, eu, el, up);
    if (!scene_barrier(bar, nw, geo.abort_all, geo.spin_limit, flag)) return false;
  }
  return true;
}

__global__ void k_batch_to_f32(long long n, const double *src, float *dst)
{
  long long i = (long long)blockIdx.x*blockDim.x + threadIdx.x;
  if (i < n) dst[i] = (float)src[i];
}

template <typename T, int G, int NW>
__global__ __launch_bounds__(NW*64) void k_batch(BatchArgs<T> b)
{
  __shared__ double red[3][NW];
  __shared__ double tot[3];
  __shared__ int flag, sweepno;
  const int wg = blockIdx.x, wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int sid = b.wg_scene[wg];
  const BatchScene sc = b.scenes[sid];
  const int w = wg - sc.wg0, nw = sc.nwg, g0 = sid*b.geo.ncol;
  const int sl_lo = b.geo.gslice[g0], sl_hi = b.geo.gslice[g0 + b.geo.ncol];
  BatchOut *o = b.out + sid;
  unsigned *bar = b.bar + sid;
  int parity = 0;

  // deterministic scene-wide sums of three values (identical in every workgroup)
  auto scene_sum = [&](double &v0, double &v1, double &v2) -> bool {
    v0 = wave_sum(v0); v1 = wave_sum(v1); v2 = wave_sum(v2);
    if (lane == 0) { red[0][wave] = v0; red[1][wave] = v1; red[2][wave] = v2; }
    __syncthreads();
    if (threadIdx.x == 0)
    {
      double t0 = red[0][0], t1 = red[1][0], t2 = red[2][0];
      for (int k = 1; k < NW; k ++) { t0 += red[0][k]; t1 += red[1][k]; t2 += red[2][k]; }
      if (nw == 1) { tot[0] = t0; tot[1] = t1; tot[2] = t2; }
      else
      {
        double *pp = b.wpart + 4*(2*(size_t)wg + parity);
        pp[0] = t0; pp[1] = t1; pp[2] = t2;
      }
    }
    bool ok = true;
    if (nw > 1)
    {
      ok = scene_barrier(bar, nw, b.geo.abort_all, b.geo.spin_limit, &flag);
      if (threadIdx.x == 0)
      {
        double t0 = 0.0, t1 = 0.0, t2 = 0.0;
        for (int q = 0; q < nw; q ++)
        {
          double *pp = b.wpart + 4*(2*(size_t)(sc.wg0 + q) + parity);
          t0 += __hip_atomic_load(pp, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          t1 += __hip_atomic_load(pp + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          t2 += __hip_atomic_load(pp + 2, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        tot[0] = t0; tot[1] = t1; tot[2] = t2;
      }
      parity ^= 1;
    }
    __syncthreads();
    v0 = tot[0]; v1 = tot[1]; v2 = tot[2];
    return ok;
  };

  // ---- free energy of the scene (ldy.c:729-736)
  double e0 = 0.0, e1 = 0.0, e2 = 0.0;
  for (int s = sl_lo + w*NW + wave; s < sl_hi; s += nw*NW) e0 += slice_energy<G>(b.d, s);
  bool alive = scene_sum(e0, e1, e2);
  const double energy = 0.5*e0, lo_energy = energy < DBL_EPSILON ? 1.0 : energy;

  // ---- Gauss-Seidel sweeps in T with the merit of the previous iterate fused in
  const int maxsweeps = b.nomerit ? b.maxiter : b.maxiter + 1;
  int done = 0, iters = 0, final_buf = 0, final_U = 0, aborted = 0, nsw = 0;
  double error = 0.0, merit = 0.0, prev_error = 0.0;
  for (int k = 0; alive && k < maxsweeps; k ++)
  {
    if (threadIdx.x == 0) sweepno = k;
    __syncthreads();
    SweepArgs<T> a = b.a;
    a.P = (k & 1) ? b.X1 : b.X0;
    a.Q = (k & 1) ? b.X0 : b.X1;
    const int back = b.reverse && (k & 1);
    double eu = 0.0, el = 0.0, up = 0.0;
    if (!b.nomerit && k > 0)
      alive = batch_sweep<T, G, NW, true>(b.geo, a, sc, g0, w, nw, back, &sweepno, o, bar, &flag, eu, el, up);
    else
      alive = batch_sweep<T, G, NW, false>(b.geo, a, sc, g0, w, nw, back, &sweepno, o, bar, &flag, eu, el, up);
    if (!alive) break;
    nsw ++;
    alive = scene_sum(eu, el, up);
    if (!alive) break;
    // termination (bgs.c:1016-1036), as k_finish
    aborted = __hip_atomic_load(&o->aborted, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0;
    const double error_k = sqrt(eu) / sqrt(el == 0.0 ? 1.0 : el);
    if (b.nomerit)
    {
      iters = k + 1; error = error_k; merit = 0.0; final_buf = (k + 1) & 1; final_U = 0;
      if (aborted || !(k + 1 < b.maxiter && error_k > b.eps)) done = 1;
    }
    else if (aborted)
    {
      iters = k + 1; final_buf = (k + 1) & 1; final_U = 0; done = 1;
    }
    else if (k == 0) prev_error = error_k;
    else
    {
      merit = 0.5*up / lo_energy;
      error = prev_error;
      iters = k;
      if (!(k < b.maxiter && (error > b.eps || merit > b.meritval))) { final_buf = k & 1; final_U = 1; done = 1; }
      else prev_error = error_k;
    }
    if (done) break;
  }
  if (!done) { final_buf = maxsweeps & 1; final_U = 0; }

  if (alive)
  {
    // the final state into the canonical buffers (own rows only)
    const T *Rf = final_buf ? b.X1 : b.X0;
    const T *Uf = final_U ? b.Umt : b.Ut;
    const int n = b.a.n;
    for (int s = sl_lo + w*NW + wave; s < sl_hi; s += nw*NW)
    {
      const int p = slice_row<G>(b.d.row0, b.d.nrows, s);
      if (p < 0) continue;
      if (final_buf)
#pragma unroll
        for (int k = 0; k < 4; k ++) b.R64[4*(size_t)p + k] = (double)Rf[4*(size_t)p + k];
      if (final_U)
#pragma unroll
        for (int k = 0; k < 3; k ++) b.U64[(size_t)k*n + p] = (double)Uf[(size_t)k*n + p];
    }
  }
  if (w == 0 && threadIdx.x == 0)
  {
    o->iters = iters;
    o->nref = 0;
    o->sweeps = nsw;
    o->done = done;
    o->error = error;
    o->merit = merit;
    o->energy = energy;
    o->timeout = alive ? 0 : 1;
  }
}

static int dev_cap(sfgs_ctx *ctx, const void *kern, int threads, int *cap)
{
  int nb = 0, cus = 0;
  HIPCHK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, kern, threads, 0));
  HIPCHK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, ctx->device));
  *cap = (nb * cus * 3) / 4;   // margin: every workgroup of a launch must be resident at once
  return 0;
}

template <typename T, int G, int NW>
static int launch_batch(sfgs_ctx *ctx, const BatchArgs<T> &b, int nwg, int *cap_out)
{
  auto kern = k_batch<T, G, NW>;
  if (cap_out) return dev_cap(ctx, (const void*)kern, NW*64, cap_out);
  hipLaunchKernelGGL(kern, dim3(nwg), dim3(NW*64), 0, ctx->stream, b);
  return 0;
}

template <typename T>
static int dispatch_batch(sfgs_ctx *ctx, int G, int NW, const BatchArgs<T> &b, int nwg, int *cap_out)
{
#define SFGS_BATCH_NW(GG) \
  switch (NW) { \
  case 1: return launch_batch<T, GG, 1>(ctx, b, nwg, cap_out); \
  case 2: return launch_batch<T, GG, 2>(ctx, b, nwg, cap_out); \
  case 4: return launch_batch<T, GG, 4>(ctx, b, nwg, cap_out); \
  default: return launch_batch<T, GG, 8>(ctx, b, nwg, cap_out); }
  if (G == 2) { SFGS_BATCH_NW(2) }
  SFGS_BATCH_NW(1)
#undef SFGS_BATCH_NW
}

// ---------------------------------------------------------------------
// The per-colour schedule of the batch (tune.schedule = 1): one launch per
// colour sweeps colour c of every scene that is still running, the scenes'
// slices of that colour listed together, and one small launch per sweep takes
// each scene's termination test.  Kernel boundaries are the barriers, so no
// workgroup waits for another; a launch holds the colour of all scenes, which
// fills the device even when each scene alone would not.  Each wavefront
// handles one slice and writes that slice's partial sums, which a scene's test
// adds up in slice order.
// ---------------------------------------------------------------------

struct UScene {
  int slice0, slice1;     // the scene's slices (all colours)
};

// sbound: per slice, the rows already updated in this sweep (forward: the
// scene's earlier colours; backward: its later ones)
template <typename T, int G, bool MERIT>
__global__ __launch_bounds__(TPB) void k_usweep(SweepArgs<T> a, const int *clist, int nlist, const int2 *sbound,
                                                const int *sscene, BatchOut *out, double *spart, int backward)
{
  const int idx = blockIdx.x*(TPB/SLICE) + (threadIdx.x >> 6);
  if (idx >= nlist) return;                       // whole wavefront
  const int s = clist[idx], sc = sscene[s];
  BatchOut *o = out + sc;
  if (o->done) return;                            // a finished scene: whole wavefront
  const int2 bd = sbound[s];
  double eu = 0.0, el = 0.0, up = 0.0;
  int sweepno = o->iters0 + o->sweeps;
  sweep_slice<T, G, MERIT>(a, s, true, bd.x, bd.y, backward, &sweepno, &o->failkey, &o->failcount, &o->aborted,
                           eu, el, up);
  eu = wave_sum(eu); el = wave_sum(el); up = wave_sum(up);
  if ((threadIdx.x & 63) == 0)
  {
    double *pt = spart + 3*(size_t)s;
    pt[0] = eu; pt[1] = el; pt[2] = up;
  }
}

// per-slice free energy terms
template <int G>
__global__ __launch_bounds__(TPB) void k_uenergy(SweepArgs<double> d, int nslices, double *spart)
{
  const int s = blockIdx.x*(TPB/SLICE) + (threadIdx.x >> 6);
  if (s >= nslices) return;
  double e = wave_sum(slice_energy<G>(d, s));
  if ((threadIdx.x & 63) == 0) spart[3*(size_t)s] = e;
}

// deterministic sum of three partials over a scene's slices (one block per scene)
__device__ __forceinline__ void scene_partials(const double *spart, int s0, int s1, double &a, double &b, double &c)
{
  a = b = c = 0.0;
  for (int s = s0 + (int)threadIdx.x; s < s1; s += TPB)
  {
    a += spart[3*(size_t)s]; b += spart[3*(size_t)s + 1]; c += spart[3*(size_t)s + 2];
  }
  block_sum3(a, b, c);
}

// mode 0: free energy of each scene (before the sweeps)
// mode 1: termination test after a sweep (bgs.c:1016-1036, as k_finish) per running scene;
//         running scenes are counted into active[sweep].  The phase's iteration bound is
//         maxiter - iters0 (iters0: iterations of an earlier phase, reported on top), and
//         no stop on convergence before miniter iterations of the phase.
__global__ __launch_bounds__(TPB) void k_uscene(int mode, const UScene *sc, const double *spart, BatchOut *out,
                                                int maxiter, int miniter, int nomerit, double eps, double meritval,
                                                int sweep, int *active)
{
  BatchOut *o = out + blockIdx.x;
  const UScene u = sc[blockIdx.x];
  if (u.slice1 == u.slice0) { if (threadIdx.x == 0 && mode == 0) o->done = 1; return; }
  if (mode == 1 && o->done) return;
  double eu, el, up;
  scene_partials(spart, u.slice0, u.slice1, eu, el, up);
  if (threadIdx.x != 0) return;
  if (mode == 0) { o->energy = 0.5*eu; return; }
  const double error_k = sqrt(eu) / sqrt(el == 0.0 ? 1.0 : el);
  const int k = o->sweeps, base = o->iters0, bound = maxiter - base;
  o->sweeps = k + 1;
  const int aborted = o->aborted != 0;
  int done = 0;
  if (nomerit)
  {
    o->iters = base + k + 1; o->error = error_k; o->merit = 0.0; o->final_buf = (k + 1) & 1; o->final_U = 0;
    done = aborted || !(k + 1 < bound && (error_k > eps || k + 1 < miniter));
  }
  else if (aborted) { o->iters = base + k + 1; o->final_buf = (k + 1) & 1; o->final_U = 0; done = 1; }
  else if (k == 0) o->prev_error = error_k;
  else
  {
    const double lo = o->energy < DBL_EPSILON ? 1.0 : o->energy;
    const double merit = 0.5*up/lo, error = o->prev_error;
    o->iters = base + k; o->error = error; o->merit = merit;
    if (!(k < bound && (error > eps || merit > meritval || k < miniter))) { o->final_buf = k & 1; o->final_U = 1; done = 1; }
    else o->prev_error = error_k;
  }
  if (done) o->done = 1;
  else atomicAdd(active + sweep, 1);
}

// start of the FP64 phase of MIXED: a scene that did not abort and has
// iterations left runs again from the state k_ufinal wrote into R64 / U64
__global__ void k_uphase(int nscenes, const UScene *sc, int maxiter, BatchOut *out)
{
  const int q = blockIdx.x*blockDim.x + threadIdx.x;
  if (q >= nscenes) return;
  BatchOut *o = out + q;
  o->iters0 = o->iters;
  o->sweeps0 = o->sweeps;
  o->sweeps = 0;
  o->prev_error = 0.0;
  o->final_buf = 0;
  o->final_U = 0;
  o->done = (sc[q].slice1 == sc[q].slice0 || o->aborted || o->iters >= maxiter) ? 1 : 0;
}

// the FP64 phase's final state into the canonical buffers, per scene's final buffers
template <int G>
__global__ __launch_bounds__(TPB) void k_ufinal64(int nslices, const int *row0, const int *nrows, const int *sscene,
                                                  const BatchOut *out, const double *Rb, const double *Um, double *R,
                                                  double *U, int n)
{
  const int s = blockIdx.x*(TPB/SLICE) + (threadIdx.x >> 6);
  if (s >= nslices) return;
  const int p = slice_row<G>(row0, nrows, s);
  if (p < 0) return;
  const BatchOut &o = out[sscene[s]];
  if (o.final_buf)
#pragma unroll
    for (int k = 0; k < 4; k ++) R[4*(size_t)p + k] = Rb[4*(size_t)p + k];
  if (o.final_U)
#pragma unroll
    for (int k = 0; k < 3; k ++) U[(size_t)k*n + p] = Um[(size_t)k*n + p];
}

// final state of the first phase into the FP64 store, per scene's final buffers
template <typename T, int G>
__global__ __launch_bounds__(TPB) void k_ufinal(int nslices, const int *row0, const int *nrows, const int *sscene,
                                                const BatchOut *out, const T *X0, const T *X1, const T *Ut,
                                                const T *Umt, double *R64, double *U64, int n, int mixed)
{
  const int s = blockIdx.x*(TPB/SLICE) + (threadIdx.x >> 6);
  if (s >= nslices) return;
  const int p = slice_row<G>(row0, nrows, s);
  if (p < 0) return;
  const BatchOut &o = out[sscene[s]];
  if (mixed || o.final_buf)
  {
    const T *Rf = o.final_buf ? X1 : X0;
#pragma unroll
    for (int k = 0; k < 3; k ++) R64[4*(size_t)p + k] = (double)Rf[4*(size_t)p + k];
    R64[4*(size_t)p + 3] = 0.0;
  }
  if (mixed || o.final_U)
  {
    const T *Uf = o.final_U ? Umt : Ut;
#pragma unroll
    for (int k = 0; k < 3; k ++) U64[(size_t)k*n + p] = (double)Uf[(size_t)k*n + p];
  }
}

// U = B + W R (the merit's update of U, mrf.c:61-76) of the rows of every
// scene whose final U is a merit pass's (final_U = 1), from the final R in the
// canonical buffer: the per-colour sweeps do not store that U every sweep
template <int G>
__global__ __launch_bounds__(TPB) void k_umerit(MeritArgs a, int nslices, const int *sscene, const BatchOut *out)
{
  const int s = blockIdx.x*(TPB/SLICE) + (threadIdx.x >> 6);
  double up = 0.0;
  merit_slice<G>(a, s, s < nslices && out[sscene[s < nslices ? s : 0]].final_U == 1, up);
}

struct Tuned { SFGS_BATCH_TUNE t; int valid; };
static Tuned g_default = {{2, 4, 0, 1}, 0};

static bool tune_ok(const SFGS_BATCH_TUNE &t)
{
  return (t.lanes_per_row == 1 || t.lanes_per_row == 2) && (t.schedule == 0 || t.schedule == 1) &&
         (t.waves_per_group == 1 || t.waves_per_group == 2 || t.waves_per_group == 4 || t.waves_per_group == 8) &&
         t.rows_per_group >= 0;
}

} // namespace

// the per-colour schedule of the batch (k_usweep, k_uscene); b holds the
// sweep data and buffers as for the persistent kernel.  FP64: Gauss-Seidel
// sweeps until every scene has met the test of bgs.c:1036 or its iteration
// bound.  MIXED (T = float), as SFGS_Solve for one system: FP32 sweeps under
// the test with epsilon and meritval held above FP32 resolution, for at most
// maxiter - refine iterations; then, per scene, FP64 sweeps from that state
// under the reference's test, at least refine of them, within the scene's
// remaining iterations.
template <typename T, int G>
static int run_union(sfgs_ctx *ctx, const SFGS_PARAM *par, const BatchArgs<T> &b)
{
  hipStream_t st = ctx->stream;
  const int nsc = ctx->nscenes, ncol = ctx->ncolours, ns = ctx->nslices, n = ctx->nc;
  const bool mixed = par->precision == SFGS_PREC_MIXED;
  const int nomerit = par->meritval >= 1.0 ? 1 : 0, maxiter = par->maxiter;
  const int refine = mixed ? std::max(par->refine, 0) : 0;
  // slices listed colour by colour (all scenes), their scene and the rows
  // already updated when they run in a forward or a backward sweep
  std::vector<int> clist, cstart((size_t)ncol + 1, 0), sscene((size_t)std::max(ns, 1), 0),
                   sbound(4*(size_t)std::max(ns, 1), 0), uscene(2*(size_t)std::max(nsc, 1), 0);
  for (int c = 0; c < ncol; c ++)
  {
    cstart[c] = (int)clist.size();
    for (int q = 0; q < nsc; q ++)
    {
      const int g = q*ncol + c;
      for (int sl = ctx->h_cslice[g]; sl < ctx->h_cslice[g + 1]; sl ++)
      {
        clist.push_back(sl);
        sscene[sl] = q;
        sbound[2*(size_t)sl] = ctx->h_colptr[(size_t)q*ncol];            // forward: the scene's earlier colours
        sbound[2*(size_t)sl + 1] = ctx->h_colptr[g];
        const size_t bb = 2*(size_t)std::max(ns, 1) + 2*(size_t)sl;      // backward: its later colours
        sbound[bb] = ctx->h_colptr[g + 1];
        sbound[bb + 1] = ctx->h_colptr[(size_t)(q + 1)*ncol];
      }
    }
  }
  cstart[ncol] = (int)clist.size();
  for (int q = 0; q < nsc; q ++)
  {
    uscene[2*(size_t)q] = ctx->h_cslice[(size_t)q*ncol];
    uscene[2*(size_t)q + 1] = ctx->h_cslice[(size_t)(q + 1)*ncol];
  }
  const int maxsweeps = nomerit ? maxiter : maxiter + 1;
  CHK(upload(ctx, ctx->u_clist, clist.data(), std::max(clist.size(), (size_t)1)));
  CHK(upload(ctx, ctx->u_sscene, sscene.data(), sscene.size()));
  CHK(upload(ctx, ctx->u_sbound, sbound.data(), sbound.size()));
  CHK(upload(ctx, ctx->u_scene, uscene.data(), uscene.size()));
  CHK(ensure(ctx, ctx->u_spart, 3*(size_t)std::max(ns, 1)));
  CHK(ensure(ctx, ctx->u_active, (size_t)maxsweeps + 1));
  const UScene *usc = (const UScene*)ctx->u_scene.p;
  const int2 *sbf = (const int2*)ctx->u_sbound.p, *sbb = sbf + std::max(ns, 1);
  BatchOut *out = (BatchOut*)ctx->bout.p;
  double *spart = ctx->u_spart.p;
  const int sblk = (ns + TPB/SLICE - 1) / (TPB/SLICE);
  if (ns == 0 || nsc == 0) return 0;

  // free energy of every scene
  hipLaunchKernelGGL(k_uenergy<G>, dim3(sblk), dim3(TPB), 0, st, b.d, ns, spart);
  hipLaunchKernelGGL(k_uscene, dim3(nsc), dim3(TPB), 0, st, 0, usc, (const double*)spart, out, maxiter, 0, nomerit,
                     par->epsilon, par->meritval, 0, ctx->u_active.p);

  // sweeps of one phase until every scene has stopped; sweep kk reads X[kk & 1], writes the other
  auto phase = [&](auto base, auto *X0, auto *X1, int bound, int miniter, double eps, double meritval) -> int {
    using TT = typename std::remove_pointer<decltype(X0)>::type;
    HIPCHK(hipMemsetAsync(ctx->u_active.p, 0, sizeof(int)*((size_t)maxsweeps + 1), st));
    int k = 0, batch = 2;
    while (k < maxsweeps)
    {
      const int nbt = std::min(batch, maxsweeps - k);
      for (int j = 0; j < nbt; j ++)
      {
        const int kk = k + j, back = par->reverse && (kk & 1);
        SweepArgs<TT> a = base;
        // a merit sweep's U is final only if FAILURE_EXIT stops it; the merit
        // pass's U is formed again from the final R (k_umerit)
        a.store_U = a.failure == SFGS_FAILURE_EXIT;
        a.store_Um = 0;
        a.P = (kk & 1) ? X1 : X0;
        a.Q = (kk & 1) ? X0 : X1;
        const bool merit = !nomerit && kk > 0;
        for (int ci = 0; ci < ncol; ci ++)
        {
          const int c = back ? ncol - 1 - ci : ci, nl = cstart[c + 1] - cstart[c];
          if (nl == 0) continue;
          const dim3 grid((nl + TPB/SLICE - 1) / (TPB/SLICE));
          if (merit)
            hipLaunchKernelGGL((k_usweep<TT, G, true>), grid, dim3(TPB), 0, st, a, ctx->u_clist.p + cstart[c], nl,
                               back ? sbb : sbf, (const int*)ctx->u_sscene.p, out, spart, back);
          else
            hipLaunchKernelGGL((k_usweep<TT, G, false>), grid, dim3(TPB), 0, st, a, ctx->u_clist.p + cstart[c], nl,
                               back ? sbb : sbf, (const int*)ctx->u_sscene.p, out, spart, back);
        }
        hipLaunchKernelGGL(k_uscene, dim3(nsc), dim3(TPB), 0, st, 1, usc, (const double*)spart, out, bound, miniter,
                           nomerit, eps, meritval, kk, ctx->u_active.p);
      }
      k += nbt;
      int act = 0;
      HIPCHK(hipMemcpyAsync(&act, ctx->u_active.p + (k - 1), sizeof(int), hipMemcpyDeviceToHost, st));
      HIPCHK(hipStreamSynchronize(st));
      HIPCHK(hipGetLastError());
      if (act == 0) break;
      batch = std::min(batch*2, 32);
    }
    return 0;
  };

  // the final U of the scenes that stopped on the test, from the canonical R
  const MeritArgs ma = merit_args(ctx, 1, nullptr);
  auto merit_U = [&]() {
    hipLaunchKernelGGL(k_umerit<G>, dim3(sblk), dim3(TPB), 0, st, ma, ns, (const int*)ctx->u_sscene.p,
                       (const BatchOut*)out);
  };
  if (!mixed)
  {
    CHK(phase(b.a, b.X0, b.X1, maxiter, 0, par->epsilon, par->meritval));
    // the final state into the canonical buffers
    hipLaunchKernelGGL((k_ufinal<T, G>), dim3(sblk), dim3(TPB), 0, st, ns, (const int*)ctx->slice_row0.p,
                       (const int*)ctx->slice_nrows.p, (const int*)ctx->u_sscene.p, (const BatchOut*)out,
                       (const T*)b.X0, (const T*)b.X1, (const T*)b.Ut, (const T*)b.Umt, b.R64, b.U64, n, 0);
    merit_U();
    HIPCHK(hipGetLastError());
    return 0;
  }
  // FP32 phase (none when maxiter <= refine: the whole solve is FP64, from the FP64 state)
  if (maxiter > refine)
  {
    CHK(phase(b.a, b.X0, b.X1, maxiter - refine, 0, std::max(par->epsilon, 1e-6), std::max(par->meritval, 1e-10)));
    hipLaunchKernelGGL((k_ufinal<T, G>), dim3(sblk), dim3(TPB), 0, st, ns, (const int*)ctx->slice_row0.p,
                       (const int*)ctx->slice_nrows.p, (const int*)ctx->u_sscene.p, (const BatchOut*)out,
                       (const T*)b.X0, (const T*)b.X1, (const T*)b.Ut, (const T*)b.Umt, b.R64, b.U64, n, 1);
    merit_U();   // in FP64, the FP64 phase's start
  }
  // FP64 phase per scene under the reference's test
  hipLaunchKernelGGL(k_uphase, dim3((nsc + TPB - 1)/TPB), dim3(TPB), 0, st, nsc, usc, maxiter, out);
  CHK(phase(b.d, b.R64, b.R64b, maxiter, refine, par->epsilon, par->meritval));
  hipLaunchKernelGGL(k_ufinal64<G>, dim3(sblk), dim3(TPB), 0, st, ns, (const int*)ctx->slice_row0.p,
                     (const int*)ctx->slice_nrows.p, (const int*)ctx->u_sscene.p, (const BatchOut*)out,
                     (const double*)b.R64b, (const double*)b.Um64, b.R64, b.U64, n);
  merit_U();
  HIPCHK(hipGetLastError());
  return 0;
}

// one launch over all scenes (several if the scenes need more workgroups than
// can be resident at once); ms receives the device time
static int batch_solve(sfgs_ctx *ctx, const SFGS_PARAM *par, const SFGS_BATCH_TUNE &tn, SFGS_SCENE_REPORT *reps,
                       SFGS_REPORT *total, float *ms_out)
{
  if (!ctx->assembled) { ctx->err = "solve before LOCDYN assembly"; return SFGS_E_STATE; }
  if (ctx->diag_only) { ctx->err = "the store holds a penalty solve, not an assembled system"; return SFGS_E_STATE; }
  if (ctx->upmin) { ctx->err = "the local dynamics hold the diagonal blocks only (SFGS_Locdyn_Update_Min)"; return SFGS_E_STATE; }
  if (tn.lanes_per_row != ctx->G_built)
  {
    ctx->err = "batch shape has " + std::to_string(tn.lanes_per_row) + " lanes per row, the system was assembled with " +
               std::to_string(ctx->G_built);
    return SFGS_E_STATE;
  }
  HIPCHK(hipSetDevice(ctx->device));
  hipStream_t st = ctx->stream;
  const int nsc = ctx->nscenes, ncol = ctx->ncolours, n = ctx->nc;
  const bool mixed = par->precision == SFGS_PREC_MIXED;
  const int G = tn.lanes_per_row, NW = tn.waves_per_group;
  const int ng = nsc*ncol;
  if ((int)ctx->h_colptr.size() != ng + 1 || (int)ctx->h_cslice.size() != ng + 1)
  {
    ctx->err = "layout tables do not match the scenes";
    return SFGS_E_STATE;
  }

  // workgroups per scene (the persistent launch runs FP64 solves only)
  int cap = 0;
  CHK((dispatch_batch<double>(ctx, G, NW, BatchArgs<double>(), 0, &cap)));
  if (cap < 1) { ctx->err = "batched kernel cannot be resident"; return SFGS_E_HIP; }
  std::vector<int> rows((size_t)nsc), want((size_t)nsc), slices((size_t)nsc);
  long long total_rows = 0;
  int busy = 0;
  for (int s = 0; s < nsc; s ++)
  {
    rows[s] = ctx->h_colptr[(size_t)(s + 1)*ncol] - ctx->h_colptr[(size_t)s*ncol];
    slices[s] = ctx->h_cslice[(size_t)(s + 1)*ncol] - ctx->h_cslice[(size_t)s*ncol];
    total_rows += rows[s];
    busy += rows[s] > 0 ? 1 : 0;
  }
  // rows_per_group 0: share the resident capacity among the scenes in proportion to their rows,
  // so that every scene runs in one launch and takes about the same time per iteration
  const bool balanced = tn.rows_per_group <= 0 && busy <= cap;
  for (int s = 0; s < nsc; s ++)
  {
    long long w;
    if (rows[s] == 0) w = 0;
    else if (balanced) w = (long long)((double)rows[s] / (double)total_rows * (double)(cap - busy)) + 1;
    else w = ((long long)rows[s] + std::max(tn.rows_per_group, 1) - 1) / std::max(tn.rows_per_group, 1);
    w = std::min<long long>(w, std::max(1, (slices[s] + NW - 1) / NW));   // no more workgroups than slice rounds
    want[s] = (int)std::max<long long>(rows[s] > 0 ? 1 : 0, std::min<long long>(w, cap));
  }

  // FP32 mirrors and the first-phase R / U
  if (mixed)
  {
    CHK(build_f32_mirror(ctx));
    if (n > 0)
    {
      hipLaunchKernelGGL(k_batch_to_f32, dim3(nblk_ll(4LL*n)), dim3(TPB), 0, st, 4LL*n, ctx->R.p, ctx->R32.p);
      hipLaunchKernelGGL(k_batch_to_f32, dim3(nblk_ll(3LL*n)), dim3(TPB), 0, st, 3LL*n, ctx->U.p, ctx->U32.p);
    }
  }
  CHK(ensure(ctx, ctx->Rb, 4*(size_t)std::max(n, 1)));
  CHK(upload(ctx, ctx->bgptr, ctx->h_colptr.data(), (size_t)ng + 1));
  CHK(upload(ctx, ctx->bgslice, ctx->h_cslice.data(), (size_t)ng + 1));
  CHK(ensure(ctx, ctx->bout, sizeof(BatchOut)*(size_t)std::max(nsc, 1)));
  CHK(ensure(ctx, ctx->bbar, (size_t)std::max(nsc, 1)));
  CHK(ensure(ctx, ctx->babort, 4));
  HIPCHK(hipMemsetAsync(ctx->bout.p, 0, sizeof(BatchOut)*(size_t)std::max(nsc, 1), st));
  HIPCHK(hipMemsetAsync(ctx->bbar.p, 0, sizeof(unsigned)*(size_t)std::max(nsc, 1), st));
  HIPCHK(hipMemsetAsync(ctx->babort.p, 0, sizeof(int)*4, st));

  SFGS_PARAM pd = *par;
  BatchArgs<double> bd;
  BatchArgs<float> bf;
  auto fill = [&](auto &b) {
    b.d = sweep_args<double>(ctx, &pd, ctx->W.p, ctx->Wd.p, ctx->B.p, ctx->rho.p, ctx->pfri.p, ctx->pcoh.p, ctx->shift.p,
                             ctx->U.p, ctx->Um.p, par->diagepsilon);
    b.geo.gslice = ctx->bgslice.p; b.geo.gptr = ctx->bgptr.p; b.geo.ncol = ncol;
    b.geo.abort_all = ctx->babort.p; b.geo.spin_limit = 1u << 22;
    b.R64 = ctx->R.p; b.R64b = ctx->Rb.p; b.U64 = ctx->U.p; b.Um64 = ctx->Um.p;
    b.out = (BatchOut*)ctx->bout.p; b.bar = ctx->bbar.p;
    b.maxiter = par->maxiter; b.nomerit = par->meritval >= 1.0 ? 1 : 0; b.reverse = par->reverse;
    b.eps = par->epsilon; b.meritval = par->meritval;
  };
  fill(bd);
  fill(bf);
  bd.a = bd.d;
  bd.X0 = ctx->R.p; bd.X1 = ctx->Rb.p; bd.Ut = ctx->U.p; bd.Umt = ctx->Um.p;
  if (mixed)
  {
    // FP32 resolves a relative change of about 1e-7: keep the local tolerance above it
    bf.a = sweep_args<float>(ctx, &pd, ctx->W32.p, ctx->Wd32.p, ctx->B32.p, ctx->rho32.p, ctx->pfri32.p, ctx->pcoh32.p,
                             ctx->shift32.p, ctx->U32.p, ctx->Um32.p, std::max(par->diagepsilon, 1e-6));
    bf.X0 = ctx->R32.p; bf.X1 = ctx->R32b.p; bf.Ut = ctx->U32.p; bf.Umt = ctx->Um32.p;
  }

  // launches over chunks of scenes that fit the resident capacity, or the per-colour
  // schedule, which every MIXED solve takes (its FP64 phase runs per scene)
  std::vector<int> wgs((size_t)nsc, 0);
  const bool per_colour = tn.schedule == 1 || mixed;
  HIPCHK(hipEventRecord(ctx->ev0, st));
  if (per_colour)
  {
    int rc = 0;
    if (mixed) rc = G == 2 ? run_union<float, 2>(ctx, par, bf) : run_union<float, 1>(ctx, par, bf);
    else rc = G == 2 ? run_union<double, 2>(ctx, par, bd) : run_union<double, 1>(ctx, par, bd);
    if (rc) return rc;
  }
  int s = per_colour ? nsc : 0;
  while (s < nsc)
  {
    std::vector<BatchScene> tab((size_t)nsc);
    std::vector<int> wg_scene;
    int used = 0, first = s;
    for (; s < nsc; s ++)
    {
      if (want[s] == 0) { tab[s] = BatchScene{0, 0, 0, 0}; continue; }
      if (used + want[s] > cap && used > 0) break;
      const int nw = std::min(want[s], cap - used);
      tab[s] = BatchScene{used, nw, ctx->h_colptr[(size_t)s*ncol], ctx->h_colptr[(size_t)(s + 1)*ncol]};
      for (int q = 0; q < nw; q ++) wg_scene.push_back(s);
      used += nw;
      wgs[s] = nw;
    }
    if (used == 0) continue;
    (void)first;
    CHK(upload(ctx, ctx->bscene_tab, (const int*)tab.data(), tab.size()*sizeof(BatchScene)/sizeof(int)));
    CHK(upload(ctx, ctx->bwg_scene, wg_scene.data(), wg_scene.size()));
    CHK(ensure(ctx, ctx->bpart, 8*(size_t)used));
    bd.scenes = (const BatchScene*)ctx->bscene_tab.p;
    bd.wg_scene = ctx->bwg_scene.p;
    bd.wpart = ctx->bpart.p;
    CHK(dispatch_batch<double>(ctx, G, NW, bd, used, nullptr));
    // the tables are re-uploaded for the next chunk only after this launch has run
    HIPCHK(hipStreamSynchronize(st));
    HIPCHK(hipGetLastError());
  }
  HIPCHK(hipEventRecord(ctx->ev1, st));
  std::vector<BatchOut> out((size_t)std::max(nsc, 1));
  int timed_out = 0;
  HIPCHK(hipMemcpyAsync(out.data(), ctx->bout.p, sizeof(BatchOut)*(size_t)nsc, hipMemcpyDeviceToHost, st));
  HIPCHK(hipMemcpyAsync(&timed_out, ctx->babort.p, sizeof(int), hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  HIPCHK(hipGetLastError());
  float ms = 0.0f;
  HIPCHK(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
  if (ms_out) *ms_out = ms;
  if (timed_out) { ctx->err = "a scene barrier of the batched kernel timed out"; return SFGS_E_STATE; }

  int maxit = 0;
  long long updates = 0;
  for (int q = 0; q < nsc; q ++)
  {
    const BatchOut &b = out[q];
    if (reps)
    {
      SFGS_SCENE_REPORT &r = reps[q];
      std::memset(&r, 0, sizeof(r));
      r.rows = rows[q];
      r.groups = wgs[q];
      r.iters = b.iters;
      r.error = b.error;
      r.merit = b.merit;
      r.energy = b.energy;
      r.fail_rows = b.failcount;
      r.aborted = b.aborted;
      r.error_code = SFGS_GS_OK;
      if (b.failkey) r.error_code = (int)(b.failkey & 3ULL);
      // bgs.c:1053: the iteration bound reached without an abort
      if (rows[q] > 0 && !b.aborted && b.iters - b.nref >= par->maxiter) r.error_code = SFGS_GS_DIVERGED;
    }
    maxit = std::max(maxit, b.iters);
    updates += (long long)(b.sweeps0 + b.sweeps) * rows[q];   // both phases of MIXED
  }
  if (total)
  {
    std::memset(total, 0, sizeof(*total));
    total->iters = maxit;
    total->ncolours = ncol;
    total->solve_ms = ms;
    total->row_updates = updates;
    total->nnz = ctx->nnz;
  }
  return 0;
}

extern "C" {

int SFGS_Batch_Solve(sfgs_ctx *ctx, const SFGS_PARAM *par, const SFGS_BATCH_TUNE *tune, SFGS_SCENE_REPORT *reports,
                     SFGS_REPORT *total)
{
  if (!ctx || !par || par->maxiter <= 0 || !(par->epsilon > 0.0) || par->diagmaxiter <= 0) return SFGS_E_ARG;
  if (par->diagsolver < 0 || par->diagsolver > 3 || par->failure < 0 || par->failure > 2) return SFGS_E_ARG;
  if (par->precision != SFGS_PREC_FP64 && par->precision != SFGS_PREC_MIXED) return SFGS_E_ARG;
  if (par->precision == SFGS_PREC_MIXED && par->refine < 0) return SFGS_E_ARG;
  SFGS_BATCH_TUNE tn = tune ? *tune : g_default.t;
  if (!tune) tn.lanes_per_row = ctx->G_built;
  if (!tune_ok(tn)) return SFGS_E_ARG;
  return batch_solve(ctx, par, tn, reports, total, nullptr);
}

int SFGS_Batch_Autotune(sfgs_ctx *ctx, const SFGS_PARAM *par, int trial_iters, SFGS_BATCH_TUNE *best, double *best_ms)
{
  if (!ctx || !par || trial_iters <= 0) return SFGS_E_ARG;
  if (!ctx->assembled || !ctx->from_bodies || ctx->diag_only) { ctx->err = "autotuning needs a system assembled from bodies"; return SFGS_E_STATE; }
  HIPCHK(hipSetDevice(ctx->device));
  hipStream_t st = ctx->stream;
  const int n = ctx->nc;
  // the trials run a fixed number of iterations from the current R and U, restored after each
  DBuf<double> R0, U0;
  CHK(ensure(ctx, R0, 4*(size_t)std::max(n, 1)));
  CHK(ensure(ctx, U0, 3*(size_t)std::max(n, 1)));
  if (n > 0)
  {
    HIPCHK(hipMemcpyAsync(R0.p, ctx->R.p, sizeof(double)*4*(size_t)n, hipMemcpyDeviceToDevice, st));
    HIPCHK(hipMemcpyAsync(U0.p, ctx->U.p, sizeof(double)*3*(size_t)n, hipMemcpyDeviceToDevice, st));
  }
  SFGS_PARAM tp = *par;
  tp.maxiter = trial_iters;
  tp.epsilon = 1e-300;        // no early stop: every shape runs the same iterations
  tp.meritval = par->meritval < 1.0 ? 1e-300 : par->meritval;
  tp.failure = SFGS_FAILURE_CONTINUE;
  const int G0 = ctx->G;
  SFGS_BATCH_TUNE bt = g_default.t;
  double bms = 1e300;
  int rc = 0;
  const int lanes[2] = {1, 2}, waves[4] = {1, 2, 4, 8}, rpg[4] = {0, 1024, 4096, 1 << 30};
  for (int li = 0; li < 2 && !rc; li ++)
  {
    if (ctx->G_built != lanes[li])
    {
      ctx->G = lanes[li];
      if ((rc = reassemble(ctx))) break;
    }
    // shape index 16 is the per-colour schedule (no workgroup shape), the only one of MIXED
    for (int wr = par->precision == SFGS_PREC_MIXED ? 16 : 0; wr <= 16 && !rc; wr ++)
      {
        const int wi = wr / 4, ri = wr % 4;
        SFGS_BATCH_TUNE t = wr == 16 ? SFGS_BATCH_TUNE{lanes[li], 4, 0, 1}
                                     : SFGS_BATCH_TUNE{lanes[li], waves[wi], rpg[ri], 0};
        if (n > 0)
        {
          HIPCHK(hipMemcpyAsync(ctx->R.p, R0.p, sizeof(double)*4*(size_t)n, hipMemcpyDeviceToDevice, st));
          HIPCHK(hipMemcpyAsync(ctx->U.p, U0.p, sizeof(double)*3*(size_t)n, hipMemcpyDeviceToDevice, st));
        }
        float ms = 0.0f;
        if ((rc = batch_solve(ctx, &tp, t, nullptr, nullptr, &ms))) break;
        if (ms < bms) { bms = ms; bt = t; }
      }
  }
  // keep the best layout and restore the state
  if (!rc && ctx->G_built != bt.lanes_per_row)
  {
    ctx->G = bt.lanes_per_row;
    rc = reassemble(ctx);
  }
  if (rc) ctx->G = G0;
  if (!rc && n > 0)
  {
    HIPCHK(hipMemcpyAsync(ctx->R.p, R0.p, sizeof(double)*4*(size_t)n, hipMemcpyDeviceToDevice, st));
    HIPCHK(hipMemcpyAsync(ctx->U.p, U0.p, sizeof(double)*3*(size_t)n, hipMemcpyDeviceToDevice, st));
    HIPCHK(hipStreamSynchronize(st));
  }
  release(R0);
  release(U0);
  if (rc) return rc;
  g_default.t = bt;
  g_default.valid = 1;
  if (best) *best = bt;
  if (best_ms) *best_ms = bms;
  return 0;
}

} // extern "C"
