// sfgs_sweep.hpp -- device code of the colour-major Gauss-Seidel sweep and of
// the merit pass, shared by the per-colour kernels (sfgs_kernels.hip) and the
// persistent batched kernel (sfgs_batch.hip).
#pragma once
#include <hip/hip_runtime.h>
#include "sfgs_ctx.hpp"
#include "sfgs_math.hpp"

namespace sfgs {

static const int TPB = 256;          // threads per block (4 wavefronts)
static const int SLICE = 64;         // rows per slice (one wavefront)

template <typename T>
__device__ __forceinline__ T wave_sum(T v)
{
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// =====================================================================
// Gauss-Seidel sweep of one colour (bgs.c:950-1019) with the merit of the
// previous iteration fused in (mrf.c:35-184)
//
// R lives in two buffers.  Sweep k reads P (R at the end of sweep k-1, all
// rows) and writes Q.  Row i gathers Q_j for neighbours already updated in
// this sweep (permuted index in [upd_lo, upd_hi)) and P_j otherwise, which is
// the serial colour-major Gauss-Seidel order.  When the merit is on, sweep
// k >= 1 also forms U_i = B_i + W_ii P_i + sum_j W_ij P_j, i.e. the merit
// update of U after iteration k-1 (mrf.c:61-76), accumulates C'AC for it, keeps
// that U in Um and starts the local solve from it, exactly as the reference's
// next sweep does after MERIT_Function.  The termination test of iteration k-1
// therefore completes after sweep k; when it stops, P and Um are the state.
// G lanes cooperate on one row (blocks k = g, g+G, ...), reduce B' across the
// group with cross-lane adds, and run the local solve redundantly; lane g = 0
// writes.
// =====================================================================

template <typename T>
struct SweepArgs {
  const int *row0, *nrows, *rowlen;
  const int *slen;      // block slots per lane of each slice (those of its first, longest row)
  const long long *off;
  const T *W;
  const int *wcol;
  const T *Wd, *B, *rho, *fri, *coh, *shift;
  const double *A;      // merit weights (FP64 sweeps)
  const float4 *rec;    // FP32 sweeps: the row records (B, law, W_ii, A; k_rec32); the merit arithmetic stays FP64
  const unsigned char *sep;
  const T *P;
  T *Q, *U, *Um;
  const T *Uin;         // local-solver start U (U itself except in the dataflow schedule)
  int n;
  T diageps;
  int diagmaxiter, diagsolver, failure;
  int store_U;          // write the local solver's U; a merit sweep's U is not the final state unless
                        // the solve can abort inside it (FAILURE_EXIT), then U is that state
  int store_Um;         // write the merit pass's U (Um); a caller that forms the final U = B + W R
                        // afterwards (k_merit with update_U) leaves it off
  double *partials;     // 3 per block: errup, errlo, merit numerator
  unsigned long long *failkey;
  int *failcount;
  Ctrl *ctrl;
};

template <int G, typename T>
__device__ __forceinline__ T gsum(T v)
{
  if (G >= 2) v += __shfl_xor(v, 1, 64);
  if (G >= 4) v += __shfl_xor(v, 2, 64);
  if (G >= 8) v += __shfl_xor(v, 4, 64);
  if (G >= 16) v += __shfl_xor(v, 8, 64);
  return v;
}

// deterministic block sum of three values over a block of NT threads; thread 0
// returns them (the shared slots may be reused after the next barrier)
template <int NT = TPB>
__device__ __forceinline__ void block_sum3(double &a, double &b, double &c)
{
  __shared__ double sa[NT/64], sb[NT/64], sc[NT/64];
  a = wave_sum(a); b = wave_sum(b); c = wave_sum(c);
  int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) { sa[w] = a; sb[w] = b; sc[w] = c; }
  __syncthreads();
  if (threadIdx.x == 0)
  {
    a = sa[0]; b = sb[0]; c = sc[0];
    for (int k = 1; k < NT/64; k ++) { a += sa[k]; b += sb[k]; c += sc[k]; }
  }
}

// Reactions shared between wavefronts of one persistent launch (the dataflow
// schedule) are read and written with device-scope relaxed atomics: those bypass
// the per-XCD L2's stale lines without the cache-wide write-back / invalidate
// that a device-scope fence costs.  Ordinary accesses otherwise.
template <bool COH, typename T>
__device__ __forceinline__ T r_ld(const T *p)
{
  if (COH) return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  return *p;
}

template <bool COH, typename T>
__device__ __forceinline__ void r_st(T *p, T v)
{
  if (COH) __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  else *p = v;
}

// One slice (64/G rows of one colour, one wavefront) of a sweep.  eu, el and
// up receive this lane's contributions to the error sums and to the merit
// numerator.  A failed local solve is recorded in failkey / failcount, keyed by
// the sweep number *sweepno, and raises *abortflag under FAILURE_EXIT.  run =
// false (a solve that has stopped) leaves the slice untouched; the caller reads
// that flag together with the slice table, so both cost one memory round trip.
// COH: P and Q are shared with other wavefronts of the same launch (r_ld / r_st).
template <typename T, int G, bool MERIT, bool COH = false>
__device__ __forceinline__ void sweep_slice(const SweepArgs<T> &a, int s, bool valid, int upd_lo, int upd_hi,
                                            int backward, const int *sweepno, unsigned long long *failkey,
                                            int *failcount, int *abortflag, double &eu, double &el, double &up,
                                            bool run = true)
{
  const int l = threadIdx.x & 63, r = l / G, g = l % G;
  const int n = a.n;
  int nr = 0, r0 = 0, p = 0, len = 0;
  long long off = 0;
  if (valid) { nr = a.nrows[s]; r0 = a.row0[s]; off = a.off[s]; }   // one round of loads
  const bool act = valid && run && r < nr;
  // the row's own data first, so that its loads overlap the gather below
  T Bi[3] = {T(0), T(0), T(0)}, Wi[9], R[3] = {T(0), T(0), T(0)}, U[3] = {T(0), T(0), T(0)};
  double Ai[9];
  RowLaw<T> L{T(0), T(0), T(0), T(0), false};
  if (act)
  {
    p = r0 + r;
    if constexpr (sizeof(T) == 4)
    {
      // the row record: five 16-byte loads (k_rec32), and R_i as one when not shared
      const float4 *rc = a.rec + 5*(size_t)p;
      const float4 q0 = rc[0], q1 = rc[1], q2 = rc[2], q3 = rc[3], q4 = rc[4];
      const int ls = __float_as_int(q1.w);
      len = ls >> 1;
      Bi[0] = q0.x; Bi[1] = q0.y; Bi[2] = q0.z;
      Wi[0] = q2.x; Wi[1] = Wi[3] = q2.y; Wi[2] = Wi[6] = q2.z; Wi[4] = q2.w; Wi[5] = Wi[7] = q3.x; Wi[8] = q3.y;
      if (MERIT)
      {
        Ai[0] = q3.z; Ai[1] = Ai[3] = q3.w; Ai[2] = Ai[6] = q4.x; Ai[4] = q4.y; Ai[5] = Ai[7] = q4.z; Ai[8] = q4.w;
      }
      L = RowLaw<T>{q1.x, q1.y, q1.z, q0.w, (ls & 1) != 0};
      if (COH)
      {
#pragma unroll
        for (int k = 0; k < 3; k ++) R[k] = r_ld<COH>(a.P + 4*(size_t)p + k);
      }
      else
      {
        const float4 rr = *reinterpret_cast<const float4 *>(a.P + 4*(size_t)p);
        R[0] = rr.x; R[1] = rr.y; R[2] = rr.z;
      }
    }
    else
    {
      len = a.rowlen[p];
      Bi[0] = a.B[p]; Bi[1] = a.B[(size_t)n + p]; Bi[2] = a.B[2*(size_t)n + p];
#pragma unroll
      for (int e = 0; e < 9; e ++) Wi[e] = a.Wd[(size_t)e*n + p];
#pragma unroll
      for (int k = 0; k < 3; k ++) R[k] = r_ld<COH>(a.P + 4*(size_t)p + k);
      if (MERIT)
      {
#pragma unroll
        for (int e = 0; e < 9; e ++) Ai[e] = a.A[(size_t)e*n + p];
      }
      L = RowLaw<T>{a.fri[p], a.coh[p], a.shift[p], a.rho[p], a.sep[p] != 0};
    }
    if (!MERIT)
    {
#pragma unroll
      for (int k = 0; k < 3; k ++) U[k] = a.Uin[(size_t)k*n + p];
    }
  }
  T b0 = T(0), b1 = T(0), b2 = T(0), u0 = T(0), u1 = T(0), u2 = T(0);
  if (act)
  {
    // The gather runs over the slice's slot count, known with the slice table,
    // so the neighbour indices and blocks load without waiting for this row's
    // length; slots past the row's own (padding: zero block, row 0) are loaded
    // but not added.  CH slots at a time: all their indices, then all their
    // blocks and reactions, in flight together.
    constexpr int CH = sizeof(T) == 8 ? 4 : 8;
    const T *Wb = a.W + 9*off + l;
    const int *cb = a.wcol + off + l;
    const int nsl = a.slen[s], mine = len > g ? (len - g + G - 1) / G : 0;
    for (int m0 = 0; m0 < nsl; m0 += CH)
    {
      // branch-free: every load of the chunk is issued before any is waited for
      // (slots past the slice's own re-read its last slot and are not added)
      int jj[CH];
#pragma unroll
      for (int q = 0; q < CH; q ++) jj[q] = m0 + q < nsl ? cb[(size_t)(m0 + q)*SLICE] : 0;
      // the blocks do not depend on the indices: issued with them, before any wait
      T wv[CH][9], rv[CH][3], pv[CH][3];
#pragma unroll
      for (int q = 0; q < CH; q ++)
      {
        const int m = min(m0 + q, nsl - 1);
        const T *w = Wb + 9*(size_t)m*SLICE;
#pragma unroll
        for (int e = 0; e < 9; e ++) wv[q][e] = w[e*SLICE];
      }
      // keep the block loads ahead of the first wait for the indices
      __builtin_amdgcn_sched_barrier(0);
#pragma unroll
      for (int q = 0; q < CH; q ++)
      {
        const int j = jj[q];
        const bool upd = j >= upd_lo && j < upd_hi;
        const T *xj = (upd ? a.Q : a.P) + 4*(size_t)j;
        rv[q][0] = r_ld<COH>(xj); rv[q][1] = r_ld<COH>(xj + 1); rv[q][2] = r_ld<COH>(xj + 2);
        if (MERIT)
        {
          // the previous iterate of every neighbour (equal to rv where not yet updated)
          const T *pj = a.P + 4*(size_t)j;
          pv[q][0] = r_ld<COH>(pj); pv[q][1] = r_ld<COH>(pj + 1); pv[q][2] = r_ld<COH>(pj + 2);
        }
      }
#pragma unroll
      for (int q = 0; q < CH; q ++)
      {
        const bool on = m0 + q < mine;
        const T w0 = wv[q][0], w1 = wv[q][1], w2 = wv[q][2], w3 = wv[q][3], w4 = wv[q][4], w5 = wv[q][5],
                w6 = wv[q][6], w7 = wv[q][7], w8 = wv[q][8];
        const T r0 = rv[q][0], r1 = rv[q][1], r2 = rv[q][2];
        // explicit fused multiply-adds: the same rounding in every schedule
        // that instantiates this gather, whatever the surrounding loads
        b0 = on ? b0 + fma(w6, r2, fma(w3, r1, w0*r0)) : b0;
        b1 = on ? b1 + fma(w7, r2, fma(w4, r1, w1*r0)) : b1;
        b2 = on ? b2 + fma(w8, r2, fma(w5, r1, w2*r0)) : b2;
        if (MERIT)
        {
          const T p0 = pv[q][0], p1 = pv[q][1], p2 = pv[q][2];
          u0 = on ? u0 + fma(w6, p2, fma(w3, p1, w0*p0)) : u0;
          u1 = on ? u1 + fma(w7, p2, fma(w4, p1, w1*p0)) : u1;
          u2 = on ? u2 + fma(w8, p2, fma(w5, p1, w2*p0)) : u2;
        }
      }
    }
  }
  // sum over the G lanes of each row; every lane of the wave takes part
  b0 = gsum<G>(b0); b1 = gsum<G>(b1); b2 = gsum<G>(b2);
  if (MERIT) { u0 = gsum<G>(u0); u1 = gsum<G>(u1); u2 = gsum<G>(u2); }

  if (act)
  {
    T Bp[3] = {Bi[0] + b0, Bi[1] + b1, Bi[2] + b2};
    const T R0[3] = {R[0], R[1], R[2]};
    if (MERIT)
    {
      // merit of the previous iteration at this row (mrf.c:63-84)
      mv_add(Bi, Wi, R, U);
      U[0] += u0; U[1] += u1; U[2] += u2;
      // the merit is evaluated in FP64 whatever the sweep type
      double C[3], Ud[3] = {U[0], U[1], U[2]}, Rd[3] = {R[0], R[1], R[2]};
      scf_residual<double>(L.fri, L.coh, L.shift, Ud, Rd, C, nullptr, nullptr);
      const double q0 = Ai[0]*C[0] + Ai[3]*C[1] + Ai[6]*C[2];
      const double q1 = Ai[1]*C[0] + Ai[4]*C[1] + Ai[7]*C[2];
      const double q2 = Ai[2]*C[0] + Ai[5]*C[1] + Ai[8]*C[2];
      if (g == 0)
      {
        up += q0*C[0] + q1*C[1] + q2*C[2];
        if (a.store_Um)
        {
#pragma unroll
          for (int k = 0; k < 3; k ++) a.Um[(size_t)k*n + p] = U[k];
        }
      }
    }

    int d = local_solve(a.diagsolver, L, a.diageps, a.diagmaxiter, Wi, Bp, U, R);
    if (d >= a.diagmaxiter || d < 0)
    {
      if (g == 0)
      {
        unsigned code = d < 0 ? SFGS_GS_DIAGONAL_FAILED : SFGS_GS_DIAGONAL_DIVERGED;
        unsigned long long pos = backward ? (unsigned long long)(n - 1 - p) : (unsigned long long)p;
        atomicMax(failkey, ((unsigned long long)(r_ld<COH>(sweepno) + 1) << 40) | (pos << 2) | code);
        atomicAdd(failcount, 1);
        if (a.failure == SFGS_FAILURE_EXIT) atomicExch(abortflag, 1);
      }
      if (a.failure == SFGS_FAILURE_CONTINUE)
      {
        const int chain[4] = {SFGS_DS_SEMISMOOTH_NEWTON, SFGS_DS_PROJECTED_GRADIENT,
                              SFGS_DS_DE_SAXCE_FENG, SFGS_DS_PROJECTED_NEWTON};
        for (int c = 0; c < 4; c ++)
        {
          if (chain[c] == a.diagsolver) continue;
          R[0] = R0[0]; R[1] = R0[1]; R[2] = R0[2];
          d = local_solve(chain[c], L, a.diageps, a.diagmaxiter, Wi, Bp, U, R);
          if (d < a.diagmaxiter && d >= 0) break;
        }
        if (d >= a.diagmaxiter || d < 0) { R[0] = R0[0]; R[1] = R0[1]; R[2] = R0[2]; }
      }
    }
    if (g == 0)
    {
      if constexpr (sizeof(T) == 4 && !COH)
        *reinterpret_cast<float4 *>(a.Q + 4*(size_t)p) = make_float4(R[0], R[1], R[2], 0.0f);
      else
      {
#pragma unroll
        for (int k = 0; k < 3; k ++) r_st<COH>(a.Q + 4*(size_t)p + k, R[k]);
        r_st<COH>(a.Q + 4*(size_t)p + 3, T(0));
      }
      if (!MERIT || a.store_U)
      {
#pragma unroll
        for (int k = 0; k < 3; k ++) a.U[(size_t)k*n + p] = U[k];
      }
      const double d0 = (double)R[0] - R0[0], d1 = (double)R[1] - R0[1], d2 = (double)R[2] - R0[2];
      eu += d0*d0 + d1*d1 + d2*d2;
      el += (double)R[0]*R[0] + (double)R[1]*R[1] + (double)R[2]*R[2];
    }
  }
}

// standalone merit (MERIT_Function with update_U, mrf.c:35-184) on the
// canonical R buffer; one launch over all slices
struct MeritArgs {
  const int *row0, *nrows, *rowlen;
  const long long *off;
  const double *W;
  const int *wcol;
  const double *Wd, *A, *B, *fri, *coh, *shift;
  const double *R;
  double *U;
  int n, update_U;
  double *mpartials;
  double *conmerit;   // may be null: per-row C'AC (colour-major)
};

// merit terms C'AC of the rows of one slice, added to up (lane g = 0 of each row)
template <int G>
__device__ __forceinline__ void merit_slice(const MeritArgs &a, int s, bool valid, double &up)
{
  const int l = threadIdx.x & 63, r = l / G, g = l % G;
  const int n = a.n;
  bool act = false;
  int p = 0, len = 0;
  long long off = 0;
  if (valid)
  {
    act = r < a.nrows[s];
    off = a.off[s];
    if (act) { p = a.row0[s] + r; len = a.update_U ? a.rowlen[p] : 0; }
  }
  double u0 = 0.0, u1 = 0.0, u2 = 0.0;
  if (act)
  {
    const double *Wb = a.W + 9*off + l;
    const int *cb = a.wcol + off + l;
    for (int k = g, m = 0; k < len; k += G, m ++)
    {
      const int j = cb[(size_t)m*SLICE];
      const double *w = Wb + 9*(size_t)m*SLICE;
      const double r0 = a.R[4*(size_t)j], r1 = a.R[4*(size_t)j + 1], r2 = a.R[4*(size_t)j + 2];
      // the sweep's multiply-adds (sweep_slice), so that U equals the fused merit pass's
      u0 = u0 + fma(w[6*SLICE], r2, fma(w[3*SLICE], r1, w[0]*r0));
      u1 = u1 + fma(w[7*SLICE], r2, fma(w[4*SLICE], r1, w[SLICE]*r0));
      u2 = u2 + fma(w[8*SLICE], r2, fma(w[5*SLICE], r1, w[2*SLICE]*r0));
    }
  }
  u0 = gsum<G>(u0); u1 = gsum<G>(u1); u2 = gsum<G>(u2);
  if (act)
  {
    double R[3] = {a.R[4*(size_t)p], a.R[4*(size_t)p + 1], a.R[4*(size_t)p + 2]}, U[3];
    if (a.update_U)
    {
      double Wi[9], Bv[3] = {a.B[p], a.B[(size_t)n + p], a.B[2*(size_t)n + p]};
#pragma unroll
      for (int e = 0; e < 9; e ++) Wi[e] = a.Wd[(size_t)e*n + p];
      mv_add(Bv, Wi, R, U);
      U[0] += u0; U[1] += u1; U[2] += u2;
      if (g == 0)
#pragma unroll
        for (int k = 0; k < 3; k ++) a.U[(size_t)k*n + p] = U[k];
    }
    else
    {
#pragma unroll
      for (int k = 0; k < 3; k ++) U[k] = a.U[(size_t)k*n + p];
    }
    double C[3], Ai[9];
    scf_residual<double>(a.fri[p], a.coh[p], a.shift[p], U, R, C, nullptr, nullptr);
#pragma unroll
    for (int e = 0; e < 9; e ++) Ai[e] = a.A[(size_t)e*n + p];
    const double q0 = Ai[0]*C[0] + Ai[3]*C[1] + Ai[6]*C[2];
    const double q1 = Ai[1]*C[0] + Ai[4]*C[1] + Ai[7]*C[2];
    const double q2 = Ai[2]*C[0] + Ai[5]*C[1] + Ai[8]*C[2];
    if (g == 0)
    {
      const double t = q0*C[0] + q1*C[1] + q2*C[2];
      up += t;
      if (a.conmerit) a.conmerit[p] = t;
    }
  }
}

} // namespace sfgs
