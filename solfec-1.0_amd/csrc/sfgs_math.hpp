// sfgs_math.hpp -- device-side 3x3 algebra, cone operations and the local
// Signorini-Coulomb solvers of one contact row, templated on the arithmetic
// type (double for the FP64 path, float for the mixed-precision sweeps).
//
// Semantics follow the reference:
//   scf_project<T>        scf.c:136-174  (SCF_Project)
//   scf_residual<T>       scf.c:28-133   (SCF_Linearize, C and optionally X, Y)
//   solve3<T>             dgesv on 3x3 as called at dbs.c:282, 317
//   projected_gradient<T> dbs.c:35-94
//   de_saxce_feng<T>      dbs.c:96-140
//   semismooth_newton<T>  dbs.c:142-303
//   projected_newton<T>   dbs.c:305-326
// Matrices are column-major (M[r + 3c]), frames [t1 t2 n].  Everything lives in
// registers; there is no memory traffic here.
#pragma once
#include <hip/hip_runtime.h>

namespace sfgs {

template <typename T> struct V3 { T x, y, z; };

template <typename T> __device__ __forceinline__ T tmax(T a, T b) { return a > b ? a : b; }
template <typename T> __device__ __forceinline__ T tmin(T a, T b) { return a < b ? a : b; }

// Division and square root of the local solvers: IEEE in FP64; in the FP32
// sweeps of the mixed mode the hardware reciprocal and square root (1 ulp),
// whose error stays below the FP32 local tolerance (the FP64 refinement sweeps
// settle the final state)
template <typename T> __device__ __forceinline__ T fdiv(T a, T b) { return a / b; }
template <> __device__ __forceinline__ float fdiv<float>(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }
template <typename T> __device__ __forceinline__ T fsqrt(T x) { return sqrt(x); }
template <> __device__ __forceinline__ float fsqrt<float>(float x) { return __builtin_amdgcn_sqrtf(x); }

// y = c + M v
template <typename T>
__device__ __forceinline__ void mv_add(const T *c, const T *M, const T *v, T *y)
{
  T y0 = c[0] + (M[0]*v[0] + M[3]*v[1] + M[6]*v[2]);
  T y1 = c[1] + (M[1]*v[0] + M[4]*v[1] + M[7]*v[2]);
  T y2 = c[2] + (M[2]*v[0] + M[5]*v[1] + M[8]*v[2]);
  y[0] = y0; y[1] = y1; y[2] = y2;
}

template <typename T>
__device__ __forceinline__ T rel_change(const T *R, const T *R0)
{
  T d0 = R[0] - R0[0], d1 = R[1] - R0[1], d2 = R[2] - R0[2];
  T rr = R[0]*R[0] + R[1]*R[1] + R[2]*R[2];
  return fsqrt(fdiv(d0*d0 + d1*d1 + d2*d2, tmax(rr, T(1))));
}

// LU with partial pivoting on a 3x3 (column-major), b overwritten with x.
// Returns nonzero on an exactly zero pivot (dgesv info > 0).
template <typename T>
__device__ __forceinline__ int solve3(T *a, T *b)
{
#pragma unroll
  for (int k = 0; k < 3; k ++)
  {
    int p = k;
    T amax = fabs(a[k + 3*k]);
#pragma unroll
    for (int i = k + 1; i < 3; i ++)
    {
      T v = fabs(a[i + 3*k]);
      if (v > amax) { amax = v; p = i; }
    }
    if (amax == T(0)) return k + 1;
    if (p != k)
    {
#pragma unroll
      for (int c = 0; c < 3; c ++)
      {
        // select-based swap keeps the arrays in registers
        T ak = a[k + 3*c], ap = a[p + 3*c];
        a[k + 3*c] = ap;
#pragma unroll
        for (int i = k + 1; i < 3; i ++) if (i == p) a[i + 3*c] = ak;
      }
      T bk = b[k], bp = b[p];
      b[k] = bp;
#pragma unroll
      for (int i = k + 1; i < 3; i ++) if (i == p) b[i] = bk;
    }
    T inv = fdiv(T(1), a[k + 3*k]);
    a[k + 3*k] = inv;   // kept for the back substitution
#pragma unroll
    for (int i = k + 1; i < 3; i ++)
    {
      T l = a[i + 3*k] * inv;
#pragma unroll
      for (int c = k + 1; c < 3; c ++) a[i + 3*c] -= l * a[k + 3*c];
      b[i] -= l * b[k];
    }
  }
#pragma unroll
  for (int k = 2; k >= 0; k --)
  {
    b[k] *= a[k + 3*k];
#pragma unroll
    for (int i = 0; i < k; i ++) b[i] -= b[k] * a[i + 3*k];
  }
  return 0;
}

// 3x3 inverse by LU with partial pivoting (dgetrf/dgetri as MX_Inverse calls
// them); 0 on success
template <typename T>
__device__ __forceinline__ int inverse3(const T *W, T *A)
{
#pragma unroll
  for (int c = 0; c < 3; c ++)
  {
    T M[9], x[3] = {c == 0 ? T(1) : T(0), c == 1 ? T(1) : T(0), c == 2 ? T(1) : T(0)};
#pragma unroll
    for (int k = 0; k < 9; k ++) M[k] = W[k];
    if (solve3(M, x)) return -1;
    A[3*c] = x[0]; A[3*c + 1] = x[1]; A[3*c + 2] = x[2];
  }
  return 0;
}

// spectral factorisation of S with respect to the cone (orientation of scf.c)
template <typename T>
__device__ __forceinline__ void spectral(T fri, const T *S, T &l1, T &l2, T *u1, T *u2, T &slen)
{
  T onefri2 = T(1) + fri*fri;
  slen = fsqrt(S[0]*S[0] + S[1]*S[1]);
  l1 = fdiv(-(S[2] + fri*slen), onefri2);
  l2 = fdiv(slen - fri*S[2], onefri2);
  if (slen != T(0)) { u2[0] = S[0]/slen; u2[1] = S[1]/slen; }
  else { u2[0] = T(1); u2[1] = T(0); }
  u2[2] = -fri;
  u1[0] = -fri*u2[0]; u1[1] = -fri*u2[1]; u1[2] = T(-1);
}

// R = projection of S onto the friction cone shifted by the cohesion
template <typename T>
__device__ __forceinline__ void scf_project(T fri, T coh, const T *S, T *R)
{
  T Sc[3] = {S[0], S[1], S[2] + coh}, l1, l2, u1[3], u2[3], sl;
  spectral(fri, Sc, l1, l2, u1, u2, sl);
  T g1 = tmax(l1, T(0)), g2 = tmax(l2, T(0));
  T m0 = g1*u1[0] + g2*u2[0], m1 = g1*u1[1] + g2*u2[1], m2 = g1*u1[2] + g2*u2[2];
  T s0 = S[0], s1 = S[1], s2 = S[2];
  R[0] = s0 - m0; R[1] = s1 - m1; R[2] = s2 - m2;
}

// C(U,R) with UT = -1 and zero smoothing (the merit form, mrf.c:82) and, when
// X != nullptr, the Jacobians X = dC/dU, Y = dC/dR (scf.c:92-132).
// 'shift' is the dashed-normal offset: dynamic e*min(V_N,0), static max(gap,0)/h.
template <typename T>
__device__ __forceinline__ void scf_residual(T fri, T coh, T shift, const T *U, const T *R, T *C, T *X, T *Y)
{
  T fri2 = fri*fri, onefri2 = T(1) + fri2;
  T udash = U[2] + shift;
  T ulen = sqrt(U[0]*U[0] + U[1]*U[1]);
  T F[3] = {U[0], U[1], udash + fri*ulen};
  T S[3] = {R[0] - F[0], R[1] - F[1], R[2] - F[2] + coh};
  T l1, l2, u1[3], u2[3], slen;
  spectral(fri, S, l1, l2, u1, u2, slen);
  T sq1 = fabs(l1), sq2 = fabs(l2);
  T g1 = T(0.5)*(sq1 + l1), g2 = T(0.5)*(sq2 + l2);
  C[0] = F[0] + (g1*u1[0] + g2*u2[0]);
  C[1] = F[1] + (g1*u1[1] + g2*u2[1]);
  C[2] = F[2] + (g1*u1[2] + g2*u2[2]);
  if (X)
  {
    T dF[9] = {T(1), T(0), fri*U[0]/ulen, T(0), T(1), fri*U[1]/ulen, T(0), T(0), T(1)};
    T dg1 = T(0.5)*(T(1) + l1/sq1), sdot = S[0]*S[0] + S[1]*S[1];
    if (slen != T(0))
    {
      T dg2 = T(0.5)*(T(1) + l2/sq2);
      T a = T(0.5)*(T(1) + (l2 + fri*l1)/(sq2 + fri*sq1));
      T b = (fri2*dg1 + dg2)/onefri2;
      T c = (fri*(dg1 - dg2))/onefri2;
      T d = (dg1 + fri2*dg2)/onefri2;
      Y[0] = a + (b - a)*S[0]*S[0]/sdot;
      Y[1] = (b - a)*S[1]*S[0]/sdot;
      Y[2] = c*S[0]/slen;
      Y[3] = Y[1];
      Y[4] = a + (b - a)*S[1]*S[1]/sdot;
      Y[5] = c*S[1]/slen;
      Y[6] = Y[2];
      Y[7] = Y[5];
      Y[8] = d;
    }
    else
    {
      Y[1] = Y[2] = Y[3] = Y[5] = Y[6] = Y[7] = T(0);
      Y[0] = Y[4] = Y[8] = dg1;
    }
#pragma unroll
    for (int c = 0; c < 3; c ++)
#pragma unroll
      for (int r = 0; r < 3; r ++)
        X[r + 3*c] = dF[r + 3*c] - (Y[r]*dF[3*c] + Y[r + 3]*dF[3*c + 1] + Y[r + 6]*dF[3*c + 2]);
  }
}

// per-row constants of the local problem
template <typename T>
struct RowLaw {
  T fri, coh, shift, rho;
  bool separated;   // dynamic && gap > 0: R = 0, U = B
};

template <typename T>
__device__ __forceinline__ int projected_gradient(const RowLaw<T> &L, T eps, int maxiter, const T *W, const T *B, T *U, T *R)
{
  if (L.separated) { R[0] = R[1] = R[2] = T(0); U[0] = B[0]; U[1] = B[1]; U[2] = B[2]; return 0; }
  int iter = 0;
  T err;
  do
  {
    T R0[3] = {R[0], R[1], R[2]};
    mv_add(B, W, R, U);
    T UN = U[2] + L.shift;
    R[0] -= L.rho*U[0];
    R[1] -= L.rho*U[1];
    R[2] -= L.rho*UN - L.coh;
    R[2] = tmax(T(0), R[2]);
    T t = sqrt(R[0]*R[0] + R[1]*R[1]);
    if (t >= L.fri*R[2])
    {
      if (t > T(0)) t = L.fri*R[2]/t;
      R[0] *= t; R[1] *= t;
    }
    R[2] -= L.coh;
    err = rel_change(R, R0);
  }
  while (++ iter < maxiter && err > eps);
  return iter;
}

template <typename T>
__device__ __forceinline__ int de_saxce_feng(const RowLaw<T> &L, T eps, int maxiter, const T *W, const T *B, T *U, T *R)
{
  if (L.separated) { R[0] = R[1] = R[2] = T(0); U[0] = B[0]; U[1] = B[1]; U[2] = B[2]; return 0; }
  int iter = 0;
  T err;
  do
  {
    T R0[3] = {R[0], R[1], R[2]}, tau[3];
    mv_add(B, W, R, U);
    T UN = U[2] + L.shift;
    tau[0] = R[0] - L.rho*U[0];
    tau[1] = R[1] - L.rho*U[1];
    tau[2] = R[2] - L.rho*(UN + L.fri*sqrt(U[0]*U[0] + U[1]*U[1]));
    scf_project(L.fri, L.coh, tau, R);
    err = rel_change(R, R0);
  }
  while (++ iter < maxiter && err > eps);
  return iter;
}

template <typename T>
__device__ __forceinline__ int semismooth_newton(const RowLaw<T> &L, T eps, int maxiter, const T *W, const T *B, T *U, T *R)
{
  if (L.separated) { R[0] = R[1] = R[2] = T(0); U[0] = B[0]; U[1] = B[1]; U[2] = B[2]; return 0; }
  const T fri = L.fri, coh = L.coh;
  T rho = L.rho;
  int divi = maxiter / 10, iter = 0;
  T err;
  do
  {
    T R0[3] = {R[0], R[1], R[2]}, RES[3], a[9], b[3], c[3], d[3];
    T UN = U[2] + L.shift;
    d[0] = R[0] - rho*U[0];
    d[1] = R[1] - rho*U[1];
    d[2] = (R[2] + coh) - rho*UN;
    mv_add(B, W, R, RES);
    RES[0] -= U[0]; RES[1] -= U[1]; RES[2] -= U[2];

    if (d[2] >= T(0))
    {
      T norm = fsqrt(d[0]*d[0] + d[1]*d[1]);
      T lim = fri * tmax(T(0), d[2]);
      if (norm >= lim)
      {
        if (lim > T(0))
        {
          T len = fsqrt(R[0]*R[0] + R[1]*R[1]);
          const T inorm = fdiv(T(1), norm);
          T den = tmax(lim, len) * norm;
          const T iden = fdiv(T(1), den);
          T e = lim * inorm, beta;
          if (len == T(0)) beta = T(1);
          else
          {
            T alfa = fdiv(R[0]*d[0] + R[1]*d[1], len*norm);
            T delta = tmin(fdiv(len, lim), T(1));
            beta = alfa < T(0) ? fdiv(T(1), T(1) - alfa*delta) : T(1);
          }
          T F0 = (R[0]*d[0])*iden, F1 = (R[1]*d[0])*iden, F2 = (R[0]*d[1])*iden, F3 = (R[1]*d[1])*iden;
          T M0 = e*(T(1) - F0), M1 = -e*F1, M2 = -e*F2, M3 = e*(T(1) - F3);
          T H0 = T(1) - beta*M0, H1 = -beta*M1, H2 = -beta*M2, H3 = T(1) - beta*M3;
          const T t0 = d[0]*inorm, t1 = d[1]*inorm;
          a[0] = H0 + rho*(M0*W[0] + M2*W[1]);
          a[1] = H1 + rho*(M1*W[0] + M3*W[1]);
          a[2] = W[2];
          a[3] = H2 + rho*(M0*W[3] + M2*W[4]);
          a[4] = H3 + rho*(M1*W[3] + M3*W[4]);
          a[5] = W[5];
          a[6] = rho*(M0*W[6] + M2*W[7]) - fri*t0;
          a[7] = rho*(M1*W[6] + M3*W[7]) - fri*t1;
          a[8] = W[8];
          b[0] = fri*t0*(R[2] + coh) - R[0] - rho*(M0*RES[0] + M2*RES[1]);
          b[1] = fri*t1*(R[2] + coh) - R[1] - rho*(M1*RES[0] + M3*RES[1]);
          b[2] = -UN - RES[2];
        }
        else
        {
          a[0] = T(1); a[1] = T(0); a[2] = W[2];
          a[3] = T(0); a[4] = T(1); a[5] = W[5];
          a[6] = T(0); a[7] = T(0); a[8] = W[8];
          b[0] = -R[0] - RES[0];
          b[1] = -R[1] - RES[1];
          b[2] = -UN - RES[2];
        }
      }
      else
      {
#pragma unroll
        for (int k = 0; k < 9; k ++) a[k] = W[k];
        const T id2 = fdiv(T(1), d[2]);
        a[6] += U[0]*id2;
        a[7] += U[1]*id2;
        T s = T(1) + rho*U[2]*id2;
        b[0] = -s*U[0] - RES[0];
        b[1] = -s*U[1] - RES[1];
        b[2] = -UN - RES[2];
      }
    }
    else
    {
      a[0] = T(1); a[1] = T(0); a[2] = T(0);
      a[3] = T(0); a[4] = T(1); a[5] = T(0);
      a[6] = T(0); a[7] = T(0); a[8] = T(1);
      b[0] = -R[0]; b[1] = -R[1]; b[2] = -R[2];
    }

    if (solve3(a, b)) return -1;
    if (!isfinite(b[0] + b[1] + b[2])) return -1;

    mv_add(RES, W, b, c);
    R[0] += b[0]; R[1] += b[1]; R[2] += b[2];
    U[0] += c[0]; U[1] += c[1]; U[2] += c[2];
    err = rel_change(R, R0);
    iter ++;
    if (divi > 0 && iter % divi == 0)
    {
      rho *= T(10);
      if (isinf(rho)) return -1;
    }
  }
  while (iter < maxiter && err > eps);
  return iter;
}

template <typename T>
__device__ __forceinline__ int projected_newton(const RowLaw<T> &L, T eps, int maxiter, const T *W, const T *B, T *U, T *R)
{
  T X[9], Y[9], C[3], Tm[9], merit;
  int iter = 0;
  mv_add(B, W, R, U);
  scf_residual(L.fri, L.coh, L.shift, U, R, C, X, Y);
  do
  {
#pragma unroll
    for (int c = 0; c < 3; c ++)
#pragma unroll
      for (int r = 0; r < 3; r ++)
        Tm[r + 3*c] = (X[r]*W[3*c] + X[r + 3]*W[3*c + 1] + X[r + 6]*W[3*c + 2]) + Y[r + 3*c];
    if (solve3(Tm, C)) return -1;
    R[0] -= C[0]; R[1] -= C[1]; R[2] -= C[2];
    scf_project(L.fri, L.coh, R, R);
    mv_add(B, W, R, U);
    scf_residual(L.fri, L.coh, L.shift, U, R, C, X, Y);
    merit = sqrt(C[0]*C[0] + C[1]*C[1] + C[2]*C[2]);
  }
  while (merit > eps && ++ iter < maxiter);
  return iter;
}

template <typename T>
__device__ __forceinline__ int local_solve(int solver, const RowLaw<T> &L, T eps, int maxiter, const T *W, const T *B, T *U, T *R)
{
  switch (solver)
  {
  case 0: return projected_gradient(L, eps, maxiter, W, B, U, R);
  case 1: return de_saxce_feng(L, eps, maxiter, W, B, U, R);
  case 3: return projected_newton(L, eps, maxiter, W, B, U, R);
  default: return semismooth_newton(L, eps, maxiter, W, B, U, R);
  }
}

// SCF_Linearize (scf.c:28-133) with UT < 0 and smoothing omega = eps > 0, as
// the Newton solver calls it (nts.c:1075): C(U, R), X = dC/dU, Y = dC/dR with
// |U_T| -> sqrt(|U_T|^2 + eps^2) and |l| -> sqrt(l^2 + 4 eps^2 (mu^2)).
__device__ __forceinline__ void scf_linearize_smooth(double fri, double coh, double shift, double eps, const double *U,
                                                     const double *R, double *C, double *X, double *Y)
{
  const double eps2 = eps*eps, fri2 = fri*fri, onefri2 = 1.0 + fri2;
  const double udash = U[2] + shift;
  const double ulen = sqrt(U[0]*U[0] + U[1]*U[1] + eps2);
  const double F[3] = {U[0], U[1], udash + fri*(ulen - eps)};
  double S[3] = {R[0] - F[0], R[1] - F[1], R[2] - F[2] + coh};
  const double sdot = S[0]*S[0] + S[1]*S[1], slen = sqrt(sdot);
  const double l1 = -(S[2] + fri*slen)/onefri2, l2 = (slen - fri*S[2])/onefri2;
  double u1[3], u2[3];
  if (slen != 0.0) { u2[0] = S[0]/slen; u2[1] = S[1]/slen; }
  else { u2[0] = 1.0; u2[1] = 0.0; }
  u2[2] = -fri;
  u1[0] = -fri*u2[0]; u1[1] = -fri*u2[1]; u1[2] = -1.0;
  const double sq1 = sqrt(l1*l1 + 4.0*eps2), sq2 = sqrt(l2*l2 + 4.0*eps2*fri2);
  const double g1 = 0.5*(sq1 + l1), g2 = 0.5*(sq2 + l2);
  C[0] = F[0] + (g1*u1[0] + g2*u2[0]);
  C[1] = F[1] + (g1*u1[1] + g2*u2[1]);
  C[2] = F[2] + (g1*u1[2] + g2*u2[2]);
  const double dF[9] = {1.0, 0.0, fri*U[0]/ulen, 0.0, 1.0, fri*U[1]/ulen, 0.0, 0.0, 1.0};
  const double dg1 = 0.5*(1.0 + l1/sq1);
  if (slen != 0.0)
  {
    const double dg2 = 0.5*(1.0 + l2/sq2);
    const double a = 0.5*(1.0 + (l2 + fri*l1)/(sq2 + fri*sq1));
    const double b = (fri2*dg1 + dg2)/onefri2, c = (fri*(dg1 - dg2))/onefri2, d = (dg1 + fri2*dg2)/onefri2;
    Y[0] = a + (b - a)*S[0]*S[0]/sdot;
    Y[1] = (b - a)*S[1]*S[0]/sdot;
    Y[2] = c*S[0]/slen;
    Y[3] = Y[1];
    Y[4] = a + (b - a)*S[1]*S[1]/sdot;
    Y[5] = c*S[1]/slen;
    Y[6] = Y[2];
    Y[7] = Y[5];
    Y[8] = d;
  }
  else
  {
    Y[1] = Y[2] = Y[3] = Y[5] = Y[6] = Y[7] = 0.0;
    Y[0] = Y[4] = Y[8] = dg1;
  }
  // X = dF - Y dF  ([I - dm/dS] dF/dU)
#pragma unroll
  for (int r = 0; r < 3; r ++)
#pragma unroll
    for (int c = 0; c < 3; c ++)
      X[r + 3*c] = dF[r + 3*c] - (Y[r]*dF[3*c] + Y[r + 3]*dF[3*c + 1] + Y[r + 6]*dF[3*c + 2]);
}

} // namespace sfgs
