/* or_dbs.c -- per-contact Signorini-Coulomb local solvers (3x3 diagonal block).
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Follows dbs.c:
 *   projected_gradient  dbs.c:35-94
 *   de_saxce_feng       dbs.c:96-140
 *   semismooth_newton   dbs.c:142-303
 *   projected_newton    dbs.c:305-326
 *   dispatcher          dbs.c:492-550 (CONTACT + SIGNORINI_COULOMB only)
 * OR_Solve3 stands for LAPACK dgesv on 3x3 (dbs.c:282, 317): LU with partial
 * pivoting, first maximal |a| in the column, failure on an exactly zero pivot.
 *
 * Deliberate deviation: the reference computes iter % (maxiter/10), which is a
 * division by zero for diagmaxiter < 10 (dbs.c:156, 294); here the rho*10
 * schedule is simply off when maxiter/10 == 0.
 *
 * Parity: pinned in tests/test_oracle_dbs.py -- converged (R, U) satisfy the
 * contact law C(U,R) = 0 (or_scf.c) to diagepsilon; the four solvers agree;
 * closed forms for separation, sticking and sliding on decoupled blocks;
 * OR_Solve3 against numpy.linalg.solve and its pivot order by a permutation
 * test.  Where oracle/_ref (the reference's own dbs.c/scf.c built against the
 * image's LAPACK) is present, exact iterate-level agreement is checked too. */
#include <math.h>
#include "oracle.h"
#include "or_la.h"

int OR_Solve3(double *A, double *b)
{
  for (int k = 0; k < 3; k ++)
  {
    int p = k;
    double amax = fabs(A[k + 3*k]);
    for (int i = k + 1; i < 3; i ++)
    {
      double v = fabs(A[i + 3*k]);
      if (v > amax) { amax = v; p = i; }
    }
    if (A[p + 3*k] == 0.0) return k + 1;  /* singular: dgesv info > 0 */
    if (p != k)
    {
      for (int c = 0; c < 3; c ++) { double t = A[k + 3*c]; A[k + 3*c] = A[p + 3*c]; A[p + 3*c] = t; }
      double t = b[k]; b[k] = b[p]; b[p] = t;
    }
    double inv = 1.0 / A[k + 3*k];
    for (int i = k + 1; i < 3; i ++)
    {
      double l = A[i + 3*k] * inv;
      A[i + 3*k] = l;
      for (int c = k + 1; c < 3; c ++) A[i + 3*c] -= l * A[k + 3*c];
      b[i] -= l * b[k];
    }
  }
  for (int k = 2; k >= 0; k --)
  {
    b[k] /= A[k + 3*k];
    for (int i = 0; i < k; i ++) b[i] -= b[k] * A[i + 3*k];
  }
  return 0;
}

static double rel_change(const double *R, const double *R0)
{
  double d[3] = {R[0] - R0[0], R[1] - R0[1], R[2] - R0[2]};
  double rr = or_dot(R, R);
  return sqrt(or_dot(d, d) / (rr > 1.0 ? rr : 1.0));
}

static double dashed_normal(int dynamic, double step, double restitution, double gap, const double *U, const double *V)
{
  if (dynamic) return U[2] + restitution * (V[2] < 0.0 ? V[2] : 0.0);
  return (gap > 0.0 ? gap : 0.0)/step + U[2];
}

/* dbs.c:35-94 */
static int projected_gradient(int dynamic, double eps, int maxiter, double step, double fri, double res,
                              double coh, double gap, double rho, const double *W, const double *B,
                              const double *V, double *U, double *R)
{
  if (dynamic && gap > 0.0) { R[0] = R[1] = R[2] = 0.0; U[0] = B[0]; U[1] = B[1]; U[2] = B[2]; return 0; }
  int iter = 0;
  double err;
  do
  {
    double R0[3] = {R[0], R[1], R[2]};
    or_mv_add(B, W, R, U);
    double UN = dashed_normal(dynamic, step, res, gap, U, V);
    R[0] -= rho*U[0];
    R[1] -= rho*U[1];
    R[2] -= rho*UN - coh;
    if (R[2] < 0.0) R[2] = 0.0;
    double t = sqrt(R[0]*R[0] + R[1]*R[1]);
    if (t >= fri*R[2])
    {
      if (t > 0.0) t = fri*R[2]/t;
      R[0] *= t; R[1] *= t;
    }
    R[2] -= coh;
    err = rel_change(R, R0);
  }
  while (++ iter < maxiter && err > eps);
  return iter;
}

/* dbs.c:96-140 */
static int de_saxce_feng(int dynamic, double eps, int maxiter, double step, double fri, double res,
                         double coh, double gap, double rho, const double *W, const double *B,
                         const double *V, double *U, double *R)
{
  if (dynamic && gap > 0.0) { R[0] = R[1] = R[2] = 0.0; U[0] = B[0]; U[1] = B[1]; U[2] = B[2]; return 0; }
  int iter = 0;
  double err;
  do
  {
    double R0[3] = {R[0], R[1], R[2]}, tau[3];
    or_mv_add(B, W, R, U);
    double UN = dashed_normal(dynamic, step, res, gap, U, V);
    tau[0] = R[0] - rho*U[0];
    tau[1] = R[1] - rho*U[1];
    tau[2] = R[2] - rho*(UN + fri*sqrt(or_dot2(U, U)));
    OR_SCF_Project(fri, coh, tau, R);
    err = rel_change(R, R0);
  }
  while (++ iter < maxiter && err > eps);
  return iter;
}

/* dbs.c:142-303 */
static int semismooth_newton(int dynamic, double eps, int maxiter, double step, double fri, double res,
                             double coh, double gap, double rho, const double *W, const double *B,
                             const double *V, double *U, double *R)
{
  if (dynamic && gap > 0.0) { R[0] = R[1] = R[2] = 0.0; U[0] = B[0]; U[1] = B[1]; U[2] = B[2]; return 0; }

  int divi = maxiter / 10, iter = 0;
  double err;
  do
  {
    double R0[3] = {R[0], R[1], R[2]}, RES[3], a[9], b[3], c[3], d[3];
    double UN = dashed_normal(dynamic, step, res, gap, U, V);

    d[0] = R[0] - rho*U[0];
    d[1] = R[1] - rho*U[1];
    d[2] = (R[2] + coh) - rho*UN;

    or_mv_add(B, W, R, RES);           /* RES = B + W R - U */
    RES[0] -= U[0]; RES[1] -= U[1]; RES[2] -= U[2];

    if (d[2] >= 0.0)
    {
      double norm = sqrt(d[0]*d[0] + d[1]*d[1]);
      double lim = fri * (d[2] > 0.0 ? d[2] : 0.0);

      if (norm >= lim)   /* sliding */
      {
        if (lim > 0.0)
        {
          double len = sqrt(R[0]*R[0] + R[1]*R[1]);
          double den = (lim > len ? lim : len) * norm;
          double e = lim / norm, beta, F[4], M[4], H[4];
          if (len == 0.0) beta = 1.0;
          else
          {
            double alfa = (R[0]*d[0] + R[1]*d[1]) / (len*norm);
            double delta = len/lim < 1.0 ? len/lim : 1.0;
            beta = alfa < 0.0 ? 1.0/(1.0 - alfa*delta) : 1.0;
          }
          F[0] = (R[0]*d[0])/den;
          F[1] = (R[1]*d[0])/den;
          F[2] = (R[0]*d[1])/den;
          F[3] = (R[1]*d[1])/den;
          M[0] = e*(1.0 - F[0]);
          M[1] = -e*F[1];
          M[2] = -e*F[2];
          M[3] = e*(1.0 - F[3]);
          H[0] = 1.0 - beta*M[0];
          H[1] = -beta*M[1];
          H[2] = -beta*M[2];
          H[3] = 1.0 - beta*M[3];
          a[0] = H[0] + rho*(M[0]*W[0] + M[2]*W[1]);
          a[1] = H[1] + rho*(M[1]*W[0] + M[3]*W[1]);
          a[2] = W[2];
          a[3] = H[2] + rho*(M[0]*W[3] + M[2]*W[4]);
          a[4] = H[3] + rho*(M[1]*W[3] + M[3]*W[4]);
          a[5] = W[5];
          a[6] = rho*(M[0]*W[6] + M[2]*W[7]) - fri*(d[0]/norm);
          a[7] = rho*(M[1]*W[6] + M[3]*W[7]) - fri*(d[1]/norm);
          a[8] = W[8];
          b[0] = fri*(d[0]/norm)*(R[2] + coh) - R[0] - rho*(M[0]*RES[0] + M[2]*RES[1]);
          b[1] = fri*(d[1]/norm)*(R[2] + coh) - R[1] - rho*(M[1]*RES[0] + M[3]*RES[1]);
          b[2] = -UN - RES[2];
        }
        else           /* degenerate: homogeneous tangential tractions */
        {
          a[0] = 1.0; a[1] = 0.0; a[2] = W[2];
          a[3] = 0.0; a[4] = 1.0; a[5] = W[5];
          a[6] = 0.0; a[7] = 0.0; a[8] = W[8];
          b[0] = -R[0] - RES[0];
          b[1] = -R[1] - RES[1];
          b[2] = -UN - RES[2];
        }
      }
      else               /* sticking */
      {
        for (int k = 0; k < 9; k ++) a[k] = W[k];
        a[6] += U[0]/d[2];
        a[7] += U[1]/d[2];
        double s = 1.0 + rho*U[2]/d[2];
        b[0] = -s*U[0] - RES[0];
        b[1] = -s*U[1] - RES[1];
        b[2] = -UN - RES[2];
      }
    }
    else                 /* separation */
    {
      a[0] = 1.0; a[1] = 0.0; a[2] = 0.0;
      a[3] = 0.0; a[4] = 1.0; a[5] = 0.0;
      a[6] = 0.0; a[7] = 0.0; a[8] = 1.0;
      b[0] = -R[0]; b[1] = -R[1]; b[2] = -R[2];
    }

    if (OR_Solve3(a, b)) return -1;
    if (!isfinite(b[0] + b[1] + b[2])) return -1;

    or_mv_add(RES, W, b, c);
    R[0] += b[0]; R[1] += b[1]; R[2] += b[2];
    U[0] += c[0]; U[1] += c[1]; U[2] += c[2];

    err = rel_change(R, R0);
    iter ++;

    if (divi > 0 && iter % divi == 0)
    {
      rho *= 10.0;
      if (isinf(rho)) return -1;
    }
  }
  while (iter < maxiter && err > eps);

  return iter;
}

/* dbs.c:305-326 */
static int projected_newton(int dynamic, double eps, int maxiter, double step, double fri, double res,
                            double coh, double gap, const double *W, const double *B,
                            const double *V, double *U, double *R)
{
  double T[9], X[9], Y[9], C[3], merit;
  int iter = 0;

  or_mv_add(B, W, R, U);
  OR_SCF_Linearize(dynamic, step, fri, res, coh, gap, V, U, R, -1.0, 0.0, C, X, Y);
  do
  {
    or_mm(X, W, T);
    for (int k = 0; k < 9; k ++) T[k] += Y[k];
    if (OR_Solve3(T, C)) return -1;
    R[0] -= C[0]; R[1] -= C[1]; R[2] -= C[2];
    OR_SCF_Project(fri, coh, R, R);
    or_mv_add(B, W, R, U);
    OR_SCF_Linearize(dynamic, step, fri, res, coh, gap, V, U, R, -1.0, 0.0, C, X, Y);
    merit = sqrt(or_dot(C, C));
  }
  while (merit > eps && ++ iter < maxiter);

  return iter;
}

int OR_Diag_Solve(int solver, double diagepsilon, int diagmaxiter, int dynamic, double step,
                  double friction, double restitution, double cohesion, double gap, double rho,
                  const double *W, const double *B, const double *V, double *U, double *R)
{
  switch (solver)
  {
  case OR_DS_PROJECTED_GRADIENT:
    return projected_gradient(dynamic, diagepsilon, diagmaxiter, step, friction, restitution, cohesion, gap, rho, W, B, V, U, R);
  case OR_DS_DE_SAXCE_FENG:
    return de_saxce_feng(dynamic, diagepsilon, diagmaxiter, step, friction, restitution, cohesion, gap, rho, W, B, V, U, R);
  case OR_DS_SEMISMOOTH_NEWTON:
    return semismooth_newton(dynamic, diagepsilon, diagmaxiter, step, friction, restitution, cohesion, gap, rho, W, B, V, U, R);
  case OR_DS_PROJECTED_NEWTON:
    return projected_newton(dynamic, diagepsilon, diagmaxiter, step, friction, restitution, cohesion, gap, W, B, V, U, R);
  }
  return 0;
}
