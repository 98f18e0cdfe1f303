/* or_nts.c -- the NEWTON_SOLVER with assembled local dynamics (LOCDYN_ON).
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * OR_Newton_Solve follows NEWTON_Solve (nts.c:1333-1379) for contact rows:
 *   - per row the smoothed linearisation of the Signorini-Coulomb residual,
 *     b = -C(U, R), X, Y = SCF_Linearize (scf.c:28-133 with UT < 0 and the
 *     smoothing omega), T = X W_ii + Y (nts.c:1072-1081);
 *   - linver GMRES (gmres_based_solve, nts.c:1184-1244): P = (T + delta I)^-1
 *     (nts.c:1095-1103), the flexible GMRES of ext/krylov/flexgmres.c
 *     (hypre_FlexGMRESSolve: k_dim 20, min_iter 1, absolute tolerance
 *     epsilon |b|, at most linmaxiter iterations, right preconditioning) on
 *     the operator y = alpha (X (W x) + Y x) + beta y + alpha delta x (Matvec,
 *     nts.c:584-652) from the previous increment dr, then R = proj(R + dr)
 *     (nts.c:1139-1154);
 *   - linver DIAG (diagonalized_solve, nts.c:1246-1303): per row T d = b,
 *     dr = (1 - theta) dr + theta d, R = proj(R + dr) (nts.c:1084-1093);
 *   - after each step U = W R + B and the merit (mrf.c); stop at merit <=
 *     meritval, maxiter steps or (GMRES) maxmatvec products; a NaN merit ends
 *     the solve unconverged;
 *   - unconverged with gsflag: R back to its initial value and a
 *     GAUSS_SEIDEL_SOLVER(1.0, maxiter, meritval, CONTINUE, 1e-9, 100,
 *     SEMISMOOTH_NEWTON) solve (nts.c:1358-1372).
 * Vectors are in row order 0..n-1; dot products are sequential sums.
 *
 * Parity: pinned in tests/test_oracle_nts.py.  OR_FlexGMRES_Dense (the same
 * Krylov routine on a dense operator) against numpy.linalg.solve on random
 * nonsymmetric systems (exact after n iterations for n <= k_dim, with and
 * without a preconditioner, restarted and with a zero right-hand side); the
 * Newton solve by the contact law at its converged state (C(U, R) = 0 with
 * U = W R + B), by agreement with the Gauss-Seidel oracle on systems with a
 * unique solution, by decoupled rows (R_N = m g for a sphere on a plane), by
 * the DIAG step in closed form on one row, and by the fallback path. */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <float.h>
#include "oracle.h"

#define NT_KDIM 20   /* hypre_FlexGMRESCreate default k_dim (flexgmres.c:91) */

typedef struct nt_op nt_op;
struct nt_op {
  int N;                                   /* vector length */
  void (*matvec)(nt_op *op, double alpha, const double *x, double beta, double *y);
  void (*precond)(nt_op *op, const double *b, double *x);
  int matvecs;
  /* Newton operator */
  const OR_SYSTEM *sys;
  const double *X, *Y, *P;
  double delta;
  double *tmp;
  /* dense operator */
  const double *A, *Pd;
};

static double dot(int N, const double *a, const double *b)
{
  double s = 0.0;
  for (int i = 0; i < N; i ++) s += a[i]*b[i];
  return s;
}

static void axpy(int N, double alpha, const double *x, double *y)
{
  for (int i = 0; i < N; i ++) y[i] += alpha*x[i];
}

/* hypre_FlexGMRESSolve (flexgmres.c:260-700) with rel_change off, cf_tol 0,
 * r_tol 0: x updated in place; returns the iterations done, or -1 for a
 * non-finite b or initial residual (x untouched) */
static int flexgmres(nt_op *op, const double *b, double *x, int k_dim, int max_iter, int min_iter, double a_tol,
                     double *rnorm_out)
{
  const int N = op->N;
  double *p = calloc((size_t)(k_dim + 1)*(size_t)(N > 0 ? N : 1), sizeof(double));
  double *pre = calloc((size_t)k_dim*(size_t)(N > 0 ? N : 1), sizeof(double));
  double *r = calloc((size_t)(N > 0 ? N : 1), sizeof(double)), *w = calloc((size_t)(N > 0 ? N : 1), sizeof(double));
  double *rs = calloc((size_t)k_dim + 1, sizeof(double)), *c = calloc((size_t)k_dim, sizeof(double));
  double *s = calloc((size_t)k_dim, sizeof(double)), *hh = calloc((size_t)(k_dim + 1)*(size_t)k_dim, sizeof(double));
#define P_(k) (p + (size_t)(k)*N)
#define PRE_(k) (pre + (size_t)(k)*N)
#define HH(i, j) hh[(size_t)(i)*k_dim + (j)]
  const double epsmac = 1.e-16;
  int iter = 0, i = 0, ret = 0;
  double r_norm, b_norm, t, gamma, epsilon;

  memcpy(P_(0), b, sizeof(double)*(size_t)N);
  op->matvec(op, -1.0, x, 1.0, P_(0));
  b_norm = sqrt(dot(N, b, b));
  if (b_norm != 0.0 && b_norm/b_norm != b_norm/b_norm) { ret = -1; goto done; }
  r_norm = sqrt(dot(N, P_(0), P_(0)));
  if (r_norm != 0.0 && r_norm/r_norm != r_norm/r_norm) { ret = -1; goto done; }
  epsilon = a_tol;

  while (iter < max_iter)
  {
    rs[0] = r_norm;
    if (r_norm == 0.0) break;
    if (r_norm <= epsilon && iter >= min_iter)
    {
      memcpy(r, b, sizeof(double)*(size_t)N);
      op->matvec(op, -1.0, x, 1.0, r);
      r_norm = sqrt(dot(N, r, r));
      if (r_norm <= epsilon) break;
    }
    t = 1.0/r_norm;
    for (int q = 0; q < N; q ++) P_(0)[q] *= t;
    i = 0;
    while (i < k_dim && iter < max_iter)
    {
      i ++;
      iter ++;
      memset(PRE_(i - 1), 0, sizeof(double)*(size_t)N);
      op->precond(op, P_(i - 1), PRE_(i - 1));
      op->matvec(op, 1.0, PRE_(i - 1), 0.0, P_(i));
      for (int j = 0; j < i; j ++)   /* modified Gram-Schmidt */
      {
        HH(j, i - 1) = dot(N, P_(j), P_(i));
        axpy(N, -HH(j, i - 1), P_(j), P_(i));
      }
      t = sqrt(dot(N, P_(i), P_(i)));
      HH(i, i - 1) = t;
      if (t != 0.0)
      {
        t = 1.0/t;
        for (int q = 0; q < N; q ++) P_(i)[q] *= t;
      }
      for (int j = 1; j < i; j ++)   /* previous Givens rotations */
      {
        t = HH(j - 1, i - 1);
        HH(j - 1, i - 1) = s[j - 1]*HH(j, i - 1) + c[j - 1]*t;
        HH(j, i - 1) = -s[j - 1]*t + c[j - 1]*HH(j, i - 1);
      }
      t = HH(i, i - 1)*HH(i, i - 1);
      t += HH(i - 1, i - 1)*HH(i - 1, i - 1);
      gamma = sqrt(t);
      if (gamma == 0.0) gamma = epsmac;
      c[i - 1] = HH(i - 1, i - 1)/gamma;
      s[i - 1] = HH(i, i - 1)/gamma;
      rs[i] = -HH(i, i - 1)*rs[i - 1];
      rs[i] /= gamma;
      rs[i - 1] = c[i - 1]*rs[i - 1];
      HH(i - 1, i - 1) = s[i - 1]*HH(i, i - 1) + c[i - 1]*HH(i - 1, i - 1);
      r_norm = fabs(rs[i]);
      if (r_norm <= epsilon && iter >= min_iter) break;
    }
    /* upper triangular solve, correction from the preconditioned vectors */
    rs[i - 1] = rs[i - 1]/HH(i - 1, i - 1);
    for (int k = i - 2; k >= 0; k --)
    {
      t = 0.0;
      for (int j = k + 1; j < i; j ++) t -= HH(k, j)*rs[j];
      t += rs[k];
      rs[k] = t/HH(k, k);
    }
    memcpy(w, PRE_(i - 1), sizeof(double)*(size_t)N);
    for (int q = 0; q < N; q ++) w[q] *= rs[i - 1];
    for (int j = i - 2; j >= 0; j --) axpy(N, rs[j], PRE_(j), w);
    axpy(N, 1.0, w, x);
    if (r_norm <= epsilon && iter >= min_iter)
    {
      memcpy(r, b, sizeof(double)*(size_t)N);
      op->matvec(op, -1.0, x, 1.0, r);
      r_norm = sqrt(dot(N, r, r));
      if (r_norm <= epsilon) break;
      memcpy(P_(0), r, sizeof(double)*(size_t)N);   /* false convergence */
      i = 0;
    }
    for (int j = i; j > 0; j --)   /* residual vector for the restart */
    {
      rs[j - 1] = -s[j - 1]*rs[j];
      rs[j] = c[j - 1]*rs[j];
    }
    if (i) axpy(N, rs[i] - 1.0, P_(i), P_(i));
    for (int j = i - 1; j > 0; j --) axpy(N, rs[j], P_(j), P_(i));
    if (i)
    {
      axpy(N, rs[0] - 1.0, P_(0), P_(0));
      axpy(N, 1.0, P_(i), P_(0));
    }
  }
  ret = iter;
  if (rnorm_out) *rnorm_out = r_norm;
done:
  free(p); free(pre); free(r); free(w); free(rs); free(c); free(s); free(hh);
#undef P_
#undef PRE_
#undef HH
  return ret;
}

/* ---- dense operator (for pinning the Krylov routine) ---- */

static void dense_matvec(nt_op *op, double alpha, const double *x, double beta, double *y)
{
  const int n = op->N;
  for (int i = 0; i < n; i ++)
  {
    double s = 0.0;
    for (int j = 0; j < n; j ++) s += op->A[(size_t)i*n + j]*x[j];
    y[i] = beta*y[i] + alpha*s;
  }
  op->matvecs ++;
}

static void dense_precond(nt_op *op, const double *b, double *x)
{
  const int n = op->N;
  if (!op->Pd) { memcpy(x, b, sizeof(double)*(size_t)n); return; }
  for (int i = 0; i < n; i ++)
  {
    double s = 0.0;
    for (int j = 0; j < n; j ++) s += op->Pd[(size_t)i*n + j]*b[j];
    x[i] = s;
  }
}

int OR_FlexGMRES_Dense(int n, const double *A, const double *P, const double *b, double *x, int kdim, int maxiter,
                       int miniter, double atol, double *rnorm, int *matvecs)
{
  nt_op op;
  memset(&op, 0, sizeof(op));
  op.N = n; op.A = A; op.Pd = P; op.matvec = dense_matvec; op.precond = dense_precond;
  int it = flexgmres(&op, b, x, kdim, maxiter, miniter, atol, rnorm);
  if (matvecs) *matvecs = op.matvecs;
  return it;
}

/* ---- the Newton operator ---- */

/* y = W x (U_WR_B with zero_B, nts.c:255-441): W_ii x_i, then the off-diagonal blocks */
static void w_times(const OR_SYSTEM *s, const double *x, double *y)
{
  for (int i = 0; i < s->n; i ++)
  {
    const double *Wi = s->Wd + 9*(size_t)i, *xi = x + 3*(size_t)i;
    double u[3];
    for (int r = 0; r < 3; r ++) u[r] = Wi[r]*xi[0] + Wi[r + 3]*xi[1] + Wi[r + 6]*xi[2];
    for (int q = s->rowptr[i]; q < s->rowptr[i + 1]; q ++)
    {
      const double *Wb = s->W + 9*(size_t)q, *xj = x + 3*(size_t)s->col[q];
      for (int r = 0; r < 3; r ++) u[r] += Wb[r]*xj[0] + Wb[r + 3]*xj[1] + Wb[r + 6]*xj[2];
    }
    for (int r = 0; r < 3; r ++) y[3*(size_t)i + r] = u[r];
  }
}

static void newton_matvec(nt_op *op, double alpha, const double *x, double beta, double *y)
{
  const OR_SYSTEM *s = op->sys;
  w_times(s, x, op->tmp);
  for (int i = 0; i < s->n; i ++)
  {
    const double *X = op->X + 9*(size_t)i, *Y = op->Y + 9*(size_t)i, *z = op->tmp + 3*(size_t)i;
    const double *xi = x + 3*(size_t)i;
    double u[3];
    for (int r = 0; r < 3; r ++)
      u[r] = (X[r]*z[0] + X[r + 3]*z[1] + X[r + 6]*z[2]) + (Y[r]*xi[0] + Y[r + 3]*xi[1] + Y[r + 6]*xi[2]);
    for (int r = 0; r < 3; r ++)
    {
      double *q = y + 3*(size_t)i + r;
      *q = beta*(*q);
      *q = *q + alpha*u[r];
    }
  }
  axpy(op->N, alpha*op->delta, x, y);
  op->matvecs ++;
}

static void newton_precond(nt_op *op, const double *b, double *x)
{
  for (int i = 0; i < op->sys->n; i ++)
  {
    const double *P = op->P + 9*(size_t)i, *q = b + 3*(size_t)i;
    for (int r = 0; r < 3; r ++) x[3*(size_t)i + r] = P[r]*q[0] + P[r + 3]*q[1] + P[r + 6]*q[2];
  }
}

static void uwrb(const OR_SYSTEM *s, const double *R, double *U)
{
  w_times(s, R, U);
  for (int k = 0; k < 3*s->n; k ++) U[k] += s->B[k];
}

int OR_Newton_Solve(const OR_SYSTEM *sys, const OR_NTPARAM *par, const int *order, double *R, double *U,
                    double *merhist, int *mvhist, OR_NTREPORT *rep)
{
  const int n = sys->n, N = 3*n;
  const size_t n1 = (size_t)(n > 0 ? n : 1);
  double *X = calloc(9*n1, sizeof(double)), *Y = calloc(9*n1, sizeof(double)), *P = calloc(9*n1, sizeof(double));
  double *dr = calloc(3*n1, sizeof(double)), *b = calloc(3*n1, sizeof(double)), *tmp = calloc(3*n1, sizeof(double));
  double *R0 = malloc(sizeof(double)*3*n1);
  int singular = 0;
  memcpy(R0, R, sizeof(double)*(size_t)N);
  memset(rep, 0, sizeof(*rep));

  double W_norm = 1.0;
  if (par->reldelta == 1 && n > 0)
  {
    double sum = 0.0, div = 0.0;
    for (int i = 0; i < n; i ++) { const double *W = sys->Wd + 9*(size_t)i; sum += W[0] + W[4] + W[8]; div += 3.0; }
    W_norm = sum/div;
  }
  else if (par->reldelta == 2)
  {
    W_norm = DBL_MAX;
    for (int i = 0; i < n; i ++)
      for (int e = 0; e < 9; e += 4) if (sys->Wd[9*(size_t)i + e] < W_norm) W_norm = sys->Wd[9*(size_t)i + e];
  }
  else if (par->reldelta == 3)
  {
    W_norm = 0.0;
    for (int i = 0; i < n; i ++)
      for (int e = 0; e < 9; e += 4) if (sys->Wd[9*(size_t)i + e] > W_norm) W_norm = sys->Wd[9*(size_t)i + e];
  }
  rep->W_norm = W_norm;
  const double delta = par->delta*W_norm, gamma = 1.0 - par->theta;

  nt_op op;
  memset(&op, 0, sizeof(op));
  op.N = N; op.sys = sys; op.X = X; op.Y = Y; op.P = P; op.delta = delta; op.tmp = tmp;
  op.matvec = newton_matvec; op.precond = newton_precond;

  uwrb(sys, R, U);
  double merit = OR_Merit(sys, R, U, 0, NULL);
  int iters = 0, converged = 0, nan = 0;
  while (iters < par->maxiter && (par->linver == 1 || op.matvecs < par->maxmatvec) && merit > par->meritval)
  {
    for (int i = 0; i < n; i ++)   /* linearisation per row */
    {
      double C[3], T[9];
      double *Xi = X + 9*(size_t)i, *Yi = Y + 9*(size_t)i, *bi = b + 3*(size_t)i, *Ri = R + 3*(size_t)i;
      const double *W = sys->Wd + 9*(size_t)i;
      OR_SCF_Linearize(sys->dynamic, sys->step, sys->mu[i], sys->rest[i], sys->coh[i], sys->gap[i],
                       sys->V + 3*(size_t)i, U + 3*(size_t)i, Ri, -1.0, par->omega, C, Xi, Yi);
      for (int r = 0; r < 3; r ++) bi[r] = -C[r];
      for (int r = 0; r < 3; r ++)
        for (int cc = 0; cc < 3; cc ++)
          T[r + 3*cc] = (Xi[r]*W[3*cc] + Xi[r + 3]*W[3*cc + 1] + Xi[r + 6]*W[3*cc + 2]) + Yi[r + 3*cc];
      if (par->linver == 1)
      {
        if (OR_Solve3(T, bi)) singular ++;
        double *d = dr + 3*(size_t)i;
        for (int r = 0; r < 3; r ++) d[r] = gamma*d[r] + par->theta*bi[r];
        for (int r = 0; r < 3; r ++) Ri[r] += d[r];
        OR_SCF_Project(sys->mu[i], sys->coh[i], Ri, Ri);
      }
      else
      {
        T[0] += delta; T[4] += delta; T[8] += delta;
        if (OR_Inverse3(T, P + 9*(size_t)i)) singular ++;
      }
    }
    if (par->linver != 1)
    {
      const double bnorm = sqrt(dot(N, b, b));
      flexgmres(&op, b, dr, NT_KDIM, par->linmaxiter, 1, par->epsilon*bnorm, NULL);
      for (int i = 0; i < n; i ++)
      {
        double *Ri = R + 3*(size_t)i;
        for (int r = 0; r < 3; r ++) Ri[r] += dr[3*(size_t)i + r];
        OR_SCF_Project(sys->mu[i], sys->coh[i], Ri, Ri);
      }
    }
    uwrb(sys, R, U);
    merit = OR_Merit(sys, R, U, 0, NULL);
    if (isnan(merit)) { nan = 1; break; }
    if (merhist) merhist[iters] = merit;
    if (mvhist) mvhist[iters] = par->linver == 1 ? 0 : op.matvecs;
    iters ++;
  }
  converged = !nan && merit <= par->meritval;
  rep->iters = iters;
  rep->matvecs = op.matvecs;
  rep->merit = merit;
  rep->converged = converged;
  rep->singular = singular;
  if (!converged && par->gsflag)
  {
    OR_GSPARAM gp = {1.0, par->maxiter, par->meritval, OR_FAILURE_CONTINUE, 1e-9, 100, OR_DS_SEMISMOOTH_NEWTON, 0};
    OR_REPORT gr;
    double *rer = calloc((size_t)par->maxiter, sizeof(double)), *mer = calloc((size_t)par->maxiter, sizeof(double));
    int *nat = NULL;
    if (!order)
    {
      nat = malloc(sizeof(int)*n1);
      for (int i = 0; i < n; i ++) nat[i] = i;
    }
    memcpy(R, R0, sizeof(double)*(size_t)N);
    OR_GS_Solve(sys, &gp, order ? order : nat, R, U, rer, mer, &gr);
    free(nat);
    rep->gs_fallback = 1;
    rep->gs_iters = gr.iters;
    rep->merit = gr.merit;
    free(rer); free(mer);
  }
  free(X); free(Y); free(P); free(dr); free(b); free(tmp); free(R0);
  return 0;
}
