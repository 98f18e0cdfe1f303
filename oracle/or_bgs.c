/* or_bgs.c -- serial block Gauss-Seidel and the merit function.
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * OR_GS_Solve follows the serial GAUSS_SEIDEL_Solve, bgs.c:913-1069:
 *   merit off when meritval >= 1 (bgs.c:927); forward/backward alternation when
 *   reverse is on (bgs.c:936-950); B' = B_i + sum_j W_ij R_j (bgs.c:957-963);
 *   local solve and the SSN, PG, DSF, PN fallback chain under FAILURE_CONTINUE,
 *   keeping the previous R when all fail (bgs.c:969-1011); errup/errlo
 *   (bgs.c:1016-1018); merit with U update (bgs.c:1022-1025); relative error
 *   (bgs.c:1028); histories (bgs.c:1031-1032); termination (bgs.c:1036);
 *   DIVERGED after maxiter (bgs.c:1053-1068).  FAILURE_EXIT returns 1 where
 *   the reference THROWs (bgs.c:1006, 1062); FAILURE_CALLBACK counts the
 *   callback invocations in rep->diag_fail.
 * The row order is an argument: reference list order (reverse insertion,
 * ldy.c:505-508) or the colour-major order of the GPU schedule.
 *
 * OR_Merit follows MERIT_Function, mrf.c:35-184 (CONTACT kind):
 *   U = B + W_ii R + sum W_ij R_j when update_U (mrf.c:61-76), C from
 *   SCF_Linearize with UT = -1, omega = 0 (mrf.c:82), up = C'AC (mrf.c:83-84),
 *   result 0.5 sum up / E with E = free energy (1 when < DBL_EPSILON, mrf.c:45,
 *   176-184), per-contact merit up/E (mrf.c:167, 181).
 *
 * Parity: pinned in tests/test_oracle_bgs.py -- at convergence every row
 * satisfies its local contact law; merit of an exact solution is 0 and
 * merit(R = 0) has the closed form sum F(B)'A F(B) / sum B'AB; forward and
 * list orders reach the same merit threshold; histories are monotone where the
 * mathematics requires it (single decoupled rows converge in one sweep). */
#include <math.h>
#include <float.h>
#include <string.h>
#include "oracle.h"
#include "or_la.h"

double OR_Merit(const OR_SYSTEM *sys, const double *R, double *U, int update_U, double *conmerit)
{
  double up_sum = 0.0, lo = sys->free_energy < DBL_EPSILON ? 1.0 : sys->free_energy;

  for (int i = 0; i < sys->n; i ++)
  {
    double *Ui = U + 3*i, C[3], Q[3];
    const double *Ri = R + 3*i;
    if (update_U)
    {
      or_mv_add(sys->B + 3*i, sys->Wd + 9*i, Ri, Ui);
      for (int p = sys->rowptr[i]; p < sys->rowptr[i + 1]; p ++)
        or_mv_add(Ui, sys->W + 9*(long)p, R + 3*sys->col[p], Ui);
    }
    OR_SCF_Linearize(sys->dynamic, sys->step, sys->mu[i], sys->rest[i], sys->coh[i], sys->gap[i],
                     sys->V + 3*i, Ui, Ri, -1.0, 0.0, C, NULL, NULL);
    or_mv(sys->A + 9*i, C, Q);
    double up = or_dot(Q, C);
    if (conmerit) conmerit[i] = up;
    up_sum += up;
  }
  up_sum *= 0.5;
  if (lo == 0.0) lo = 1.0;
  if (conmerit) for (int i = 0; i < sys->n; i ++) conmerit[i] /= lo;
  return up_sum / lo;
}

/* one row update; returns 0, or 1 when a FAILURE_EXIT must be thrown */
static int row_update(const OR_SYSTEM *sys, const OR_GSPARAM *par, int i, double *R, double *U,
                      double *errup, double *errlo, int *gserror, OR_REPORT *rep)
{
  double Bp[3], R0[3], *Ri = R + 3*i;
  memcpy(Bp, sys->B + 3*i, sizeof(Bp));
  for (int p = sys->rowptr[i]; p < sys->rowptr[i + 1]; p ++)
    or_mv_add(Bp, sys->W + 9*(long)p, R + 3*sys->col[p], Bp);
  memcpy(R0, Ri, sizeof(R0));

  int d = OR_Diag_Solve(par->diagsolver, par->diagepsilon, par->diagmaxiter, sys->dynamic, sys->step,
                        sys->mu[i], sys->rest[i], sys->coh[i], sys->gap[i], sys->rho[i],
                        sys->Wd + 9*i, Bp, sys->V + 3*i, U + 3*i, Ri);
  if (d > 0) rep->newton_iters += d;

  if (d >= par->diagmaxiter || d < 0)
  {
    *gserror = d < 0 ? OR_GS_DIAGONAL_FAILED : OR_GS_DIAGONAL_DIVERGED;
    rep->diag_fail ++;
    switch (par->failure)
    {
    case OR_FAILURE_CONTINUE:
    {
      static const int chain[4] = {OR_DS_SEMISMOOTH_NEWTON, OR_DS_PROJECTED_GRADIENT,
                                   OR_DS_DE_SAXCE_FENG, OR_DS_PROJECTED_NEWTON};
      for (int k = 0; k < 4; k ++)
      {
        if (chain[k] == par->diagsolver) continue;
        memcpy(Ri, R0, sizeof(R0));
        d = OR_Diag_Solve(chain[k], par->diagepsilon, par->diagmaxiter, sys->dynamic, sys->step,
                          sys->mu[i], sys->rest[i], sys->coh[i], sys->gap[i], sys->rho[i],
                          sys->Wd + 9*i, Bp, sys->V + 3*i, U + 3*i, Ri);
        if (d > 0) rep->newton_iters += d;
        if (d < par->diagmaxiter && d >= 0) break;
      }
      if (d >= par->diagmaxiter || d < 0) memcpy(Ri, R0, sizeof(R0));
    }
    break;
    case OR_FAILURE_EXIT:
      return 1;
    case OR_FAILURE_CALLBACK:
      break;
    }
  }

  double dR[3] = {Ri[0] - R0[0], Ri[1] - R0[1], Ri[2] - R0[2]};
  *errup += or_dot(dR, dR);
  *errlo += or_dot(Ri, Ri);
  rep->row_updates ++;
  return 0;
}

int OR_GS_Solve(const OR_SYSTEM *sys, const OR_GSPARAM *par, const int *order,
                double *R, double *U, double *rerhist, double *merhist, OR_REPORT *rep)
{
  int n = sys->n, nomerit = par->meritval >= 1.0 ? 1 : 0, gserror = OR_GS_OK, iters = 0;
  double merit = 0.0, error = 0.0;

  memset(rep, 0, sizeof(*rep));
  do
  {
    double errup = 0.0, errlo = 0.0;
    int backward = par->reverse && (iters % 2);
    for (int k = 0; k < n; k ++)
    {
      int i = order[backward ? n - 1 - k : k];
      if (row_update(sys, par, i, R, U, &errup, &errlo, &gserror, rep))
      {
        rep->iters = iters; rep->error = gserror; rep->rel_error = error; rep->merit = merit;
        return 1;
      }
    }
    if (!nomerit) merit = OR_Merit(sys, R, U, 1, NULL);
    error = sqrt(errup) / sqrt(errlo == 0.0 ? 1.0 : errlo);
    if (rerhist) rerhist[iters] = error;
    if (merhist) merhist[iters] = merit;
  }
  while (++ iters < par->maxiter && (error > par->epsilon || merit > par->meritval));

  rep->iters = iters;
  rep->rel_error = error;
  rep->merit = merit;
  rep->error = gserror;
  if (iters >= par->maxiter)
  {
    rep->error = OR_GS_DIVERGED;
    if (par->failure == OR_FAILURE_EXIT) return 1;
  }
  return 0;
}

double OR_GS_Sweep(const OR_SYSTEM *sys, const OR_GSPARAM *par, const int *order, int backward,
                   double *R, double *U, double *errlo_out)
{
  OR_REPORT rep;
  double errup = 0.0, errlo = 0.0;
  int gserror = 0, n = sys->n;
  memset(&rep, 0, sizeof(rep));
  OR_GSPARAM p = *par;
  p.failure = OR_FAILURE_CONTINUE;
  for (int k = 0; k < n; k ++) row_update(sys, &p, order[backward ? n - 1 - k : k], R, U, &errup, &errlo, &gserror, &rep);
  if (errlo_out) *errlo_out = errlo;
  return errup;
}
