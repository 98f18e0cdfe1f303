/* bgs_host.c -- the GAUSS_SEIDEL solver object of the host layer.
 *
 * Creation and defaults follow GAUSS_SEIDEL_Create (bgs.c:443-472) and the
 * Python constructor defaults (lng.c:4258-4342): diagepsilon defaults to
 * 0.01 min(epsilon, meritval, 1e-4) clamped at DBL_EPSILON, diagmaxiter to
 * max(100, maxiter/100), reverse OFF, variant FULL, innerloops 1.  The solve
 * itself runs on the GPU (SFGS_Solve); failure actions follow
 * bgs.c:972-1011 and 1053-1068: CONTINUE keeps going, EXIT ends the run with
 * the reference's error, CALLBACK calls the user function once per local
 * failure and once on divergence, and a zero return stops the program. */
#include <float.h>
#include <stdlib.h>
#include <string.h>
#include <stdio.h>
#include "solfec.h"
#include "sol_internal.h"

GAUSS_SEIDEL* GAUSS_SEIDEL_Create(double epsilon, int maxiter, double meritval, int failure,
                                  double diagepsilon, int diagmaxiter, int diagsolver,
                                  void *data, GAUSS_SEIDEL_Callback callback)
{
  if (!(epsilon > 0.0) || maxiter <= 0 || failure < 0 || failure > 2 || diagsolver < 0 || diagsolver > 3) return NULL;
  if (failure == SFGS_FAILURE_CALLBACK && !callback) return NULL;
  GAUSS_SEIDEL *gs = calloc(1, sizeof(GAUSS_SEIDEL));
  if (!gs) return NULL;
  if (!(diagepsilon > 0.0))
  {
    double m = meritval < 1e-4 ? meritval : 1e-4;
    diagepsilon = 0.01 * (epsilon < m ? epsilon : m);
    if (diagepsilon < DBL_EPSILON) diagepsilon = DBL_EPSILON;
  }
  if (diagmaxiter <= 0) diagmaxiter = maxiter/100 > 100 ? maxiter/100 : 100;
  gs->par.epsilon = epsilon;
  gs->par.maxiter = maxiter;
  gs->par.meritval = meritval;
  gs->par.failure = failure;
  gs->par.diagepsilon = diagepsilon;
  gs->par.diagmaxiter = diagmaxiter;
  gs->par.diagsolver = diagsolver;
  gs->par.reverse = 0;
  gs->par.precision = SFGS_PREC_FP64;
  gs->par.refine = 2;
  gs->variant = 0;
  gs->innerloops = 1;
  gs->data = data;
  gs->callback = callback;
  gs->error = SFGS_GS_OK;
  gs->itershistcount = -1;
  return gs;
}

void GAUSS_SEIDEL_Destroy(GAUSS_SEIDEL *gs)
{
  if (!gs) return;
  free(gs->rerhist);
  free(gs->merhist);
  free(gs->itershist);
  free(gs);
}

double GAUSS_SEIDEL_Get(GAUSS_SEIDEL *gs, const char *m)
{
  if (!strcmp(m, "epsilon")) return gs->par.epsilon;
  if (!strcmp(m, "maxiter")) return gs->par.maxiter;
  if (!strcmp(m, "meritval")) return gs->par.meritval;
  if (!strcmp(m, "diagepsilon")) return gs->par.diagepsilon;
  if (!strcmp(m, "diagmaxiter")) return gs->par.diagmaxiter;
  if (!strcmp(m, "diagsolver")) return gs->par.diagsolver;
  if (!strcmp(m, "failure")) return gs->par.failure;
  if (!strcmp(m, "reverse")) return gs->par.reverse;
  if (!strcmp(m, "variant")) return gs->variant;
  if (!strcmp(m, "innerloops")) return gs->innerloops;
  if (!strcmp(m, "iters")) return gs->iters;
  if (!strcmp(m, "error")) return gs->error;
  if (!strcmp(m, "precision")) return gs->par.precision;
  if (!strcmp(m, "refine")) return gs->par.refine;
  if (!strcmp(m, "warmstart")) return gs->warmstart;
  return -1.0;
}

int GAUSS_SEIDEL_Set(GAUSS_SEIDEL *gs, const char *m, double v)
{
  if (!strcmp(m, "epsilon")) { if (!(v > 0.0)) return -1; gs->par.epsilon = v; return 0; }
  if (!strcmp(m, "maxiter")) { if (!(v > 0.0)) return -1; gs->par.maxiter = (int)v; return 0; }
  if (!strcmp(m, "meritval")) { if (!(v > 0.0)) return -1; gs->par.meritval = v; return 0; }
  if (!strcmp(m, "diagepsilon")) { if (!(v > 0.0)) return -1; gs->par.diagepsilon = v; return 0; }
  if (!strcmp(m, "diagmaxiter")) { if (!(v > 0.0)) return -1; gs->par.diagmaxiter = (int)v; return 0; }
  if (!strcmp(m, "diagsolver")) { if (v < 0 || v > 3) return -1; gs->par.diagsolver = (int)v; return 0; }
  if (!strcmp(m, "reverse")) { gs->par.reverse = v != 0.0; return 0; }
  if (!strcmp(m, "variant")) { if (v < 0 || v > 2) return -1; gs->variant = (int)v; return 0; }
  if (!strcmp(m, "innerloops")) { if (v < 1) return -1; gs->innerloops = (int)v; return 0; }
  if (!strcmp(m, "itershist")) { gs->itershistcount = v != 0.0 ? 0 : -1; return 0; }
  if (!strcmp(m, "precision")) { if (v != SFGS_PREC_FP64 && v != SFGS_PREC_MIXED) return -1; gs->par.precision = (int)v; return 0; }
  if (!strcmp(m, "refine")) { if (v < 0) return -1; gs->par.refine = (int)v; return 0; }
  if (!strcmp(m, "warmstart")) { if (v != 0 && v != 1) return -1; gs->warmstart = (int)v; return 0; }
  return -1;
}

const char* GAUSS_SEIDEL_Failure(GAUSS_SEIDEL *gs)
{
  switch (gs->par.failure)
  {
  case SFGS_FAILURE_CONTINUE: return "FAILURE_CONTINUE";
  case SFGS_FAILURE_EXIT: return "FAILURE_EXIT";
  case SFGS_FAILURE_CALLBACK: return "FAILURE_CALLBACK";
  }
  return NULL;
}

const char* GAUSS_SEIDEL_Diagsolver(GAUSS_SEIDEL *gs)
{
  switch (gs->par.diagsolver)
  {
  case SFGS_DS_PROJECTED_GRADIENT: return "PROJECTED_GRADIENT";
  case SFGS_DS_DE_SAXCE_FENG: return "DE_SAXCE_FENG";
  case SFGS_DS_SEMISMOOTH_NEWTON: return "SEMISMOOTH_NEWTON";
  case SFGS_DS_PROJECTED_NEWTON: return "PROJECTED_NEWTON";
  }
  return NULL;
}

const char* GAUSS_SEIDEL_Error(GAUSS_SEIDEL *gs)
{
  switch (gs->error)
  {
  case SFGS_GS_OK: return "OK";
  case SFGS_GS_DIVERGED: return "DIVERGED";
  case SFGS_GS_DIAGONAL_DIVERGED: return "DIAGONAL_DIVERGED";
  case SFGS_GS_DIAGONAL_FAILED: return "DIAGONAL_FAILED";
  }
  return NULL;
}

const char* GAUSS_SEIDEL_Reverse(GAUSS_SEIDEL *gs) { return gs->par.reverse ? "ON" : "OFF"; }

const char* GAUSS_SEIDEL_Variant(GAUSS_SEIDEL *gs)
{
  switch (gs->variant)
  {
  case 0: return "FULL";
  case 1: return "MIDDLE_JACOBI";
  case 2: return "BOUNDARY_JACOBI";
  }
  return NULL;
}

int GAUSS_SEIDEL_History(GAUSS_SEIDEL *gs, double *rerhist, double *merhist)
{
  if (rerhist && gs->rerhist) memcpy(rerhist, gs->rerhist, sizeof(double)*(size_t)gs->iters);
  if (merhist && gs->merhist) memcpy(merhist, gs->merhist, sizeof(double)*(size_t)gs->iters);
  return gs->iters;
}

int GAUSS_SEIDEL_Itershist(GAUSS_SEIDEL *gs, int *out, int cap)
{
  int n = gs->itershistcount > 0 ? gs->itershistcount : 0;
  if (out) memcpy(out, gs->itershist, sizeof(int)*(size_t)(n < cap ? n : cap));
  return n;
}

int GAUSS_SEIDEL_Report(GAUSS_SEIDEL *gs, SFGS_REPORT *rep)
{
  if (rep) *rep = gs->rep;
  return 0;
}

int GAUSS_SEIDEL_Solve_Device(GAUSS_SEIDEL *gs, sfgs_ctx *g, char *err, size_t errlen)
{
  int maxiter = gs->par.maxiter;
  if (gs->hist_cap < maxiter)
  {
    /* the histories hold at most maxiter values, MIXED included */
    double *r = realloc(gs->rerhist, sizeof(double)*(size_t)maxiter);
    if (r) gs->rerhist = r;
    double *m = r ? realloc(gs->merhist, sizeof(double)*(size_t)maxiter) : NULL;
    if (m) gs->merhist = m;
    if (!r || !m)
    {
      snprintf(err, errlen, "out of memory for the histories of %d iterations", maxiter);
      return SOL_ERR_MEM;
    }
    gs->hist_cap = maxiter;
  }
  int rc = SFGS_Solve(g, &gs->par, gs->rerhist, gs->merhist, &gs->rep);
  if (rc)
  {
    snprintf(err, errlen, "GAUSS_SEIDEL_Solve on GPU failed: %s (%s)", SFGS_Error_String(rc), SFGS_Last_Error(g));
    return SOL_ERR_GPU;
  }
  gs->iters = gs->rep.iters;
  gs->error = gs->rep.error;
  if (gs->itershistcount >= 0)
  {
    if (gs->itershistcount >= gs->itershistsize)
    {
      int *h = realloc(gs->itershist, sizeof(int)*((size_t)gs->itershistsize + 1024));
      if (!h)
      {
        snprintf(err, errlen, "out of memory for the iteration history");
        return SOL_ERR_MEM;
      }
      gs->itershist = h;
      gs->itershistsize += 1024;
    }
    gs->itershist[gs->itershistcount ++] = gs->iters;
  }
  switch (gs->par.failure)
  {
  case SFGS_FAILURE_EXIT:
    if (gs->rep.aborted) return SOL_ERR_GAUSS_SEIDEL_DIAGONAL_DIVERGED;
    if (gs->error == SFGS_GS_DIVERGED) return SOL_ERR_GAUSS_SEIDEL_DIVERGED;
    break;
  case SFGS_FAILURE_CALLBACK:
    for (int k = 0; k < gs->rep.fail_rows; k ++) if (gs->callback(gs->data) == 0) return SOL_ERR_CALLBACK_EXIT;
    if (gs->error == SFGS_GS_DIVERGED && gs->callback(gs->data) == 0) return SOL_ERR_CALLBACK_EXIT;
    break;
  }
  return SOL_OK;
}
