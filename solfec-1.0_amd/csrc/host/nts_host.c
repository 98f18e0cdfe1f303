/* nts_host.c -- the NEWTON_SOLVER object of the host layer.
 *
 * Defaults are those of the Python constructor (lng.c:4740-4857): meritval
 * 1e-8, maxiter 1000, locdyn ON, linver GMRES, linmaxiter 10, maxmatvec
 * 10000, epsilon 0.25, delta 0, theta 0.25, omega 1e-10, gsflag ON, reldelta
 * OFF.  The solve runs on the GPU (SFGS_Newton_Solve); histories and the
 * iteration count are kept as the reference keeps them in struct newton
 * (nts.h:31-66): merhist and mvhist of the last solve, itershist of every
 * solve once it is switched on. */
#include <stdlib.h>
#include <string.h>
#include <stdio.h>
#include "solfec.h"
#include "sol_internal.h"

NEWTON* NEWTON_Create(double meritval, int maxiter)
{
  if (!(meritval > 0.0) || maxiter <= 0) return NULL;
  NEWTON *ns = calloc(1, sizeof(NEWTON));
  if (!ns) return NULL;
  ns->par.meritval = meritval;
  ns->par.maxiter = maxiter;
  ns->par.linver = SFGS_NT_GMRES;
  ns->par.linmaxiter = 10;
  ns->par.maxmatvec = 10000;
  ns->par.epsilon = 0.25;
  ns->par.delta = 0.0;
  ns->par.theta = 0.25;
  ns->par.omega = 1e-10;
  ns->par.gsflag = 1;
  ns->par.reldelta = 0;
  ns->par.bodyspace = 0;   /* locdyn ON (nts.c:1312) */
  ns->itershistcount = -1;
  return ns;
}

void NEWTON_Destroy(NEWTON *ns)
{
  if (!ns) return;
  free(ns->merhist);
  free(ns->mvhist);
  free(ns->itershist);
  free(ns);
}

double NEWTON_Get(NEWTON *ns, const char *m)
{
  if (!strcmp(m, "meritval")) return ns->par.meritval;
  if (!strcmp(m, "maxiter")) return ns->par.maxiter;
  if (!strcmp(m, "linver")) return ns->par.linver;
  if (!strcmp(m, "linmaxiter")) return ns->par.linmaxiter;
  if (!strcmp(m, "maxmatvec")) return ns->par.maxmatvec;
  if (!strcmp(m, "epsilon")) return ns->par.epsilon;
  if (!strcmp(m, "delta")) return ns->par.delta;
  if (!strcmp(m, "theta")) return ns->par.theta;
  if (!strcmp(m, "omega")) return ns->par.omega;
  if (!strcmp(m, "gsflag")) return ns->par.gsflag;
  if (!strcmp(m, "reldelta")) return ns->par.reldelta;
  if (!strcmp(m, "locdyn")) return ns->par.bodyspace ? 0.0 : 1.0;
  if (!strcmp(m, "iters")) return ns->iters;
  return -1.0;
}

/* the setters validate as lng.c:4871-5120 does */
int NEWTON_Set(NEWTON *ns, const char *m, double v)
{
  if (!strcmp(m, "meritval")) { if (!(v > 0.0)) return -1; ns->par.meritval = v; return 0; }
  if (!strcmp(m, "maxiter")) { if (!(v >= 1.0)) return -1; ns->par.maxiter = (int)v; return 0; }
  if (!strcmp(m, "linver")) { if (v != SFGS_NT_GMRES && v != SFGS_NT_DIAG) return -1; ns->par.linver = (int)v; return 0; }
  if (!strcmp(m, "linmaxiter")) { if (!(v >= 1.0)) return -1; ns->par.linmaxiter = (int)v; return 0; }
  if (!strcmp(m, "maxmatvec")) { if (!(v >= 1.0)) return -1; ns->par.maxmatvec = (int)v; return 0; }
  if (!strcmp(m, "epsilon")) { if (!(v > 0.0)) return -1; ns->par.epsilon = v; return 0; }
  if (!strcmp(m, "delta")) { if (!(v >= 0.0)) return -1; ns->par.delta = v; return 0; }
  if (!strcmp(m, "theta")) { if (!(v > 0.0 && v <= 1.0)) return -1; ns->par.theta = v; return 0; }
  if (!strcmp(m, "omega")) { if (!(v > 0.0)) return -1; ns->par.omega = v; return 0; }
  if (!strcmp(m, "gsflag")) { ns->par.gsflag = v != 0.0; return 0; }
  if (!strcmp(m, "reldelta")) { if (v < 0 || v > 3) return -1; ns->par.reldelta = (int)v; return 0; }
  if (!strcmp(m, "locdyn")) { ns->par.bodyspace = v == 0.0; return 0; }   /* 1 ON, 0 OFF (nts.h:40) */
  if (!strcmp(m, "itershist")) { ns->itershistcount = v != 0.0 ? 0 : -1; return 0; }
  return -1;
}

int NEWTON_History(NEWTON *ns, double *merhist, int *mvhist)
{
  if (merhist && ns->merhist) memcpy(merhist, ns->merhist, sizeof(double)*(size_t)ns->iters);
  if (mvhist && ns->mvhist) memcpy(mvhist, ns->mvhist, sizeof(int)*(size_t)ns->iters);
  return ns->iters;
}

int NEWTON_Itershist(NEWTON *ns, int *out, int cap)
{
  int n = ns->itershistcount > 0 ? ns->itershistcount : 0;
  if (out) memcpy(out, ns->itershist, sizeof(int)*(size_t)(n < cap ? n : cap));
  return n;
}

int NEWTON_Report(NEWTON *ns, SFGS_NEWTON_REPORT *rep)
{
  if (!ns || !rep) return -1;
  *rep = ns->rep;
  return 0;
}

int NEWTON_Solve_Device(NEWTON *ns, sfgs_ctx *g, char *err, size_t errlen)
{
  int maxiter = ns->par.maxiter;
  if (ns->hist_cap < maxiter)
  {
    ns->merhist = realloc(ns->merhist, sizeof(double)*(size_t)maxiter);
    ns->mvhist = realloc(ns->mvhist, sizeof(int)*(size_t)maxiter);
    ns->hist_cap = maxiter;
  }
  int rc = SFGS_Newton_Solve(g, &ns->par, ns->merhist, ns->mvhist, &ns->rep);
  if (rc)
  {
    snprintf(err, errlen, "NEWTON_Solve on GPU failed: %s (%s)", SFGS_Error_String(rc), SFGS_Last_Error(g));
    return SOL_ERR_GPU;
  }
  ns->iters = ns->rep.iters;
  if (ns->itershistcount >= 0)   /* nts.c:1220-1229 */
  {
    if (ns->itershistcount >= ns->itershistsize)
    {
      ns->itershistsize += 1024;
      ns->itershist = realloc(ns->itershist, sizeof(int)*(size_t)ns->itershistsize);
    }
    ns->itershist[ns->itershistcount ++] = ns->iters;
  }
  return SOL_OK;
}
