/* sol_internal.h -- private state of the host domain and the solver object. */
#ifndef SOL_INTERNAL_H
#define SOL_INTERNAL_H

#include <stddef.h>
#include "solfec.h"

typedef struct {
  int s1, s2;
  int model;                         /* SOLFEC_SIGNORINI_COULOMB or SOLFEC_SPRING_DASHPOT */
  double friction, restitution, cohesion;
  double spring, dashpot, hpow;      /* the penalty solver's spring-dashpot law */
} MATPAIR;

struct solfec {
  int dynamic;
  double step, time;
  double gravity[3];
  char outpath[512];
  char err[512];
  int device;
  sfgs_ctx *gpu;

  /* bodies */
  int nb, bcap;
  int *kind, *shape, *surf;        /* shape: 0 sphere, 1 plane, 2 ellipsoid */
  double *rad, *plane;             /* rad: radius, or bounding radius of an ellipsoid; plane: point 3, unit normal 3 */
  double *sca, *rot0;              /* ellipsoid: semi-axes 3, principal axes in the referential frame 9 */
  double *rot, *x, *x0, *J, *jinv, *minv, *mass;
  double *velo, *velo0, *aux, *force, *tmp_velo;

  /* materials */
  MATPAIR defmat;
  MATPAIR *mat;
  int nmat;

  /* contacts */
  int nc, ccap;
  int *c_master, *c_slave;
  double *c_mpnt, *c_spnt, *c_point, *c_base, *c_gap, *c_fri, *c_rest, *c_coh, *c_R, *c_U;

  /* detection scratch */
  int *grid_head, *grid_next;
  size_t grid_cap, gnext_cap;
  unsigned long long *pair_keys;
  int *pair_vals;
  size_t pair_cap, pair_mask;
  double *upd_geo;
  int *upd_ok;
  size_t upd_cap;
  int **cand, *cand_n, *cand_cap, cand_nthr;

  /* learned warm start: features of the last assembled step (snapshot writer) */
  float *capture;
  int capture_cap, captured;

  /* device loop (SOLFEC_Device_Loop): the step functions keep bodies and
   * contacts on the device; dev_valid: the device holds the current state;
   * host_stale: the host arrays above are behind it */
  int devloop, dev_valid, host_stale;
  int locdyn_current;   /* the device holds the local dynamics of the current contact list */

  /* phase timers (sol.c:392-439 labels), seconds */
  double t_timint, t_conupd, t_condet, t_locdyn, t_consol;
};

struct gauss_seidel {
  SFGS_PARAM par;
  int variant, innerloops;
  int warmstart;          /* 0: previous-step R (the reference), 1: learned (SFGS_Warm_Start) */
  void *data;
  GAUSS_SEIDEL_Callback callback;
  int iters, error;
  double *rerhist, *merhist;
  int hist_cap;
  int *itershist;
  int itershistcount, itershistsize;
  SFGS_REPORT rep;
};

struct newton {
  SFGS_NEWTON_PARAM par;
  int iters;
  double *merhist;
  int *mvhist;
  int hist_cap;
  int *itershist;
  int itershistcount, itershistsize;
  SFGS_NEWTON_REPORT rep;
};

struct penalty {
  int implicit;
  double *kcp;            /* spring, dashpot, hpow per contact (host loop) */
  int kcp_cap;
};

/* host-layer internals shared by sol_host.c and sol_state.c */
int sol_dev_pull(SOLFEC *sol);             /* device state into the host arrays, when behind */
void sol_dev_edit(SOLFEC *sol);            /* before a host-side change of the state */
int sol_ensure_contacts(SOLFEC *sol, long long n);

/* run the device Newton solve for the current system of g (nts_host.c) */
int NEWTON_Solve_Device(NEWTON *ns, sfgs_ctx *g, char *err, size_t errlen);

/* run the device solve for the current system of g; maps failures to SOL_ERR_* */
int GAUSS_SEIDEL_Solve_Device(GAUSS_SEIDEL *gs, sfgs_ctx *g, char *err, size_t errlen);

#endif
