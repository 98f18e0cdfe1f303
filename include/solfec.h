/* solfec.h -- host C layer (libsolfec.so) of the MI355X contact-solve build.
 *
 * It keeps the reference's solver object and time-stepping contract:
 *   GAUSS_SEIDEL_Create / Destroy / Solve and the string queries replace
 *   bgs.h:114-140 (bgs.c:443-472, 913-1069, 1072-1155);
 *   SOLFEC_Run replaces sol.c:665-825 for a GAUSS_SEIDEL_SOLVER, one time step
 *   being DOM_Update_Begin, LOCDYN_Update_Begin, GAUSS_SEIDEL_Solve,
 *   LOCDYN_Update_End, DOM_Update_End (sol.c:746-806).
 * The scene it steps is a minimal rigid-body domain: spheres (RIG) and fixed
 * half-space planes (OBS), Signorini-Coulomb contacts, gravity.  The contact
 * solve itself (assembly, Gauss-Seidel, body velocity update) runs on the GPU
 * through sfgs.h; this layer fails loudly when the GPU library cannot run.
 *
 * Errors: functions returning int give 0 on success; SOLFEC_Run returns the
 * SOL_ERR_* code that the reference would THROW (err.h), and RUN in the
 * Python layer turns it into RuntimeError(SOLFEC_Error_String(code)). */
#ifndef SOLFEC_H
#define SOLFEC_H

#include "sfgs.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct solfec SOLFEC;
typedef struct gauss_seidel GAUSS_SEIDEL;
typedef int (*GAUSS_SEIDEL_Callback)(void *data);  /* returns 0 to stop the program (lng.c:4213-4225) */

/* error codes carried out of SOLFEC_Run (err.h names) */
#define SOL_OK                              0
#define SOL_ERR_GAUSS_SEIDEL_DIAGONAL_DIVERGED 1
#define SOL_ERR_GAUSS_SEIDEL_DIVERGED          2
#define SOL_ERR_DOM_DEPTH                      3
#define SOL_ERR_GPU                            4
#define SOL_ERR_CALLBACK_EXIT                  5
#define SOL_ERR_ARG                            6
#define SOL_ERR_MEM                            7   /* a host array could not be sized: count out of range or no memory */

/* body kinds */
#define SOL_RIG 0
#define SOL_OBS 1

/* ---- GAUSS_SEIDEL object (bgs.h) ---- */
GAUSS_SEIDEL* GAUSS_SEIDEL_Create(double epsilon, int maxiter, double meritval, int failure,
                                  double diagepsilon, int diagmaxiter, int diagsolver,
                                  void *data, GAUSS_SEIDEL_Callback callback);
void GAUSS_SEIDEL_Destroy(GAUSS_SEIDEL *gs);
/* member access (struct gs, bgs.h:64-111) */
double GAUSS_SEIDEL_Get(GAUSS_SEIDEL *gs, const char *member);
int GAUSS_SEIDEL_Set(GAUSS_SEIDEL *gs, const char *member, double value);
const char* GAUSS_SEIDEL_Failure(GAUSS_SEIDEL *gs);
const char* GAUSS_SEIDEL_Diagsolver(GAUSS_SEIDEL *gs);
const char* GAUSS_SEIDEL_Error(GAUSS_SEIDEL *gs);
const char* GAUSS_SEIDEL_Reverse(GAUSS_SEIDEL *gs);
const char* GAUSS_SEIDEL_Variant(GAUSS_SEIDEL *gs);
/* histories of the last solve: rerhist/merhist have 'iters' entries */
int GAUSS_SEIDEL_History(GAUSS_SEIDEL *gs, double *rerhist, double *merhist);
/* iteration-count history across solves (itershist, enabled by setting "itershist" = 1) */
int GAUSS_SEIDEL_Itershist(GAUSS_SEIDEL *gs, int *out, int cap);
/* report of the last device solve */
int GAUSS_SEIDEL_Report(GAUSS_SEIDEL *gs, SFGS_REPORT *rep);

/* ---- domain ---- */
SOLFEC* SOLFEC_Create(int dynamic, double step, const char *outpath);
void SOLFEC_Destroy(SOLFEC *sol);
const char* SOLFEC_Error_String(int code);
const char* SOLFEC_Last_Error(SOLFEC *sol);
void SOLFEC_Gravity(SOLFEC *sol, double gx, double gy, double gz);
/* surface pair material (SURFACE_MATERIAL, SIGNORINI_COULOMB); s1 = s2 = -1 sets the default */
int SOLFEC_Surface_Material(SOLFEC *sol, int s1, int s2, double friction, double restitution, double cohesion);
/* surface material models (sps.h:52-54) */
#define SOLFEC_SIGNORINI_COULOMB 0
#define SOLFEC_SPRING_DASHPOT 1
/* SURFACE_MATERIAL with every constant (lng.c:2128-2190): model, friction,
 * restitution in [0, 1], cohesion, spring, dashpot (negative: critical
 * damping) and the Hertz power hpow (1 for a linear spring).  The constants
 * spring, dashpot and hpow are read by the PENALTY_SOLVER for every contact
 * (pes.c:272); a GAUSS_SEIDEL_SOLVER step refuses a domain that has a
 * SPRING_DASHPOT material (the reference applies the spring law inside its
 * sweep, dbs.c:521-526; this library does not) */
int SOLFEC_Surface_Material_Ex(SOLFEC *sol, int s1, int s2, int model, double friction, double restitution,
                               double cohesion, double spring, double dashpot, double hpow);
/* rigid sphere of given density; returns the body index */
int SOLFEC_Sphere(SOLFEC *sol, const double *centre, double radius, double density, int surface);
/* rigid ellipsoid (ELLIP, eli.c) with semi-axes 3 along the columns of axes
 * (column-major rotation of the referential frame, NULL = identity) and the
 * given density; returns the body index */
int SOLFEC_Ellipsoid(SOLFEC *sol, const double *centre, const double *semi_axes, const double *axes,
                     double density, int surface);
/* fixed half-space obstacle {x : (x - point).normal >= 0}; returns the body index */
int SOLFEC_Plane(SOLFEC *sol, const double *point, const double *normal, int surface);
/* set the velocity [W v] of a body (W referential angular, v spatial linear) */
int SOLFEC_Set_Velocity(SOLFEC *sol, int body, const double *velo6);
/* jittered cubic lattice of n spheres, radii U[rmin,rmax], in the box [0,L]^3
 * bounded by 6 planes of surface 0; seeded (splitmix64). Returns 0. */
int SOLFEC_Lattice_Scene(SOLFEC *sol, int n, double L, double rmin, double rmax, double jitter,
                         double density, unsigned long long seed);
/* the same lattice of ellipsoids: semi-axes U[amin,amax] each, uniformly random
 * orientations; seeded. Returns 0. */
int SOLFEC_Ellipsoid_Lattice_Scene(SOLFEC *sol, int n, double L, double amin, double amax, double jitter,
                                   double density, unsigned long long seed);

/* SOLFEC_Run (sol.c:665-825): step until time >= t0 + duration */
int SOLFEC_Run(SOLFEC *sol, GAUSS_SEIDEL *gs, double duration);

/* ---- NEWTON_SOLVER (nts.c, locdyn 'ON') ---- */
typedef struct newton NEWTON;
/* NEWTON_Create (nts.c:1305-1331) with the Python constructor's defaults
 * (lng.c:4740-4857) for the other fields (see SFGS_NEWTON_PARAM) */
NEWTON* NEWTON_Create(double meritval, int maxiter);
void NEWTON_Destroy(NEWTON *ns);   /* nts.c:1388-1394 */
/* members by name: meritval, maxiter, linver (SFGS_NT_*), linmaxiter,
 * maxmatvec, epsilon, delta, theta, omega, gsflag, reldelta (0 OFF, 1 avgWii,
 * 2 minWii, 3 maxWii), locdyn (1 ON, 0 OFF: body-space products), iters; "itershist" (set only) switches the per-solve
 * iteration history on or off.  NEWTON_Set returns -1 for an invalid value */
double NEWTON_Get(NEWTON *ns, const char *member);
int NEWTON_Set(NEWTON *ns, const char *member, double value);
/* merit and operator-product histories of the last solve; returns iters */
int NEWTON_History(NEWTON *ns, double *merhist, int *mvhist);
int NEWTON_Itershist(NEWTON *ns, int *out, int cap);
int NEWTON_Report(NEWTON *ns, SFGS_NEWTON_REPORT *rep);
/* LOCDYN_Update_Begin + NEWTON_Solve (sol.c:472) on the GPU + the velocity
 * update, as SOLFEC_Step_Solve */
int SOLFEC_Step_Solve_Newton(SOLFEC *sol, NEWTON *ns);
/* SOLFEC_Run with a NEWTON_SOLVER */
int SOLFEC_Run_Newton(SOLFEC *sol, NEWTON *ns, double duration);

/* ---- PENALTY_SOLVER (pes.c) ---- */
typedef struct penalty PENALTY;
/* PENALTY_Create (pes.c:236-245): implicit 1 = variant 'IMPLICIT', 0 = 'EXPLICIT' */
PENALTY* PENALTY_Create(int implicit);
void PENALTY_Destroy(PENALTY *ps);   /* pes.c:300-303 */
/* LOCDYN_Update_Begin in its penalty form + PENALTY_Solve (pes.c:247-281) on
 * the GPU (SFGS_Penalty_Solve) + the velocity update, as SOLFEC_Step_Solve */
int SOLFEC_Step_Solve_Penalty(SOLFEC *sol, PENALTY *ps);
/* SOLFEC_Run with a PENALTY_SOLVER (sol.c:471) */
int SOLFEC_Run_Penalty(SOLFEC *sol, PENALTY *ps, double duration);
/* single step phases (sol.c:761-773) for drivers that interpose work */
/* 1: the step functions run DOM_Update_Begin / DOM_Update_End on the device
 * (SFGS_Dom_Update_Begin / _End) and keep bodies and contacts there; the host
 * arrays are refreshed when read (SOLFEC_Get_*, dumps) and uploaded again after
 * a change on the host.  Same results as 0, the host loop, bit for bit. */
int SOLFEC_Device_Loop(SOLFEC *sol, int on);
int SOLFEC_Step_Begin(SOLFEC *sol);                    /* DOM_Update_Begin */
int SOLFEC_Step_Solve(SOLFEC *sol, GAUSS_SEIDEL *gs);  /* LOCDYN_Update_Begin + solve (+ Update_End) */
int SOLFEC_Step_End(SOLFEC *sol);                      /* DOM_Update_End */

double SOLFEC_Time(SOLFEC *sol);
int SOLFEC_Nbod(SOLFEC *sol);
int SOLFEC_Ncon(SOLFEC *sol);
/* wall-clock seconds accumulated per phase: "TIMINT", "CONUPD", "CONDET", "LOCDYN", "CONSOL" */
double SOLFEC_Timing(SOLFEC *sol, const char *label);
/* device context used by the domain (created on first use) */
sfgs_ctx* SOLFEC_Gpu(SOLFEC *sol);

/* state export (natural order); NULL skips an array */
int SOLFEC_Get_Bodies(SOLFEC *sol, int *kind, double *rot, double *x, double *x0, double *jinv,
                      double *minv, double *velo, double *velo0, double *radius);
int SOLFEC_Get_Contacts(SOLFEC *sol, int *master, int *slave, double *mpnt, double *spnt, double *base,
                        double *gap, double *friction, double *restitution, double *cohesion,
                        double *R, double *U);
/* replace the carried reactions / velocities (warm start) of the current contacts */
int SOLFEC_Set_RU(SOLFEC *sol, const double *R, const double *U);
/* LOCDYN_Dump (ldy.c:892-1002) and LOCDYN_W_MatrixMarket (ldy.c:1005-1052) of
 * the last assembled step, in the reference's text formats (rows in list
 * order, newest contact first); the _Write forms take host arrays in contact
 * order (natural CSR adjacency as SFGS_Get_Locdyn returns it) */
int SOLFEC_Locdyn_Dump(SOLFEC *sol, const char *path);
int SOLFEC_W_MatrixMarket(SOLFEC *sol, const char *path);
int SOLFEC_Locdyn_Write(const char *path, int n, const double *point, const int *master, const int *slave,
                        const double *base, const double *Wd, const double *B, const double *V, const double *U,
                        const double *R, const int *rowptr, const int *col, const double *W);
int SOLFEC_W_MatrixMarket_Write(const char *path, int n, const double *Wd, const int *rowptr, const int *col,
                                const double *W);
/* snapshot writer: after each assembly, the features of every contact (18
 * floats, natural order, see SFGS_Features) are copied to buffer when the step
 * has at most cap contacts; SOLFEC_Captured gives the count of the last step */
/* State file for restart and warm-start persistence: body records as
 * BODY_Write_State (bod.c:2202-2221) and constraint records as write_constraint
 * (dio.c:19-59, IOVER 4), XDR-encoded, after a header (see sol_state.c).  Read
 * into a domain with the same bodies (same scene construction) and step; the
 * contacts, their R and U, the body configurations, velocities and the time are
 * replaced. */
int SOLFEC_Write_State(SOLFEC *sol, const char *path);
int SOLFEC_Read_State(SOLFEC *sol, const char *path);

int SOLFEC_Capture(SOLFEC *sol, float *buffer, int cap);
int SOLFEC_Captured(SOLFEC *sol);

#ifdef __cplusplus
}
#endif
#endif
