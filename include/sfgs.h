/* sfgs.h -- C-ABI of the MI355X (gfx950) contact-solve library libsfgs.so.
 *
 * The library replaces, for rigid bodies and Signorini-Coulomb contacts, the
 * Solfec-1.0 hot path between DOM_Update_Begin and DOM_Update_End
 * (sol.c:764-773): LOCDYN assembly, the block Gauss-Seidel solve with its
 * local 3x3 solvers, the merit function, and the body velocity update by the
 * reactions.  All signatures use plain pointers and sizes; host arrays are in
 * natural contact / body order; 3x3 matrices are column-major, contact frames
 * have columns [t1 t2 n] as in the reference (alg.h:881-953, dom.c:133-151).
 *
 * Every call is synchronous with respect to the host arrays it reads or
 * writes.  A context owns its device memory and one HIP stream of default
 * priority; calls on one context must not overlap.  Functions return 0 on
 * success and a negative SFGS_E* code on failure (SFGS_Error_String). */
#ifndef SFGS_H
#define SFGS_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct sfgs_ctx sfgs_ctx;

/* error codes */
#define SFGS_OK            0
#define SFGS_E_HIP        -1   /* a HIP runtime call failed */
#define SFGS_E_ARG        -2   /* invalid argument or size */
#define SFGS_E_STATE      -3   /* call out of order (e.g. solve before assembly) */
#define SFGS_E_COLOURS    -4   /* contact graph needs more colours than supported */
#define SFGS_E_SINGULAR   -5   /* a diagonal block W_ii is singular */

/* local solvers: numbering of DIAS, dbs.h:27-33 */
#define SFGS_DS_PROJECTED_GRADIENT 0
#define SFGS_DS_DE_SAXCE_FENG      1
#define SFGS_DS_SEMISMOOTH_NEWTON  2
#define SFGS_DS_PROJECTED_NEWTON   3

/* failure actions: GSFAIL, bgs.h:39-44 */
#define SFGS_FAILURE_CONTINUE 0
#define SFGS_FAILURE_EXIT     1
#define SFGS_FAILURE_CALLBACK 2

/* solver status: GSERROR, bgs.h:31-37 */
#define SFGS_GS_OK                0
#define SFGS_GS_DIVERGED          1
#define SFGS_GS_DIAGONAL_DIVERGED 2
#define SFGS_GS_DIAGONAL_FAILED   3

/* sweep precision */
#define SFGS_PREC_FP64  0   /* FP64 storage and arithmetic (the reference) */
#define SFGS_PREC_MIXED 1   /* FP32 W/B/R sweeps, FP64 merit and termination, then FP64 sweeps */

/* Gauss-Seidel parameters: the fields of struct gs, bgs.h:64-111, with the
 * defaults of GAUSS_SEIDEL_SOLVER, lng.c:4243-4349 */
typedef struct {
  double epsilon;      /* relative error bound */
  int maxiter;         /* iteration bound */
  double meritval;     /* merit bound; >= 1 switches the merit off (bgs.c:927) */
  int failure;         /* SFGS_FAILURE_* */
  double diagepsilon;  /* local solver accuracy */
  int diagmaxiter;     /* local solver iteration bound */
  int diagsolver;      /* SFGS_DS_* */
  int reverse;         /* 1: alternate forward / backward sweeps (bgs.c:936-950) */
  int precision;       /* SFGS_PREC_* */
  int refine;          /* MIXED: FP32 sweeps stop refine iterations before maxiter (or at the FP32
                          floor: relative error 1e-6, merit 1e-10), then FP64 sweeps, at least
                          refine (default 2), run until the termination test holds in FP64 or the
                          maxiter iterations are spent; maxiter <= refine solves in FP64 */
  int profile;         /* 1: time every sweep-kernel launch with HIP events (slower) */
} SFGS_PARAM;

typedef struct {
  int iters;           /* sweeps performed (gs->iters) */
  int error;           /* SFGS_GS_* (gs->error) */
  double rel_error;    /* last relative error */
  double merit;        /* last merit value (dom->merit) */
  int ncolours;        /* colours of the contact graph */
  int fail_rows;       /* local-solver failures (callback invocations) */
  int aborted;         /* 1 when FAILURE_EXIT stopped the solve */
  double solve_ms;     /* device time of the solve (HIP events) */
  double sweep_ms;     /* device time of the sweep-kernel launches (profile mode) */
  long long sweep_launches; /* sweep-kernel launches timed in profile mode */
  long long row_updates; /* rows visited = sweeps x rows */
  long long nnz;       /* off-diagonal blocks */
  int fp64_sweeps;     /* MIXED: sweeps of the FP64 phase (the rest of row_updates / n are FP32) */
  int launches_per_sweep; /* sweep-kernel launches per sweep of the colour-launch schedules (the
                             trailing colours of a few slices each share one launch) */
} SFGS_REPORT;

/* context: LOCDYN_Create, ldy.h:93 (one device-resident W/B/R store) */
sfgs_ctx* SFGS_Create(int device);
void SFGS_Destroy(sfgs_ctx *ctx);
const char* SFGS_Error_String(int code);
const char* SFGS_Last_Error(sfgs_ctx *ctx);

/* Rigid bodies, natural order (bod.h BODY of kind RIG/OBS): kind 0 = RIG,
 * 1 = OBS; rot 9, x0 3 (referential centre), jinv 9 (inverse referential
 * inertia), minv 1, velo 6 = [W v] after BODY_Dynamic_Step_Begin (bod.c:1313-1373),
 * velo0 6 = [W v] at the start of the step. */
int SFGS_Bodies_Upload(sfgs_ctx *ctx, int nb, const int *kind, const double *rot, const double *x0,
                       const double *jinv, const double *minv, const double *velo, const double *velo0);

/* Contacts, natural order (CON, dom.h:77-137): master/slave body indices,
 * referential points mpnt/spnt 3 each, base 9, gap, friction, restitution,
 * cohesion (times area), and the carried-over reactions R and relative
 * velocities U, 3 each (zero for new contacts, mem.c:77,86). */
int SFGS_Contacts_Upload(sfgs_ctx *ctx, int nc, const int *master, const int *slave,
                         const double *mpnt, const double *spnt, const double *base,
                         const double *gap, const double *friction, const double *restitution,
                         const double *cohesion, const double *R, const double *U);

/* LOCDYN_Update_Begin (ldy.h:103, ldy.c:626-876): contact adjacency through
 * shared non-obstacle bodies (LOCDYN_Insert, ldy.c:492-568), contact-graph
 * colouring, colour-major block-sparse W, W_ii, A = inv(W_ii), rho, B, V and
 * the free energy, all on the device.  dynamic = 1 for DYNAMIC analysis. */
int SFGS_Locdyn_Update_Begin(sfgs_ctx *ctx, double step, int dynamic);
/* LOCDYN_Update_Begin with UPMIN (ldy.c:59, 644), as the reference runs it for
 * NEWTON_SOLVER with locdyn OFF: W_ii, A, rho, B, V and the free energy per
 * contact and the body incidence, no adjacency and no off-diagonal blocks
 * (every row in one colour; SFGS_Get_Locdyn returns nnz = 0).  Only
 * SFGS_Newton_Solve with bodyspace = 1, the merit without the U update, the
 * body update and the cohesion update accept such a store; the other solves
 * return SFGS_E_STATE. */
int SFGS_Locdyn_Update_Min(sfgs_ctx *ctx, double step, int dynamic);

/* Alternative to the two uploads + assembly: install an already assembled
 * system (natural order, off-diagonal blocks in CSR rowptr/col, 9 per block)
 * together with R and U.  Used to solve identical W, B, R0 on CPU and GPU. */
int SFGS_System_Upload(sfgs_ctx *ctx, int n, const int *rowptr, const int *col, const double *W,
                       const double *Wd, const double *A, const double *B, const double *V,
                       const double *rho, const double *friction, const double *restitution,
                       const double *cohesion, const double *gap, double free_energy,
                       double step, int dynamic, const double *R, const double *U);

/* GAUSS_SEIDEL_Solve (bgs.h:119, serial bgs.c:913-1069) on the current
 * system.  rerhist / merhist (may be NULL) receive par->maxiter doubles.  On
 * return R and U (SFGS_Get_RU) are the reference's final state: when the solve
 * stops on its test or at maxiter with the merit on, U is the merit's update
 * U = B + W R of the final R (mrf.c:61-76); otherwise (no merit, or
 * FAILURE_EXIT) the local solver's U of the last sweep. */
int SFGS_Solve(sfgs_ctx *ctx, const SFGS_PARAM *par, double *rerhist, double *merhist, SFGS_REPORT *rep);

/* MERIT_Function (mrf.h:31, mrf.c:35-184) for the current R; update_U as in
 * the reference (U recomputed from all R); conmerit (may be NULL) receives the
 * per-contact merit (con->merit). */
int SFGS_Merit(sfgs_ctx *ctx, int update_U, double *merit, double *conmerit);

/* Download the current R, U (natural order, 3 per contact); NULL skips. */
int SFGS_Get_RU(sfgs_ctx *ctx, double *R, double *U);

/* update_cohesion of LOCDYN_Update_End (ldy.c:69-95, 879-889), after a
 * Gauss-Seidel or Newton solve (not after SFGS_Penalty_Solve, whose law breaks
 * cohesion itself): a contact with cohesion state c > 0 (the cohesive force,
 * cohesion times area) loses it, c = 0, when R_N + c < e (mode I) or
 * |R_T| + e >= friction (R_N + c) (mode II), e = 1e-8 c (COHESION_EPSILON,
 * sps.h:34), with the solved R.  The states are those of SFGS_Contacts_Upload
 * or of the device loop; SFGS_Get_Cohesion reads them back. */
int SFGS_Update_Cohesion(sfgs_ctx *ctx);

/* Download the contacts' cohesion states (natural order, 1 per contact): what
 * SURFACE_MATERIAL_Cohesion_Get (sps.h:127, sps.c:356) returns per CON. */
int SFGS_Get_Cohesion(sfgs_ctx *ctx, double *cohesion);

/* Upload the contacts' cohesion states (natural order, 1 per contact), as
 * SURFACE_MATERIAL_Cohesion_Set (sps.h:130, sps.c:363) sets them per CON; the
 * host layer restores with it the states that update_cohesion leaves alone
 * (SPRING_DASHPOT contacts, ldy.c:78) after a NEWTON_SOLVER step. */
int SFGS_Set_Cohesion(sfgs_ctx *ctx, const double *cohesion);

/* Download the assembled LOCDYN in natural order for inspection: rowptr
 * (n+1), col/via (nnz), W (9 nnz), Wd, A (9 n), B, V (3 n), rho (n), the free
 * energy and the colour of every contact; NULL skips an array. */
int SFGS_Get_Locdyn(sfgs_ctx *ctx, int *rowptr, int *col, int *via, double *W, double *Wd,
                    double *A, double *B, double *V, double *rho, double *free_energy, int *colour);
int SFGS_Get_Sizes(sfgs_ctx *ctx, int *n, long long *nnz, int *ncolours);

/* Lanes of a wavefront that serve one contact row in the sweep layout (1, 2,
 * 4 or 8; default 2).  Takes effect at the next LOCDYN assembly or system
 * upload.  No counterpart in the reference (a tuning knob of the GPU layout). */
int SFGS_Set_Layout(sfgs_ctx *ctx, int lanes_per_row);

/* Schedule of the Gauss-Seidel sweeps of SFGS_Solve (no counterpart in the
 * reference, which sweeps serially): 0 = one kernel launch per colour, each
 * XCD taking a contiguous run of the colour's rows (default); 1 = one
 * persistent dataflow launch in which each row waits only for its own
 * neighbours, with the same iterates (forward sweeps; a solve with reverse = 1,
 * or with 4 or 8 lanes per row, uses schedule 0); 2 = one launch per colour in
 * the hardware's workgroup order; 3 = one persistent launch whose workgroups
 * meet at a device-wide barrier after every colour, with the same iterates
 * (1 or 2 lanes per row; otherwise schedule 0). */
int SFGS_Set_Schedule(sfgs_ctx *ctx, int schedule);
/* Wavefronts per workgroup of the per-colour sweep launches (1, 2 or 4; one
 * slice of 64 lanes per wavefront), or 0 (default): 2 for FP64 solves and 4
 * for MIXED, the faster of each at the 200k-sphere packing.  No counterpart
 * in the reference.  Same iterates for every value. */
int SFGS_Set_Sweep_Block(sfgs_ctx *ctx, int waves_per_block);

/* Order of the rows inside a colour (no counterpart in the reference): 0 =
 * decreasing row length (least slice padding); 1 = contact index; 2 = master
 * body in blocks of 64 bodies, then decreasing row length; 3 (default) = the
 * Z-order (Morton code on a 1024^3 grid over the bodies' bounding box) of the
 * master body's position, which keeps the rows of bodies close in space
 * together, so that their neighbours' reactions share cache lines.  Orders 2
 * and 3 need a system assembled from bodies (otherwise 0).  Takes effect at the
 * next assembly; the iterates do not depend on it (rows of a colour are
 * independent). */
int SFGS_Set_Row_Order(sfgs_ctx *ctx, int order);

/* Colour an arbitrary graph given in CSR (no self loops) with the device
 * colouring kernel; returns the number of colours or a negative code. */
int SFGS_Colour(sfgs_ctx *ctx, int n, const int *rowptr, const int *col, int *colour);

/* SFGS_Colour without any allocation.  No counterpart in the reference: it is
 * for callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  rowptr (n + 1), col
 * (nnz) and colour (n) are DEVICE pointers the caller owns; colour receives the
 * same colours as SFGS_Colour (bit for bit).  The rounds' scratch is a
 * file-scope device array of this library, one per GPU, held until the process
 * ends: n is limited to SFGS_STATIC_COLOUR_MAX (larger n returns SFGS_E_ARG),
 * and a graph that is not valid CSR of nnz entries (rowptr[0] = 0,
 * non-decreasing, rowptr[n] = nnz, columns in [0, n)) is refused with
 * SFGS_E_ARG before any colouring round runs.  Calls that overlap on one GPU
 * share that array and race; each call clears what it needs.  Returns the
 * number of colours or a negative code. */
#define SFGS_STATIC_COLOUR_MAX (1 << 20)
int SFGS_Colour_Static(sfgs_ctx *ctx, int n, long long nnz, const int *rowptr, const int *col, int *colour);

/* Body velocity update of BODY_Dynamic_Step_End (bod.c:1446-1447 with
 * rig_constraints_force, bod.c:418-435): velo += h iM r, r = sum(+-H' R) over
 * the contacts of every rigid body, using the current R; velo_out receives 6
 * per body, force_out the generalised reaction r, 6 per body; either may be
 * NULL, in which case that result stays on the device. */
int SFGS_Body_Update(sfgs_ctx *ctx, double step, double *velo_out, double *force_out);

/* ---- learned warm start (no counterpart in the reference) ----
 * A 18 -> 128 -> 128 -> 128 -> 3 GELU perceptron (input padded to 32, output to
 * 32), bf16 weights, FP32 accumulation, trained and run with the library's own
 * bf16 matrix-core kernels.  Features of contact i: gap, B (3), W_ii (9),
 * R_prev (3), number of off-diagonal blocks, friction.  Output: R - R_prev in
 * units of the output scale.  Parameters are FP32, layout per layer [in][out]:
 * W1 32x128, b1 128, W2 128x128, b2 128, W3 128x128, b3 128, W4 128x32, b4 32
 * (41,376 values); norm = 18 feature means, 18 feature scales, 3 output scales. */
int SFGS_Mlp_Set(sfgs_ctx *ctx, const float *params, long long nparams, const float *norm);
int SFGS_Mlp_Get(sfgs_ctx *ctx, float *params, long long *nparams);
/* prediction for host features (n x 18): out n x 3, units of R */
int SFGS_Mlp_Forward(sfgs_ctx *ctx, const float *features, int n, float *out);
/* loss (mean squared error in normalised units) and its gradient for one batch */
int SFGS_Mlp_Grad(sfgs_ctx *ctx, const float *features, const float *targets, int n, float *grad, float *loss);
/* Adam training on the device over 'epochs' passes of seeded random minibatches */
int SFGS_Mlp_Train(sfgs_ctx *ctx, const float *features, const float *targets, long long n, int epochs, int batch,
                   float lr, unsigned long long seed, float *loss_hist);
/* features of the current system (natural order, n x 18): the snapshot writer */
int SFGS_Features(sfgs_ctx *ctx, float *features);
/* replace the carried R of the current system by R_prev + MLP(features) (call
 * after SFGS_Locdyn_Update_Begin, before SFGS_Solve) */
int SFGS_Warm_Start(sfgs_ctx *ctx);
/* what SFGS_Warm_Start changes: new_only = 1 applies the prediction only to
 * contacts without a carried reaction (R_prev = 0, new contacts, mem.c:77)
 * and keeps R_prev elsewhere; scale multiplies the prediction (default 0, 1) */
int SFGS_Warm_Start_Mode(sfgs_ctx *ctx, int new_only, float scale);

/* ---- batched solve of independent scenes (no counterpart in the reference,
 * which solves one domain per GAUSS_SEIDEL_Solve call, bgs.c:913-1069) ----
 * The uploaded bodies are split into scenes by a scene id per body; contacts
 * belong to their master body's scene and may not couple rigid bodies of two
 * scenes.  Each scene runs its own termination test (bgs.c:1036), so a scene
 * that converges stops while the others go on.  FP64 solves run either as one
 * persistent kernel launch, each scene owning a set of workgroups that sweep
 * its colours and synchronise among themselves only, or with one launch per
 * colour over every running scene (the tune's schedule).  MIXED solves take
 * the per-colour schedule and, per scene, the two phases of SFGS_Solve's
 * MIXED: FP32 sweeps under the test with epsilon and meritval held at or
 * above 1e-6 and 1e-10, for at most maxiter - refine iterations, then FP64
 * sweeps under the reference's test, at least refine of them, within the
 * scene's remaining iterations.  Call SFGS_Batch_Scenes after
 * SFGS_Bodies_Upload and before SFGS_Locdyn_Update_Begin (a body upload
 * resets the split); SFGS_Solve refuses a split system. */
typedef struct {
  int lanes_per_row;    /* layout the system was assembled with (SFGS_Set_Layout): 1 or 2 */
  int waves_per_group;  /* wavefronts per workgroup: 1, 2, 4 or 8 */
  int rows_per_group;   /* target rows of a scene per workgroup; 0 = share the resident workgroups
                           among the scenes in proportion to their rows (one launch) */
  int schedule;         /* 0 = one persistent launch (the fields above; FP64 only, MIXED uses 1);
                           1 = one launch per colour over the colour's slices of every running
                           scene, a scene-wise termination test after each sweep */
} SFGS_BATCH_TUNE;

typedef struct {
  int rows;             /* contacts of the scene */
  int groups;           /* workgroups that solved it */
  int iters;            /* Gauss-Seidel iterations (MIXED: FP32 + FP64 iterations) */
  int error_code;       /* SFGS_GS_* */
  int fail_rows;        /* local solver failures */
  int aborted;          /* FAILURE_EXIT raised */
  double error;         /* last relative error */
  double merit;         /* last merit (0 with the merit off) */
  double energy;        /* free energy 1/2 sum B'AB of the scene */
} SFGS_SCENE_REPORT;

int SFGS_Batch_Scenes(sfgs_ctx *ctx, int nscenes, const int *body_scene);
/* solve every scene from its current R; tune == NULL uses the last autotuned
 * (or default) shape; reports receives nscenes records, total (may be NULL)
 * the launch totals (iters = largest scene count, solve_ms = device time) */
int SFGS_Batch_Solve(sfgs_ctx *ctx, const SFGS_PARAM *par, const SFGS_BATCH_TUNE *tune, SFGS_SCENE_REPORT *reports,
                     SFGS_REPORT *total);
/* time trial_iters iterations of every candidate shape (lanes 1 and 2, which
 * re-assembles the system; for FP64 the persistent launch with waves 1..8 and
 * rows per group proportional, 1024, 4096 or all, and the per-colour schedule;
 * for MIXED the per-colour schedule) and keep the fastest; R and U are
 * restored after each trial */
int SFGS_Batch_Autotune(sfgs_ctx *ctx, const SFGS_PARAM *par, int trial_iters, SFGS_BATCH_TUNE *best, double *best_ms);

/* ---- device-resident time stepping: DOM_Update_Begin / DOM_Update_End on the
 * device, so that a run loop keeps bodies and contacts in HBM ----
 * Rigid spheres and ellipsoids with fixed planes (OBS).  The body state is the
 * one of SFGS_Bodies_Upload (kind, rot, x0, J^-1, 1/m, velo, velo0) plus the
 * fields below; the contact list is the one of SFGS_Contacts_Upload, with the
 * spatial points given here.  Same arithmetic as the host layer of this
 * library (one shared source, no floating-point contraction), so the device
 * loop reproduces the host loop. */
typedef struct {
  const int *shape;      /* 0 sphere, 1 plane (an OBS body), 2 ellipsoid */
  const int *surf;       /* surface id, for the material table */
  const double *rad;     /* radius; the largest semi-axis of an ellipsoid */
  const double *plane;   /* planes: a point and the unit normal, 6 per body (others ignored) */
  const double *sca;     /* ellipsoids: semi-axes, 3 per body */
  const double *rot0;    /* ellipsoids: principal axes in the referential frame, 9 per body */
  const double *x;       /* current centre, 3 per body */
  const double *J;       /* inertia tensor, 9 per body */
  const double *mass;    /* mass */
  const double *aux;     /* exp[-(h/2)W] J W + h T carried from step begin to step end, 3 per body */
} SFGS_SHAPES;

/* surface pair material (SURFACE_MATERIAL, sps.h:40-66); the pair is
 * unordered; spring, dashpot and hpow (Hertz power, 1 for a linear spring) are
 * the constants of the penalty solver's spring-dashpot law */
typedef struct {
  int s1, s2;
  double friction, restitution, cohesion;
  double spring, dashpot, hpow;
} SFGS_MATERIAL;

/* body shapes, the material table (defmat for pairs not in it) and the
 * contacts' spatial points (3 per uploaded contact); call after
 * SFGS_Bodies_Upload and SFGS_Contacts_Upload */
int SFGS_Dom_Upload(sfgs_ctx *ctx, const SFGS_SHAPES *shapes, int nmat, const SFGS_MATERIAL *mat,
                    const SFGS_MATERIAL *defmat, const double *point);
/* DOM_Update_Begin (dom.c:3630) for this body set: the rigid half step
 * of every body with gravity (BODY_Dynamic_Step_Begin, bod.c:1313-1373); the
 * update of every contact, which keeps its slot, R and U, or is removed when
 * the bodies separate (update_contact, dom.c:430-483); detection of new
 * contacts (overlap_create, dom.c:348-416; sphere pairs goc.c:421-455,
 * plane-sphere goc.c:308-348), appended in ascending order of the body pair,
 * convex pairs before plane pairs, with R = U = 0 (mem.c:77,86).  *ncon
 * receives the new contact count.  SFGS_Locdyn_Update_Begin follows. */
int SFGS_Dom_Update_Begin(sfgs_ctx *ctx, double step, const double *gravity, int *ncon);
/* The domain time of the next SFGS_Dom_Update_Begin (dom->time): contacts it
 * creates at time 0 take their material's cohesion as cohesion state, later
 * ones none (SURFACE_MATERIAL_Transfer, sps.c:264-282).  A cohesive contact
 * (state > 0) is kept when its bodies separate, with its referential points,
 * its spatial point following the master body and its previous frame and gap
 * (update_contact, dom.c:452-458). */
int SFGS_Dom_Set_Time(sfgs_ctx *ctx, double time);
/* Analysis of the device loop (dom->dynamic): 1 (default) integrates the
 * rigid bodies with BODY_Dynamic_Step_Begin/End (bod.c:1313-1373, 1425-1492),
 * 0 with BODY_Static_Step_Begin/End (QUASI_STATIC, bod.c:1588-1645: u(t+h) =
 * h M^-1 f(t+h) from no carried velocity, x += h v, R = R exp[h W]).  The
 * assembly takes the same flag from SFGS_Locdyn_Update_Begin. */
int SFGS_Dom_Set_Analysis(sfgs_ctx *ctx, int dynamic);
/* DOM_Update_End (dom.c:3818) after SFGS_Solve: R and U back into the
 * contact list, the velocity update by the reactions (as SFGS_Body_Update)
 * and the rigid end of step (BODY_Dynamic_Step_End, bod.c:1425-1492) */
int SFGS_Dom_Update_End(sfgs_ctx *ctx, double step);
/* current body state (any pointer may be NULL): rot 9, x 3, velo 6, velo0 6,
 * aux 3 and the last generalised reaction force 6 per body */
int SFGS_Dom_Get_Bodies(sfgs_ctx *ctx, double *rot, double *x, double *velo, double *velo0, double *aux,
                        double *force);
/* current contact list, natural order (any pointer may be NULL) */
int SFGS_Dom_Get_Contacts(sfgs_ctx *ctx, int *master, int *slave, double *point, double *mpnt, double *spnt,
                          double *base, double *gap, double *friction, double *restitution, double *cohesion,
                          double *R, double *U);

/* NEWTON_SOLVER (nts.h:31-66): linear solver of the Newton step */
#define SFGS_NT_GMRES 0   /* PQN_GMRES: flexible GMRES with a block-diagonal preconditioner */
#define SFGS_NT_DIAG  1   /* PQN_DIAG: per-row steps with theta averaging */

/* the fields of struct newton (nts.h:31-66) with the defaults of
 * NEWTON_SOLVER, lng.c:4740-4857: meritval 1e-8, maxiter 1000, GMRES,
 * linmaxiter 10, maxmatvec 10000, epsilon 0.25, delta 0, theta 0.25,
 * omega 1e-10, gsflag 1, reldelta 0 */
typedef struct {
  double meritval;     /* merit value sufficient for termination */
  int maxiter;         /* Newton steps bound */
  int linver;          /* SFGS_NT_* */
  int linmaxiter;      /* GMRES iterations bound per step */
  int maxmatvec;       /* operator products bound (GMRES) */
  double epsilon;      /* GMRES tolerance, relative to |b| */
  double delta;        /* diagonal regularisation of the step operator */
  double theta;        /* DIAG relaxation, (0, 1] */
  double omega;        /* smoothing of the cone residual, > 0 */
  int gsflag;          /* 1: Gauss-Seidel from the initial R after an unconverged solve */
  int reldelta;        /* delta relative to W: 0 off, 1 mean, 2 min, 3 max of the W_ii diagonals */
  int bodyspace;       /* 1: locdyn 'OFF' -- W x and U = B + W R through the bodies, H (h M^-1 H' x)
                          (nts.c:396-431, 683-852); 0: the assembled W blocks (locdyn 'ON') */
} SFGS_NEWTON_PARAM;

typedef struct {
  int iters;           /* Newton steps (ns->iters) */
  int matvecs;         /* operator products */
  int converged;       /* 1: merit <= meritval */
  int gs_fallback;     /* 1: the Gauss-Seidel fallback ran */
  int gs_iters;        /* its iterations */
  int singular;        /* rows with a singular step block (the reference aborts there) */
  double merit;        /* final merit */
  double W_norm;       /* the W norm of reldelta (1 when off) */
} SFGS_NEWTON_REPORT;

/* NEWTON_Solve (nts.h:83, nts.c:1333-1379) with assembled local dynamics
 * (locdyn 'ON') on the system of SFGS_Locdyn_Update_Begin or
 * SFGS_System_Upload, from its current R: gmres_based_solve (nts.c:1184-1244)
 * or diagonalized_solve (nts.c:1246-1303), then, when unconverged and gsflag
 * is set, R reset and GAUSS_SEIDEL_SOLVER(1.0, maxiter, meritval, CONTINUE,
 * 1e-9, 100, SEMISMOOTH_NEWTON) (nts.c:1358-1372).  merhist and mvhist take
 * maxiter entries (may be NULL).  R and U follow from SFGS_Get_RU. */
int SFGS_Newton_Solve(sfgs_ctx *ctx, const SFGS_NEWTON_PARAM *par, double *merhist, int *mvhist,
                      SFGS_NEWTON_REPORT *rep);

/* z = W x formed through the bodies as NEWTON_SOLVER's locdyn 'OFF' mode forms
 * it (nts.c:396-431): v_b = h M_b^-1 sum of +-H' x over the contacts of body b,
 * then z_i = H_master v_master - H_slave v_slave per contact.  x and z are in
 * natural contact order, 3 per contact; needs a LOCDYN assembly from bodies
 * (SFGS_Locdyn_Update_Begin).  No counterpart of its own in the reference (the
 * operator of that mode, exposed so that it can be checked against W). */
int SFGS_Bodyspace_Product(sfgs_ctx *ctx, const double *x, double *z);

/* PENALTY_Solve (pes.c:247-281) for the uploaded bodies and contacts (or the
 * device loop's): per contact W_ii, B and V as SFGS_Locdyn_Update_Begin forms
 * them, without off-diagonal blocks (the penalty update, ldy.c:744), then the
 * spring-dashpot law PENALTY_Spring_Dashpot_Contact (pes.c:33-106), IMPLICIT
 * (implicit = 1) or EXPLICIT, from the carried R.  spring, dashpot and hpow
 * are given per contact (natural order), or all three NULL: then they come
 * from the material table of SFGS_Dom_Upload by the bodies' surfaces.  R and
 * U follow from SFGS_Get_RU, velocities from SFGS_Body_Update or
 * SFGS_Dom_Update_End.  The context holds no assembled system afterwards:
 * SFGS_Solve, SFGS_Merit and SFGS_Get_Locdyn return SFGS_E_STATE until the
 * next SFGS_Locdyn_Update_Begin.  A cohesive contact (cohesion state c > 0)
 * may take tension down to R_N = -c; it loses its cohesion (c = 0) below that
 * or when its tangential reaction reaches the cone (pes.c:59-99). */
int SFGS_Penalty_Solve(sfgs_ctx *ctx, double step, int implicit, const double *spring, const double *dashpot,
                       const double *hpow);

#ifdef __cplusplus
}
#endif
#endif
