/* oracle.h -- CPU FP64 restatement of the Solfec-1.0 contact-solve hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: it is loaded by
 * tests/, by __graft_entry__.smoke() and by the cpu_baseline leg of bench.py,
 * and by nothing else.  The product (libsfgs.so, libsolfec.so) never links it.
 *
 * It restates, in plain serial C, the reference algorithms:
 *   or_ldy.c  -- ldy.c:364-397 (update_V_and_B), ldy.c:626-844 (LOCDYN_Update_Begin),
 *                ldy.c:492-568 (LOCDYN_Insert adjacency), bod.c:216-258 (H, inverse M)
 *   or_dbs.c  -- dbs.c:35-326 (local 3x3 solvers), dbs.c:492-550 (dispatcher)
 *   or_scf.c  -- scf.c:28-174 (cone linearisation / projection)
 *   or_bgs.c  -- bgs.c:913-1069 (serial Gauss-Seidel), mrf.c:35-184 (merit)
 *   or_bod.c  -- bod.c:418-435, 1425-1447 (body velocity update by reactions)
 *   or_pes.c  -- pes.c:33-106 (spring-dashpot contact law of the penalty solver),
 *                ldy.c:69-95 (update_cohesion)
 *   or_nts.c  -- nts.c:937-1379 and ext/krylov/flexgmres.c:82-700 (NEWTON_SOLVER, LOCDYN_ON)
 *   or_dom.c  -- goc.c:50-59, 308-348, 421-455, gjk.c:560-584 (sphere and plane
 *                contacts), dom.c:348-483 (contact update and creation, brute
 *                force), bod.c:1313-1373, 1425-1492 (rigid step), alg.h:1138-1204
 *   or_colour.c -- greedy colouring in priority order (no reference counterpart;
 *                 it fixes the sweep order used by the GPU schedule)
 *
 * Parity status of every function is stated in its source file header.
 * Conventions (as the reference, alg.h:881-953): 3x3 matrices are column-major,
 * local contact frame columns are [t1 t2 n] (dom.c:133-151), index 2 = normal. */
#ifndef SFGS_ORACLE_H
#define SFGS_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* body kinds (bod.h) */
enum { OR_RIG = 0, OR_OBS = 1 };

/* diagonal solvers, numbered as DIAS in dbs.h:27-33 */
enum { OR_DS_PROJECTED_GRADIENT = 0, OR_DS_DE_SAXCE_FENG = 1,
       OR_DS_SEMISMOOTH_NEWTON = 2, OR_DS_PROJECTED_NEWTON = 3 };

/* failure actions GSFAIL bgs.h:39-44 and error codes GSERROR bgs.h:31-37 */
enum { OR_FAILURE_CONTINUE = 0, OR_FAILURE_EXIT = 1, OR_FAILURE_CALLBACK = 2 };
enum { OR_GS_OK = 0, OR_GS_DIVERGED = 1, OR_GS_DIAGONAL_DIVERGED = 2, OR_GS_DIAGONAL_FAILED = 3 };

typedef struct {
  int nb;
  const int *kind;        /* OR_RIG or OR_OBS */
  const double *rot;      /* 9 per body: rotation, column-major */
  const double *x0;       /* 3: referential mass centre */
  const double *jinv;     /* 9: inverse referential inertia */
  const double *minv;     /* 1: inverse mass */
  const double *velo;     /* 6: [W v] free velocity at t+h (referential angular, spatial linear) */
  const double *velo0;    /* 6: [W v] at t */
} OR_BODIES;

typedef struct {
  int nc;
  const int *master, *slave;  /* body indices */
  const double *mpnt, *spnt;  /* 3 each: referential points on master / slave */
  const double *base;         /* 9: local frame, columns t1 t2 n */
  const double *gap;
  const double *friction, *restitution, *cohesion;  /* cohesion already times area */
} OR_CONTACTS;

typedef struct {
  int n;
  const int *rowptr, *col;   /* off-diagonal block CSR (col = contact index) */
  const double *W;           /* 9 per off-diagonal block */
  const double *Wd, *A;      /* 9 per row: W_ii and its inverse */
  const double *B, *V;       /* 3 per row */
  const double *rho;         /* 1 per row */
  const double *mu, *rest, *coh, *gap;
  double free_energy;        /* 0.5 sum B'AB */
  double step;
  int dynamic;
} OR_SYSTEM;

typedef struct {
  double epsilon;
  int maxiter;
  double meritval;
  int failure;
  double diagepsilon;
  int diagmaxiter;
  int diagsolver;
  int reverse;
} OR_GSPARAM;

typedef struct {
  int iters;
  int error;          /* OR_GS_* */
  double rel_error;   /* last sqrt(errup)/sqrt(errlo) */
  double merit;       /* last merit value (0 when merit is off) */
  long diag_fail;     /* local-solver failures seen (all sweeps) */
  long newton_iters;  /* total local iterations (all rows, all sweeps) */
  long row_updates;   /* rows visited */
} OR_REPORT;

/* ---- or_ldy.c ---- */
/* row lengths of the contact adjacency (ldy.c:492-568): rowptr[n+1]; returns nnz */
long OR_Adjacency(const OR_BODIES *bod, const OR_CONTACTS *con, int *rowptr);
/* fill col/bod for every row; order: master's contacts then slave's, ascending index */
void OR_Adjacency_Fill(const OR_BODIES *bod, const OR_CONTACTS *con, const int *rowptr, int *col, int *via);
/* assemble V, B, W_ii, rho, A, off-diagonal W blocks and free energy (ldy.c:364-844) */
int OR_Locdyn_Assemble(const OR_BODIES *bod, const OR_CONTACTS *con, double step,
                       const int *rowptr, const int *col, const int *via,
                       double *W, double *Wd, double *A, double *B, double *V,
                       double *rho, double *free_energy);
/* 3x6 H operator of body b at referential point X in frame base (bod.c:216-228) */
void OR_Operator_H(const OR_BODIES *bod, int b, const double *X, const double *base, double *H);
/* symmetric 3x3 eigenvalues, ascending (stands for LAPACK dsyev 'N') */
void OR_Eig3(const double *W, double *lambda);
/* 3x3 inverse by LU with partial pivoting (stands for dgetrf/dgetri); 0 on success */
int OR_Inverse3(const double *W, double *A);

/* ---- or_dbs.c ---- */
int OR_Diag_Solve(int solver, double diagepsilon, int diagmaxiter, int dynamic, double step,
                  double friction, double restitution, double cohesion, double gap, double rho,
                  const double *W, const double *B, const double *V, double *U, double *R);
/* 3x3 A x = b with partial pivoting (stands for dgesv); A column-major, b overwritten */
int OR_Solve3(double *A, double *b);

/* ---- or_scf.c ---- */
void OR_SCF_Project(double friction, double cohesion, double *S, double *R);
void OR_SCF_Linearize(int dynamic, double step, double friction, double restitution,
                      double cohesion, double gap, const double *V, const double *U,
                      const double *R, double UT, double omega, double *C, double *X, double *Y);

/* ---- or_bgs.c ---- */
double OR_Merit(const OR_SYSTEM *sys, const double *R, double *U, int update_U, double *conmerit);
int OR_GS_Solve(const OR_SYSTEM *sys, const OR_GSPARAM *par, const int *order,
                double *R, double *U, double *rerhist, double *merhist, OR_REPORT *rep);
/* one plain sweep in the given order (no merit, no termination): for timing */
double OR_GS_Sweep(const OR_SYSTEM *sys, const OR_GSPARAM *par, const int *order, int backward,
                   double *R, double *U, double *errlo_out);

/* ---- or_colour.c ---- */
unsigned OR_Hash32(unsigned x);
int OR_Colour(int n, const int *rowptr, const int *col, int *colour);
void OR_Colour_Order(int n, const int *colour, int ncolours, int *order, int *colptr);

/* ---- or_bod.c ---- */
void OR_Body_Reaction(const OR_BODIES *bod, const OR_CONTACTS *con, const double *R, double *force);
void OR_Body_Velocity_Update(const OR_BODIES *bod, const double *force, double step, double *velo);

/* contact detection and the rigid step (or_dom.c).  Sphere contact out[11]:
 * swap (1 when b is the slave), slave point, master point, normal, gap; plane
 * contact out[10]: plane point, sphere point, normal, gap.  OR_Detect takes
 * shape 0 sphere / 1 plane, returns the contact count (writing at most cap):
 * master, slave, the index of the kept old contact (-1 for new ones) and 10
 * geometry values per contact (slave point, master point, normal, gap). */
int OR_Sphere_Contact(const double *ca, double ra, const double *cb, double rb, double *out);
int OR_Plane_Contact(const double *plane, const double *c, double r, double *out);
long long OR_Detect(int nb, const int *shape, const double *rad, const double *plane, const double *x,
                    int nold, const int *master_old, const int *slave_old, long long cap,
                    int *master, int *slave, int *old_index, double *geo);
void OR_Expmap(const double *O, double *R);
void OR_Rig_Step_Begin(int kind, double h, const double *gravity, double mass, double minv, const double *J,
                       const double *I, double *velo, double *velo0, double *R, double *x, double *A);
void OR_Rig_Step_End(int kind, double h, const double *I, const double *force, double *velo, double *R, double *x,
                     double *A);
void OR_Rig_Static_Begin(int kind, double h, const double *gravity, double mass, double minv, const double *J,
                         const double *I, double *velo);
void OR_Rig_Static_End(int kind, double h, double *velo, double *R, double *x);

/* spring-dashpot contact law (or_pes.c); W, B, V of the contact; R in: the
 * carried reaction (its tangential part enters B_N), out: the reaction; coh
 * (NULL: not cohesive) in: the cohesion state, out: 0 when broken; returns 0,
 * or -1 for a singular tangential block */
int OR_Penalty_Contact(int implicit, double step, double gap, double spring, double dashpot, double hpow,
                       double friction, const double *W, const double *B, const double *V, double *U, double *R,
                       double *coh);
/* update_cohesion of LOCDYN_Update_End (or_pes.c): R 3 per contact, coh in/out */
void OR_Update_Cohesion(int n, const double *R, const double *friction, double *coh);

/* NEWTON_SOLVER parameters (nts.h:31-66) and results (or_nts.c) */
typedef struct {
  double meritval;
  int maxiter;
  int linver;        /* 0 GMRES (PQN_GMRES), 1 DIAG (PQN_DIAG) */
  int linmaxiter, maxmatvec;
  double epsilon, delta, theta, omega;
  int gsflag;        /* 1: Gauss-Seidel after an unconverged solve */
  int reldelta;      /* 0 OFF, 1 avgWii, 2 minWii, 3 maxWii */
} OR_NTPARAM;

typedef struct {
  int iters, matvecs, converged, gs_fallback, gs_iters, singular;
  double merit, W_norm;
} OR_NTREPORT;

/* NEWTON_Solve with assembled local dynamics; order: the Gauss-Seidel row
 * order of the fallback (NULL: natural); merhist, mvhist: maxiter entries or NULL */
int OR_Newton_Solve(const OR_SYSTEM *sys, const OR_NTPARAM *par, const int *order, double *R, double *U,
                    double *merhist, int *mvhist, OR_NTREPORT *rep);
/* the flexible GMRES of the Newton solve on a dense row-major n x n operator A
 * with a dense preconditioner P (NULL: identity); x: initial guess in, solution
 * out; returns the iterations (-1: non-finite input) */
int OR_FlexGMRES_Dense(int n, const double *A, const double *P, const double *b, double *x, int kdim, int maxiter,
                       int miniter, double atol, double *rnorm, int *matvecs);

#ifdef __cplusplus
}
#endif
#endif
