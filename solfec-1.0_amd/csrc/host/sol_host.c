/* sol_host.c -- minimal rigid-body domain and the time-stepping loop.
 *
 * Time step (sol.c:746-806):
 *   DOM_Update_Begin   -> SOLFEC_Step_Begin: rigid half step and free velocity
 *                         (BODY_Dynamic_Step_Begin, scheme RIG_NEG, bod.c:1313-1373),
 *                         update of existing contacts (update_contact, dom.c:430-483),
 *                         detection of new ones (overlap_create, dom.c:348-416;
 *                         sphere pairs goc.c:421-455, plane-sphere goc.c:308-348,
 *                         ellipsoids: see ellip_pair and plane_ellip),
 *                         new contacts start with R = U = 0 (mem.c:77,86);
 *   LOCDYN_Update_Begin + GAUSS_SEIDEL_Solve -> SOLFEC_Step_Solve (GPU, sfgs.h);
 *   DOM_Update_End     -> SOLFEC_Step_End: body velocity update by the reactions
 *                         (GPU) and the rigid end-of-step (BODY_Dynamic_Step_End,
 *                         bod.c:1425-1492).
 * Contacts are kept in an array in insertion order; persisting ones keep their
 * slot, removed ones are compacted away, new ones are appended. */
#define _GNU_SOURCE
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#include "solfec.h"
#include "sol_internal.h"
#include "sol_geom.h"

#ifdef _OPENMP
#include <omp.h>
static int omp_get_num_threads_(void) { return omp_get_num_threads(); }
static int omp_get_thread_num_(void) { return omp_get_thread_num(); }
#else
static int omp_get_num_threads_(void) { return 1; }
static int omp_get_thread_num_(void) { return 0; }
#endif

#define GEOMETRIC_EPSILON SOL_GEOMETRIC_EPSILON

static double now_s(void)
{
  struct timespec t;
  clock_gettime(CLOCK_MONOTONIC, &t);
  return (double)t.tv_sec + 1e-9*(double)t.tv_nsec;
}

/* no host array of a scene comes near this many bytes: a larger request comes
 * from a corrupted or overflowed count */
#define SOL_MAX_BYTES ((size_t)1 << 40)

/* resizes the array *pp (any pointer type, passed by address) to count
 * elements of elem bytes.  A negative count, one beyond SOL_MAX_BYTES, or a
 * failed realloc leaves the array as it was, records the reason and returns
 * SOL_ERR_MEM; nothing here ends the process. */
static int grow(char *err, size_t errlen, void *pp, long long count, size_t elem)
{
  void **p = (void**)pp;
  if (count < 0 || (unsigned long long)count > SOL_MAX_BYTES / elem)
  {
    if (err) snprintf(err, errlen, "refused to allocate %lld elements of %zu bytes", count, elem);
    return SOL_ERR_MEM;
  }
  void *q = realloc(*p, count > 0 ? (size_t)count*elem : 1);
  if (!q)
  {
    if (err) snprintf(err, errlen, "out of memory (%lld elements of %zu bytes)", count, elem);
    return SOL_ERR_MEM;
  }
  *p = q;
  return 0;
}
#define GROW(arr, count) grow(sol->err, sizeof(sol->err), &(arr), (long long)(count), sizeof(*(arr)))

static int dev_pull(SOLFEC *sol);

/* a change of the host state: bring the host arrays up to date first, and
 * upload again at the next device step */
static void dev_edit(SOLFEC *sol)
{
  if (sol->devloop) (void)dev_pull(sol);
  sol->dev_valid = 0;
  sol->locdyn_current = 0;
}

/* ---------------- domain construction ---------------- */

SOLFEC* SOLFEC_Create(int dynamic, double step, const char *outpath)
{
  if (!(step > 0.0)) return NULL;
  SOLFEC *sol = calloc(1, sizeof(SOLFEC));
  if (!sol) return NULL;
  sol->dynamic = dynamic ? 1 : 0;
  sol->step = step;
  sol->time = 0.0;
  sol->defmat.s1 = sol->defmat.s2 = -1;
  sol->defmat.model = SOLFEC_SIGNORINI_COULOMB;
  sol->defmat.hpow = 1.0;   /* the other constants 0 (lng.c:2137-2146) */
  sol->device = 0;
  snprintf(sol->outpath, sizeof(sol->outpath), "%s", outpath ? outpath : "");
  return sol;
}

void SOLFEC_Destroy(SOLFEC *sol)
{
  if (!sol) return;
  if (sol->gpu) SFGS_Destroy(sol->gpu);
  free(sol->kind); free(sol->shape); free(sol->rad); free(sol->surf); free(sol->plane);
  free(sol->sca); free(sol->rot0);
  free(sol->rot); free(sol->x); free(sol->x0); free(sol->J); free(sol->jinv); free(sol->minv);
  free(sol->velo); free(sol->velo0); free(sol->aux); free(sol->force); free(sol->mass);
  free(sol->mat);
  free(sol->c_master); free(sol->c_slave); free(sol->c_mpnt); free(sol->c_spnt); free(sol->c_base);
  free(sol->c_gap); free(sol->c_fri); free(sol->c_rest); free(sol->c_coh); free(sol->c_R); free(sol->c_U);
  free(sol->c_point);
  free(sol->grid_head); free(sol->grid_next); free(sol->pair_keys); free(sol->pair_vals);
  free(sol->tmp_velo);
  free(sol->upd_geo); free(sol->upd_ok);
  for (int t = 0; t < sol->cand_nthr; t ++) free(sol->cand[t]);
  free(sol->cand); free(sol->cand_n); free(sol->cand_cap);
  free(sol);
}

const char* SOLFEC_Error_String(int code)
{
  switch (code)
  {
  case SOL_OK: return "OK";
  case SOL_ERR_GAUSS_SEIDEL_DIAGONAL_DIVERGED: return "GAUSS_SEIDEL_SOLVER failed with error code DIAGONAL_DIVERGED";
  case SOL_ERR_GAUSS_SEIDEL_DIVERGED: return "GAUSS_SEIDEL_SOLVER failed with error code DIVERGED";
  case SOL_ERR_DOM_DEPTH: return "Unphysical interpenetration has occurred";
  case SOL_ERR_GPU: return "GPU contact solver failed";
  case SOL_ERR_CALLBACK_EXIT: return "GAUSS_SEIDEL_SOLVER failure callback requested exit";
  case SOL_ERR_ARG: return "invalid argument";
  case SOL_ERR_MEM: return "host memory allocation refused or failed";
  }
  return "unknown error";
}

const char* SOLFEC_Last_Error(SOLFEC *sol) { return sol ? sol->err : "null"; }

void SOLFEC_Gravity(SOLFEC *sol, double gx, double gy, double gz)
{
  sol->gravity[0] = gx; sol->gravity[1] = gy; sol->gravity[2] = gz;
}

int SOLFEC_Surface_Material(SOLFEC *sol, int s1, int s2, double friction, double restitution, double cohesion)
{
  return SOLFEC_Surface_Material_Ex(sol, s1, s2, SOLFEC_SIGNORINI_COULOMB, friction, restitution, cohesion, 0.0, 0.0, 1.0);
}

int SOLFEC_Surface_Material_Ex(SOLFEC *sol, int s1, int s2, int model, double friction, double restitution,
                               double cohesion, double spring, double dashpot, double hpow)
{
  if (!sol || friction < 0.0 || restitution < 0.0 || restitution > 1.0 || cohesion < 0.0 || spring < 0.0 ||
      hpow < 0.0 || (model != SOLFEC_SIGNORINI_COULOMB && model != SOLFEC_SPRING_DASHPOT)) return SOL_ERR_ARG;
  dev_edit(sol);
  MATPAIR m;
  m.s1 = s1 < s2 ? s1 : s2;
  m.s2 = s1 < s2 ? s2 : s1;
  m.model = model;
  m.friction = friction; m.restitution = restitution; m.cohesion = cohesion;
  m.spring = spring; m.dashpot = dashpot; m.hpow = hpow;
  if (s1 < 0 && s2 < 0)
  {
    m.s1 = m.s2 = -1;
    sol->defmat = m;
    return 0;
  }
  if (GROW(sol->mat, sol->nmat + 1LL)) return SOL_ERR_MEM;
  sol->mat[sol->nmat ++] = m;
  return 0;
}

static const MATPAIR* find_material(SOLFEC *sol, int s1, int s2)
{
  int a = s1 < s2 ? s1 : s2, b = s1 < s2 ? s2 : s1;
  for (int k = 0; k < sol->nmat; k ++) if (sol->mat[k].s1 == a && sol->mat[k].s2 == b) return &sol->mat[k];
  return &sol->defmat;
}

static int add_body(SOLFEC *sol)
{
  dev_edit(sol);
  if (sol->nb == sol->bcap)
  {
    long long cap = sol->bcap ? 2LL*sol->bcap : 64;
    if (cap > 0x3fffffff)
    {
      snprintf(sol->err, sizeof(sol->err), "too many bodies");
      return -1;
    }
    /* arrays already grown keep their new size when a later one fails; bcap
     * then stays at the size all of them have */
    if (GROW(sol->kind, cap) || GROW(sol->shape, cap) || GROW(sol->surf, cap) || GROW(sol->rad, cap) ||
        GROW(sol->plane, 6*cap) || GROW(sol->sca, 3*cap) || GROW(sol->rot0, 9*cap) || GROW(sol->rot, 9*cap) ||
        GROW(sol->x, 3*cap) || GROW(sol->x0, 3*cap) || GROW(sol->J, 9*cap) || GROW(sol->jinv, 9*cap) ||
        GROW(sol->minv, cap) || GROW(sol->mass, cap) || GROW(sol->velo, 6*cap) || GROW(sol->velo0, 6*cap) ||
        GROW(sol->aux, 3*cap) || GROW(sol->force, 6*cap) || GROW(sol->tmp_velo, 6*cap)) return -1;
    sol->bcap = (int)cap;
  }
  int b = sol->nb ++;
  static const double I[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
  memcpy(sol->rot + 9*b, I, sizeof(I));
  memset(sol->x + 3*b, 0, 3*sizeof(double));
  memset(sol->x0 + 3*b, 0, 3*sizeof(double));
  memset(sol->J + 9*b, 0, 9*sizeof(double));
  memset(sol->jinv + 9*b, 0, 9*sizeof(double));
  memset(sol->velo + 6*b, 0, 6*sizeof(double));
  memset(sol->velo0 + 6*b, 0, 6*sizeof(double));
  memset(sol->aux + 3*b, 0, 3*sizeof(double));
  memset(sol->force + 6*b, 0, 6*sizeof(double));
  memset(sol->plane + 6*b, 0, 6*sizeof(double));
  memset(sol->sca + 3*b, 0, 3*sizeof(double));
  memcpy(sol->rot0 + 9*b, I, sizeof(I));
  sol->minv[b] = 0.0;
  sol->mass[b] = 0.0;
  sol->rad[b] = 0.0;
  return b;
}

int SOLFEC_Sphere(SOLFEC *sol, const double *centre, double radius, double density, int surface)
{
  if (!sol || !(radius > 0.0) || !(density > 0.0)) return -1;
  int b = add_body(sol);
  if (b < 0) return -1;
  /* sphere mass characteristics (sph.c:167-238): m = rho 4/3 pi r^3, J = 2/5 m r^2 I */
  double m = density * (4.0/3.0) * M_PI * radius*radius*radius, j = 0.4*m*radius*radius;
  sol->kind[b] = SOL_RIG;
  sol->shape[b] = 0;
  sol->surf[b] = surface;
  sol->rad[b] = radius;
  memcpy(sol->x + 3*b, centre, 3*sizeof(double));
  memcpy(sol->x0 + 3*b, centre, 3*sizeof(double));
  sol->J[9*b] = sol->J[9*b + 4] = sol->J[9*b + 8] = j;
  sol->jinv[9*b] = sol->jinv[9*b + 4] = sol->jinv[9*b + 8] = 1.0/j;
  sol->mass[b] = m;
  sol->minv[b] = 1.0/m;
  return b;
}

int SOLFEC_Ellipsoid(SOLFEC *sol, const double *centre, const double *semi_axes, const double *axes,
                     double density, int surface)
{
  if (!sol || !(density > 0.0)) return -1;
  for (int k = 0; k < 3; k ++) if (!(semi_axes[k] > 0.0)) return -1;
  int b = add_body(sol);
  if (b < 0) return -1;
  const double a = semi_axes[0], bb = semi_axes[1], c = semi_axes[2];
  /* ellipsoid mass characteristics (eli.c:170-227): m = rho 4/3 pi abc,
   * principal inertia m/5 (b^2+c^2, a^2+c^2, a^2+b^2), J = R0 diag R0' */
  double m = density * (4.0/3.0) * M_PI * a*bb*c, P[9] = {0}, T[9], J[9];
  P[0] = 0.2*m*(bb*bb + c*c); P[4] = 0.2*m*(a*a + c*c); P[8] = 0.2*m*(a*a + bb*bb);
  double *R0 = sol->rot0 + 9*b;
  if (axes) memcpy(R0, axes, 9*sizeof(double));
  for (int col = 0; col < 3; col ++)        /* T = P R0' */
    for (int r = 0; r < 3; r ++) T[r + 3*col] = P[r + 3*r] * R0[col + 3*r];
  geom_mm(R0, T, J);
  sol->kind[b] = SOL_RIG;
  sol->shape[b] = 2;
  sol->surf[b] = surface;
  memcpy(sol->sca + 3*b, semi_axes, 3*sizeof(double));
  sol->rad[b] = a > bb ? (a > c ? a : c) : (bb > c ? bb : c);
  memcpy(sol->x + 3*b, centre, 3*sizeof(double));
  memcpy(sol->x0 + 3*b, centre, 3*sizeof(double));
  memcpy(sol->J + 9*b, J, sizeof(J));
  /* J^-1 = R0 diag^-1 R0' */
  for (int col = 0; col < 3; col ++)
    for (int r = 0; r < 3; r ++) T[r + 3*col] = R0[col + 3*r] / P[r + 3*r];
  geom_mm(R0, T, sol->jinv + 9*b);
  sol->mass[b] = m;
  sol->minv[b] = 1.0/m;
  return b;
}

int SOLFEC_Plane(SOLFEC *sol, const double *point, const double *normal, int surface)
{
  double len = sqrt(normal[0]*normal[0] + normal[1]*normal[1] + normal[2]*normal[2]);
  if (!sol || !(len > 0.0)) return -1;
  int b = add_body(sol);
  if (b < 0) return -1;
  sol->kind[b] = SOL_OBS;
  sol->shape[b] = 1;
  sol->surf[b] = surface;
  memcpy(sol->x + 3*b, point, 3*sizeof(double));
  memcpy(sol->x0 + 3*b, point, 3*sizeof(double));
  memcpy(sol->plane + 6*b, point, 3*sizeof(double));
  for (int k = 0; k < 3; k ++) sol->plane[6*b + 3 + k] = normal[k]/len;
  return b;
}

int SOLFEC_Set_Velocity(SOLFEC *sol, int body, const double *velo6)
{
  if (!sol || body < 0 || body >= sol->nb) return SOL_ERR_ARG;
  dev_edit(sol);
  memcpy(sol->velo + 6*body, velo6, 6*sizeof(double));
  return 0;
}

/* splitmix64 (S. Vigna) -> uniform double in [0,1) */
static unsigned long long splitmix64(unsigned long long *s)
{
  unsigned long long z = (*s += 0x9e3779b97f4a7c15ULL);
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
  return z ^ (z >> 31);
}
static double urand(unsigned long long *s) { return (double)(splitmix64(s) >> 11) * (1.0/9007199254740992.0); }

int SOLFEC_Lattice_Scene(SOLFEC *sol, int n, double L, double rmin, double rmax, double jitter,
                         double density, unsigned long long seed)
{
  if (!sol || n <= 0 || !(L > 0.0) || !(rmin > 0.0) || rmax < rmin) return SOL_ERR_ARG;
  int side = (int)ceil(cbrt((double)n) - 1e-9);
  double a = L / side;
  unsigned long long s = seed;
  for (int i = 0; i < n; i ++)
  {
    int ix = i % side, iy = (i / side) % side, iz = i / (side*side);
    double c[3] = {(ix + 0.5)*a, (iy + 0.5)*a, (iz + 0.5)*a};
    for (int k = 0; k < 3; k ++) c[k] += (2.0*urand(&s) - 1.0)*jitter;
    double r = rmin + (rmax - rmin)*urand(&s);
    SOLFEC_Sphere(sol, c, r, density, 1);
  }
  const double P[6][6] = {
    {0, 0, 0, 1, 0, 0}, {L, 0, 0, -1, 0, 0},
    {0, 0, 0, 0, 1, 0}, {0, L, 0, 0, -1, 0},
    {0, 0, 0, 0, 0, 1}, {0, 0, L, 0, 0, -1}};
  for (int k = 0; k < 6; k ++) SOLFEC_Plane(sol, P[k], P[k] + 3, 0);
  return 0;
}

/* uniformly random rotation from a uniformly random unit quaternion (K. Shoemake,
 * "Uniform random rotations", Graphics Gems III, 1992) */
static void random_rotation(unsigned long long *s, double *R)
{
  double u1 = urand(s), u2 = 2.0*M_PI*urand(s), u3 = 2.0*M_PI*urand(s);
  double a = sqrt(1.0 - u1), b = sqrt(u1);
  double w = a*sin(u2), x = a*cos(u2), y = b*sin(u3), z = b*cos(u3);
  R[0] = 1.0 - 2.0*(y*y + z*z); R[3] = 2.0*(x*y - w*z);       R[6] = 2.0*(x*z + w*y);
  R[1] = 2.0*(x*y + w*z);       R[4] = 1.0 - 2.0*(x*x + z*z); R[7] = 2.0*(y*z - w*x);
  R[2] = 2.0*(x*z - w*y);       R[5] = 2.0*(y*z + w*x);       R[8] = 1.0 - 2.0*(x*x + y*y);
}

int SOLFEC_Ellipsoid_Lattice_Scene(SOLFEC *sol, int n, double L, double amin, double amax, double jitter,
                                   double density, unsigned long long seed)
{
  if (!sol || n <= 0 || !(L > 0.0) || !(amin > 0.0) || amax < amin) return SOL_ERR_ARG;
  int side = (int)ceil(cbrt((double)n) - 1e-9);
  double a = L / side;
  unsigned long long s = seed;
  for (int i = 0; i < n; i ++)
  {
    int ix = i % side, iy = (i / side) % side, iz = i / (side*side);
    double c[3] = {(ix + 0.5)*a, (iy + 0.5)*a, (iz + 0.5)*a}, sca[3], R[9];
    for (int k = 0; k < 3; k ++) c[k] += (2.0*urand(&s) - 1.0)*jitter;
    for (int k = 0; k < 3; k ++) sca[k] = amin + (amax - amin)*urand(&s);
    random_rotation(&s, R);
    SOLFEC_Ellipsoid(sol, c, sca, R, density, 1);
  }
  const double P[6][6] = {
    {0, 0, 0, 1, 0, 0}, {L, 0, 0, -1, 0, 0},
    {0, 0, 0, 0, 1, 0}, {0, L, 0, 0, -1, 0},
    {0, 0, 0, 0, 0, 1}, {0, 0, L, 0, 0, -1}};
  for (int k = 0; k < 6; k ++) SOLFEC_Plane(sol, P[k], P[k] + 3, 0);
  return 0;
}

/* ---------------- rigid body integration (scheme RIG_NEG) ---------------- */

/* bod.c:1313-1373 for RIG, no point forces: gravity only; QUASI_STATIC
 * analysis: bod.c:1588-1614 */
static void body_step_begin(SOLFEC *sol, int b)
{
  if (sol->kind[b] == SOL_OBS)   /* fixed obstacles */
  {
    if (sol->dynamic) memcpy(sol->velo0 + 6*b, sol->velo + 6*b, 6*sizeof(double));
    return;
  }
  if (!sol->dynamic)
  {
    geom_rig_static_begin(sol->step, sol->gravity, sol->mass[b], sol->minv[b], sol->J + 9*b, sol->jinv + 9*b,
                          sol->velo + 6*b);
    return;
  }
  geom_rig_step_begin(sol->step, sol->gravity, sol->mass[b], sol->minv[b], sol->J + 9*b, sol->jinv + 9*b,
                      sol->velo + 6*b, sol->velo0 + 6*b, sol->rot + 9*b, sol->x + 3*b, sol->aux + 3*b);
}

/* bod.c:1425-1492 for RIG (QUASI_STATIC: bod.c:1616-1645): velocity already
 * updated by the reactions (GPU) */
static void body_step_end(SOLFEC *sol, int b)
{
  if (sol->kind[b] == SOL_OBS) return;
  if (!sol->dynamic)
  {
    geom_rig_static_end(sol->step, sol->velo + 6*b, sol->rot + 9*b, sol->x + 3*b);
    return;
  }
  geom_rig_step_end(sol->step, sol->jinv + 9*b, sol->force + 6*b, sol->velo + 6*b, sol->rot + 9*b, sol->x + 3*b,
                    sol->aux + 3*b);
}

/* ---------------- contacts ---------------- */

/* room for n contacts; a count outside [0, 2^30) (a corrupted size) is
 * refused with SOL_ERR_MEM, as is a failed allocation */
static int ensure_contacts(SOLFEC *sol, long long n)
{
  if (n < 0 || n > 0x3fffffff)
  {
    snprintf(sol->err, sizeof(sol->err), "refused a contact count of %lld", n);
    return SOL_ERR_MEM;
  }
  if (n <= sol->ccap) return 0;
  long long cap = sol->ccap ? sol->ccap : 1024;
  while (cap < n) cap *= 2;
  if (GROW(sol->c_master, cap) || GROW(sol->c_slave, cap) || GROW(sol->c_mpnt, 3*cap) || GROW(sol->c_spnt, 3*cap) ||
      GROW(sol->c_point, 3*cap) || GROW(sol->c_base, 9*cap) || GROW(sol->c_gap, cap) || GROW(sol->c_fri, cap) ||
      GROW(sol->c_rest, cap) || GROW(sol->c_coh, cap) || GROW(sol->c_R, 3*cap) || GROW(sol->c_U, 3*cap))
    return SOL_ERR_MEM;
  sol->ccap = (int)cap;
  return 0;
}

/* referential image of a spatial point (BODY_Ref_Point): X = X0 + R'(x - xc) */
static void ref_point(SOLFEC *sol, int b, const double *p, double *X)
{
  geom_ref_point(sol->rot + 9*b, sol->x + 3*b, sol->x0 + 3*b, p, X);
}

static void convex_frame(SOLFEC *sol, int b, double *Q, double *S)
{
  geom_frame(sol->shape[b], sol->rot + 9*b, sol->rot0 + 9*b, sol->sca + 3*b, sol->rad[b], Q, S);
}

/* convex pair as (slave a, master b): the sphere routine for sphere pairs, the
 * ellipsoid one otherwise */
static int convex_pair(SOLFEC *sol, int a, int b, double *spnt, double *mpnt, double *normal, double *gap)
{
  if (sol->shape[a] == 0 && sol->shape[b] == 0)
    return geom_sphere_pair(sol->x + 3*a, sol->rad[a], sol->x + 3*b, sol->rad[b], spnt, mpnt, normal, gap);
  double Qa[9], Sa[3], Qb[9], Sb[3];
  convex_frame(sol, a, Qa, Sa);
  convex_frame(sol, b, Qb, Sb);
  return geom_ellip_pair(sol->x + 3*a, Qa, Sa, sol->x + 3*b, Qb, Sb, spnt, mpnt, normal, gap);
}

/* plane (slave) and sphere or ellipsoid (master) */
static int plane_convex(SOLFEC *sol, int pl, int s, double *spnt, double *mpnt, double *normal, double *gap)
{
  if (sol->shape[s] == 2)
  {
    double Q[9], S[3];
    convex_frame(sol, s, Q, S);
    return geom_plane_ellip(sol->plane + 6*pl, sol->x + 3*s, Q, S, spnt, mpnt, normal, gap);
  }
  return geom_plane_sphere(sol->plane + 6*pl, sol->x + 3*s, sol->rad[s], spnt, mpnt, normal, gap);
}

/* open-addressing map from body pair to contact slot */
static int map_build(SOLFEC *sol)
{
  size_t cap = 16;
  while (cap < 2*(size_t)sol->nc + 16) cap *= 2;
  if (cap > sol->pair_cap)
  {
    if (GROW(sol->pair_keys, cap) || GROW(sol->pair_vals, cap)) return SOL_ERR_MEM;
    sol->pair_cap = cap;
  }
  sol->pair_mask = cap - 1;
  for (size_t k = 0; k <= sol->pair_mask; k ++) sol->pair_keys[k] = ~0ULL;
  for (int i = 0; i < sol->nc; i ++)
  {
    unsigned long long key = geom_pair_key(sol->c_master[i], sol->c_slave[i]);
    size_t h = geom_mix64(key) & sol->pair_mask;
    while (sol->pair_keys[h] != ~0ULL) h = (h + 1) & sol->pair_mask;
    sol->pair_keys[h] = key;
    sol->pair_vals[h] = i;
  }
  return 0;
}
static int map_find(SOLFEC *sol, int a, int b)
{
  unsigned long long key = geom_pair_key(a, b);
  size_t h = geom_mix64(key) & sol->pair_mask;
  while (sol->pair_keys[h] != ~0ULL)
  {
    if (sol->pair_keys[h] == key) return sol->pair_vals[h];
    h = (h + 1) & sol->pair_mask;
  }
  return -1;
}

/* contact data from its geometry; the cohesion state is set apart (new_contact) */
static void set_contact(SOLFEC *sol, int i, int master, int slave, const double *mp, const double *sp,
                        const double *normal, double gap)
{
  sol->c_master[i] = master;
  sol->c_slave[i] = slave;
  memcpy(sol->c_point + 3*i, mp, 3*sizeof(double));
  ref_point(sol, master, mp, sol->c_mpnt + 3*i);
  ref_point(sol, slave, sp, sol->c_spnt + 3*i);
  geom_localbase(normal, sol->c_base + 9*i);
  sol->c_gap[i] = gap;
  const MATPAIR *m = find_material(sol, sol->surf[master], sol->surf[slave]);
  sol->c_fri[i] = m->friction;
  sol->c_rest[i] = m->restitution;
}

/* a cohesive contact keeps its material points (dom.c:452-458): the spatial
 * point is the master's original point in the current configuration, and the
 * frame and gap are those of the geometry update (the previous ones when the
 * surfaces have separated) */
static void set_cohesive(SOLFEC *sol, int i, const double *normal, double gap)
{
  const int m = sol->c_master[i];
  geom_cur_point(sol->rot + 9*m, sol->x + 3*m, sol->x0 + 3*m, sol->c_mpnt + 3*i, sol->c_point + 3*i);
  geom_localbase(normal, sol->c_base + 9*i);
  sol->c_gap[i] = gap;
  const MATPAIR *mp = find_material(sol, sol->surf[m], sol->surf[sol->c_slave[i]]);
  sol->c_fri[i] = mp->friction;
  sol->c_rest[i] = mp->restitution;
}

/* a contact found by the detection: zero reactions and the cohesion state of
 * SURFACE_MATERIAL_Transfer (sps.c:264-282), the material's cohesion for a
 * contact created at time 0 and none later; area = 1 for these contacts
 * (goc.c:450), so the state is the cohesive force itself */
static void new_contact(SOLFEC *sol, int i, int master, int slave, const double *mp, const double *sp,
                        const double *normal, double gap)
{
  set_contact(sol, i, master, slave, mp, sp, normal, gap);
  const MATPAIR *m = find_material(sol, sol->surf[master], sol->surf[slave]);
  sol->c_coh[i] = sol->time == 0.0 && m->cohesion > 0.0 ? m->cohesion : 0.0;
  memset(sol->c_R + 3*i, 0, 3*sizeof(double));
  memset(sol->c_U + 3*i, 0, 3*sizeof(double));
}

/* update existing contacts (keep slot, R, U; remove separated), then detect
 * new ones with a uniform grid over sphere centres.  New convex pairs are
 * appended in ascending order of (a, b), a < b the body indices, then new
 * plane contacts in ascending order of (plane, body): an order fixed by the
 * pairs alone, which the device loop reproduces. */
static int update_and_detect(SOLFEC *sol)
{
  double t0 = now_s();
  int nc = sol->nc, kept = 0;
  double sp[3], mp[3], nrm[3], g;

  /* geometry of every existing contact (independent per contact) */
  if ((size_t)nc > sol->upd_cap)
  {
    if (GROW(sol->upd_geo, 10LL*nc) || GROW(sol->upd_ok, nc)) return SOL_ERR_MEM;
    sol->upd_cap = (size_t)nc;
  }
  #pragma omp parallel for schedule(static)
  for (int i = 0; i < nc; i ++)
  {
    int m = sol->c_master[i], s = sol->c_slave[i];
    double *G = sol->upd_geo + 10*(size_t)i;
    if (sol->shape[s] == 1) sol->upd_ok[i] = plane_convex(sol, s, m, G, G + 3, G + 6, G + 9);
    else sol->upd_ok[i] = convex_pair(sol, s, m, G, G + 3, G + 6, G + 9);
  }
  /* compaction keeps the insertion order; upd_ok receives the old index of each
   * kept slot.  Separated contacts are removed unless cohesive (dom.c:452-480),
   * and a separated cohesive contact keeps its previous normal and gap. */
  for (int i = 0; i < nc; i ++)
  {
    const int cohesive = sol->c_coh[i] > 0.0;
    if (!sol->upd_ok[i] && !cohesive) continue;
    if (!sol->upd_ok[i])
    {
      double *G = sol->upd_geo + 10*(size_t)i;
      memcpy(G + 6, sol->c_base + 9*i + 6, 3*sizeof(double));
      G[9] = sol->c_gap[i];
    }
    if (kept != i)
    {
      sol->c_master[kept] = sol->c_master[i];
      sol->c_slave[kept] = sol->c_slave[i];
      memcpy(sol->c_R + 3*kept, sol->c_R + 3*i, 3*sizeof(double));
      memcpy(sol->c_U + 3*kept, sol->c_U + 3*i, 3*sizeof(double));
      memcpy(sol->c_mpnt + 3*kept, sol->c_mpnt + 3*i, 3*sizeof(double));
      memcpy(sol->c_spnt + 3*kept, sol->c_spnt + 3*i, 3*sizeof(double));
      sol->c_coh[kept] = sol->c_coh[i];
    }
    sol->upd_ok[kept ++] = i;
  }
  #pragma omp parallel for schedule(static)
  for (int k = 0; k < kept; k ++)
  {
    const double *G = sol->upd_geo + 10*(size_t)sol->upd_ok[k];
    if (sol->c_coh[k] > 0.0) set_cohesive(sol, k, G + 6, G[9]);
    else set_contact(sol, k, sol->c_master[k], sol->c_slave[k], G + 3, G, G + 6, G[9]);
  }
  sol->nc = kept;
  double t1 = now_s();
  sol->t_conupd += t1 - t0;

  /* broad phase: spheres hashed into cells of size 2 rmax + eps */
  int nb = sol->nb;
  double rmax = 0.0, lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
  for (int b = 0; b < nb; b ++)
    if (sol->shape[b] != 1)   /* spheres and ellipsoids, by their bounding spheres */
    {
      if (sol->rad[b] > rmax) rmax = sol->rad[b];
      for (int k = 0; k < 3; k ++)
      {
        if (sol->x[3*b + k] < lo[k]) lo[k] = sol->x[3*b + k];
        if (sol->x[3*b + k] > hi[k]) hi[k] = sol->x[3*b + k];
      }
    }
  if (map_build(sol)) return SOL_ERR_MEM;
  if (rmax > 0.0)
  {
    /* at most 4 cells per body (very sparse scenes get a coarser grid); the
     * count is formed without integer overflow however far a body has gone */
    double cell = 2.0*rmax + 2.0*GEOMETRIC_EPSILON;
    int dim[3];
    long long ncell = 1;
    if (geom_grid_dims(lo, hi, 4LL*nb + 1024, &cell, dim, &ncell))
    {
      snprintf(sol->err, sizeof(sol->err), "body positions are not finite (the time integration diverged)");
      return SOL_ERR_ARG;
    }
    if ((size_t)ncell > sol->grid_cap)
    {
      if (GROW(sol->grid_head, ncell)) return SOL_ERR_MEM;
      sol->grid_cap = (size_t)ncell;
    }
    if ((size_t)nb > sol->gnext_cap)
    {
      if (GROW(sol->grid_next, nb)) return SOL_ERR_MEM;
      sol->gnext_cap = (size_t)nb;
    }
    for (long long c = 0; c < ncell; c ++) sol->grid_head[c] = -1;
    /* insert in descending body order so that cell lists run in ascending order */
    for (int b = nb - 1; b >= 0; b --)
    {
      if (sol->shape[b] == 1) continue;
      int ci[3];
      for (int k = 0; k < 3; k ++) ci[k] = geom_cell_coord((sol->x[3*b + k] - lo[k])/cell, dim[k]);
      long long c = ((long long)ci[2]*dim[1] + ci[1])*dim[0] + ci[0];
      sol->grid_next[b] = sol->grid_head[c];
      sol->grid_head[c] = b;
    }
    /* candidate new pairs per thread, merged in ascending order of a */
    int nthr = 1;
    #pragma omp parallel
    {
      #pragma omp single
      nthr = omp_get_num_threads_();
    }
    if (sol->cand_nthr < nthr)
    {
      if (GROW(sol->cand, nthr) || GROW(sol->cand_n, nthr) || GROW(sol->cand_cap, nthr)) return SOL_ERR_MEM;
      for (int t = sol->cand_nthr; t < nthr; t ++) { sol->cand[t] = NULL; sol->cand_cap[t] = 0; }
      sol->cand_nthr = nthr;
    }
    int cand_failed = 0;   /* a candidate buffer that could not grow: the step fails after the region */
    #pragma omp parallel reduction(|:cand_failed)
    {
      int t = omp_get_thread_num_();
      int chunk = (nb + nthr - 1) / nthr, a0 = t*chunk, a1 = a0 + chunk < nb ? a0 + chunk : nb;
      sol->cand_n[t] = 0;
      for (int a = a0; a < a1; a ++)
      {
        if (sol->shape[a] == 1) continue;
        int ci[3], first = sol->cand_n[t];
        for (int k = 0; k < 3; k ++) ci[k] = geom_cell_coord((sol->x[3*a + k] - lo[k])/cell, dim[k]);
        for (int dz = -1; dz <= 1; dz ++)
          for (int dy = -1; dy <= 1; dy ++)
            for (int dx = -1; dx <= 1; dx ++)
            {
              int cx = ci[0] + dx, cy = ci[1] + dy, cz = ci[2] + dz;
              if (cx < 0 || cy < 0 || cz < 0 || cx >= dim[0] || cy >= dim[1] || cz >= dim[2]) continue;
              for (int b = sol->grid_head[((long long)cz*dim[1] + cy)*dim[0] + cx]; b >= 0; b = sol->grid_next[b])
              {
                if (b <= a) continue;
                double dx_ = sol->x[3*b] - sol->x[3*a], dy_ = sol->x[3*b + 1] - sol->x[3*a + 1],
                       dz_ = sol->x[3*b + 2] - sol->x[3*a + 2], e_ = sol->rad[a] + sol->rad[b] + GEOMETRIC_EPSILON;
                if (dx_*dx_ + dy_*dy_ + dz_*dz_ >= e_*e_) continue;
                if (map_find(sol, a, b) >= 0) continue;
                if (sol->cand_n[t] + 2 > sol->cand_cap[t])
                {
                  long long cap = sol->cand_cap[t] ? 2LL*sol->cand_cap[t] : 1024;
                  if (cap > 0x3fffffff || grow(NULL, 0, &sol->cand[t], cap, sizeof(int))) { cand_failed = 1; break; }
                  sol->cand_cap[t] = (int)cap;
                }
                sol->cand[t][sol->cand_n[t] ++] = a;
                sol->cand[t][sol->cand_n[t] ++] = b;
              }
            }
        /* the pairs of a in ascending order of b, whatever the cell layout */
        int *cb = sol->cand[t];
        for (int q = first + 3; q < sol->cand_n[t]; q += 2)
        {
          int v = cb[q], r = q;
          while (r > first + 1 && cb[r - 2] > v) { cb[r] = cb[r - 2]; r -= 2; }
          cb[r] = v;
        }
      }
    }
    if (cand_failed)
    {
      snprintf(sol->err, sizeof(sol->err), "out of memory for the candidate contact pairs");
      return SOL_ERR_MEM;
    }
    for (int t = 0; t < nthr; t ++)
      for (int q = 0; q < sol->cand_n[t]; q += 2)
      {
        int a = sol->cand[t][q], b = sol->cand[t][q + 1];
        /* orientation rule of detect_sphere_sphere (goc.c:421-455): the body with the
         * lexicographically larger centre is the master */
        int slave = a, master = b;
        if (geom_pntcmp(sol->x + 3*a, sol->x + 3*b) > 0) { slave = b; master = a; }
        if (!convex_pair(sol, slave, master, sp, mp, nrm, &g)) continue;
        if (ensure_contacts(sol, sol->nc + 1LL)) return SOL_ERR_MEM;
        new_contact(sol, sol->nc ++, master, slave, mp, sp, nrm, g);
      }
  }
  /* planes against all spheres */
  for (int p = 0; p < nb; p ++)
  {
    if (sol->shape[p] != 1) continue;
    for (int s = 0; s < nb; s ++)
    {
      if (sol->shape[s] == 1) continue;
      const double *P = sol->plane + 6*p, *c = sol->x + 3*s;
      double dist = (c[0] - P[0])*P[3] + (c[1] - P[1])*P[4] + (c[2] - P[2])*P[5];
      if (dist >= sol->rad[s] + GEOMETRIC_EPSILON) continue;
      if (map_find(sol, p, s) >= 0) continue;
      if (!plane_convex(sol, p, s, sp, mp, nrm, &g)) continue;
      if (ensure_contacts(sol, sol->nc + 1LL)) return SOL_ERR_MEM;
      new_contact(sol, sol->nc ++, s, p, mp, sp, nrm, g);
    }
  }
  sol->t_condet += now_s() - t1;
  return 0;
}

/* ---------------- GPU context ---------------- */

sfgs_ctx* SOLFEC_Gpu(SOLFEC *sol)
{
  if (!sol->gpu)
  {
    sol->gpu = SFGS_Create(sol->device);
    if (!sol->gpu) snprintf(sol->err, sizeof(sol->err), "no HIP device available for the contact solver");
  }
  return sol->gpu;
}

/* ---------------- device loop ---------------- */

/* bodies, shapes, materials and contacts to the device */
static int dev_push(SOLFEC *sol, sfgs_ctx *g)
{
  int rc = SFGS_Bodies_Upload(g, sol->nb, sol->kind, sol->rot, sol->x0, sol->jinv, sol->minv, sol->velo, sol->velo0);
  if (!rc) rc = SFGS_Contacts_Upload(g, sol->nc, sol->c_master, sol->c_slave, sol->c_mpnt, sol->c_spnt, sol->c_base,
                                     sol->c_gap, sol->c_fri, sol->c_rest, sol->c_coh, sol->c_R, sol->c_U);
  if (!rc)
  {
    SFGS_SHAPES sh = {sol->shape, sol->surf, sol->rad, sol->plane, sol->sca, sol->rot0, sol->x, sol->J, sol->mass,
                      sol->aux};
    SFGS_MATERIAL *mat = NULL, def;
    if (GROW(mat, sol->nmat + 1LL)) return SOL_ERR_MEM;
    for (int k = 0; k < sol->nmat; k ++)
    {
      mat[k].s1 = sol->mat[k].s1; mat[k].s2 = sol->mat[k].s2;
      mat[k].friction = sol->mat[k].friction; mat[k].restitution = sol->mat[k].restitution;
      mat[k].cohesion = sol->mat[k].cohesion;
      mat[k].spring = sol->mat[k].spring; mat[k].dashpot = sol->mat[k].dashpot; mat[k].hpow = sol->mat[k].hpow;
    }
    def.s1 = def.s2 = -1;
    def.friction = sol->defmat.friction; def.restitution = sol->defmat.restitution; def.cohesion = sol->defmat.cohesion;
    def.spring = sol->defmat.spring; def.dashpot = sol->defmat.dashpot; def.hpow = sol->defmat.hpow;
    rc = SFGS_Dom_Upload(g, &sh, sol->nmat, mat, &def, sol->c_point);
    if (!rc) rc = SFGS_Dom_Set_Analysis(g, sol->dynamic);
    free(mat);
  }
  if (rc)
  {
    snprintf(sol->err, sizeof(sol->err), "device loop upload failed: %s (%s)", SFGS_Error_String(rc), SFGS_Last_Error(g));
    return SOL_ERR_GPU;
  }
  sol->dev_valid = 1;
  sol->host_stale = 0;
  return 0;
}

/* the device's bodies and contacts into the host arrays, when they are behind */
static int dev_pull(SOLFEC *sol)
{
  if (!sol->devloop || !sol->host_stale || !sol->gpu) return 0;
  sfgs_ctx *g = sol->gpu;
  int n = 0;
  int rc = SFGS_Dom_Get_Bodies(g, sol->rot, sol->x, sol->velo, sol->velo0, sol->aux, sol->force);
  if (!rc) rc = SFGS_Get_Sizes(g, &n, NULL, NULL);
  if (!rc)
  {
    rc = ensure_contacts(sol, n > 0 ? n : 1);
    if (rc) return rc;
    sol->nc = n;
    rc = SFGS_Dom_Get_Contacts(g, sol->c_master, sol->c_slave, sol->c_point, sol->c_mpnt, sol->c_spnt, sol->c_base,
                               sol->c_gap, sol->c_fri, sol->c_rest, sol->c_coh, sol->c_R, sol->c_U);
  }
  if (rc)
  {
    snprintf(sol->err, sizeof(sol->err), "device loop download failed: %s (%s)", SFGS_Error_String(rc), SFGS_Last_Error(g));
    return SOL_ERR_GPU;
  }
  sol->host_stale = 0;
  return 0;
}

int sol_dev_pull(SOLFEC *sol) { return dev_pull(sol); }
void sol_dev_edit(SOLFEC *sol) { dev_edit(sol); }
int sol_ensure_contacts(SOLFEC *sol, long long n) { return ensure_contacts(sol, n); }

int SOLFEC_Device_Loop(SOLFEC *sol, int on)
{
  if (!sol) return SOL_ERR_ARG;
  if (sol->devloop && !on)
  {
    int rc = dev_pull(sol);
    if (rc) return rc;
  }
  sol->devloop = on ? 1 : 0;
  sol->dev_valid = 0;
  sol->host_stale = 0;
  return 0;
}

static int dev_step_begin(SOLFEC *sol)
{
  sfgs_ctx *g = SOLFEC_Gpu(sol);
  if (!g) return SOL_ERR_GPU;
  double t0 = now_s();
  if (!sol->dev_valid)
  {
    int rc = dev_push(sol, g);
    if (rc) return rc;
  }
  int nc = 0, rc = SFGS_Dom_Set_Time(g, sol->time);
  if (!rc) rc = SFGS_Dom_Update_Begin(g, sol->step, sol->gravity, &nc);
  if (rc)
  {
    snprintf(sol->err, sizeof(sol->err), "device step begin failed: %s (%s)", SFGS_Error_String(rc), SFGS_Last_Error(g));
    return SOL_ERR_GPU;
  }
  sol->nc = nc;
  sol->host_stale = 1;
  sol->t_condet += now_s() - t0;   /* the half step, contact update and detection in one call */
  return 0;
}

static int dev_step_end(SOLFEC *sol)
{
  sfgs_ctx *g = SOLFEC_Gpu(sol);
  if (!g) return SOL_ERR_GPU;
  double t0 = now_s();
  int rc = SFGS_Dom_Update_End(g, sol->step);
  if (rc)
  {
    snprintf(sol->err, sizeof(sol->err), "device step end failed: %s (%s)", SFGS_Error_String(rc), SFGS_Last_Error(g));
    return SOL_ERR_GPU;
  }
  sol->time += sol->step;
  sol->host_stale = 1;
  sol->t_timint += now_s() - t0;
  return 0;
}

/* ---------------- time stepping ---------------- */

/* stopfile (sol.c:344-367): a file named STOP in the output directory */
static int stop_file(SOLFEC *sol)
{
  char path[sizeof(sol->outpath) + 8];
  snprintf(path, sizeof(path), "%s/STOP", sol->outpath);
  FILE *f = fopen(path, "r");
  if (!f) return 0;
  fclose(f);
  return 1;
}

int SOLFEC_Step_Begin(SOLFEC *sol)
{
  sol->locdyn_current = 0;
  if (sol->devloop) return dev_step_begin(sol);
  double t0 = now_s();
  #pragma omp parallel for schedule(static)
  for (int b = 0; b < sol->nb; b ++) body_step_begin(sol, b);
  sol->t_timint += now_s() - t0;
  int rc = ensure_contacts(sol, 1);
  return rc ? rc : update_and_detect(sol);
}

int SOLFEC_Step_End(SOLFEC *sol)
{
  if (sol->devloop) return dev_step_end(sol);
  double t0 = now_s();
  #pragma omp parallel for schedule(static)
  for (int b = 0; b < sol->nb; b ++) body_step_end(sol, b);
  sol->time += sol->step;
  sol->t_timint += now_s() - t0;
  return 0;
}

int SOLFEC_Capture(SOLFEC *sol, float *buffer, int cap)
{
  sol->capture = buffer;
  sol->capture_cap = buffer ? cap : 0;
  sol->captured = 0;
  return 0;
}

int SOLFEC_Captured(SOLFEC *sol) { return sol->captured; }

double SOLFEC_Time(SOLFEC *sol) { return sol->time; }
int SOLFEC_Nbod(SOLFEC *sol) { return sol->nb; }
int SOLFEC_Ncon(SOLFEC *sol) { return sol->nc; }

double SOLFEC_Timing(SOLFEC *sol, const char *label)
{
  if (!strcmp(label, "TIMINT")) return sol->t_timint;
  if (!strcmp(label, "CONUPD")) return sol->t_conupd;
  if (!strcmp(label, "CONDET")) return sol->t_condet;
  if (!strcmp(label, "LOCDYN")) return sol->t_locdyn;
  if (!strcmp(label, "CONSOL")) return sol->t_consol;
  return 0.0;
}

int SOLFEC_Get_Bodies(SOLFEC *sol, int *kind, double *rot, double *x, double *x0, double *jinv,
                      double *minv, double *velo, double *velo0, double *radius)
{
  int rc = dev_pull(sol);
  if (rc) return rc;
  int nb = sol->nb;
  if (kind) memcpy(kind, sol->kind, sizeof(int)*(size_t)nb);
  if (rot) memcpy(rot, sol->rot, sizeof(double)*9*(size_t)nb);
  if (x) memcpy(x, sol->x, sizeof(double)*3*(size_t)nb);
  if (x0) memcpy(x0, sol->x0, sizeof(double)*3*(size_t)nb);
  if (jinv) memcpy(jinv, sol->jinv, sizeof(double)*9*(size_t)nb);
  if (minv) memcpy(minv, sol->minv, sizeof(double)*(size_t)nb);
  if (velo) memcpy(velo, sol->velo, sizeof(double)*6*(size_t)nb);
  if (velo0) memcpy(velo0, sol->velo0, sizeof(double)*6*(size_t)nb);
  if (radius) memcpy(radius, sol->rad, sizeof(double)*(size_t)nb);
  return 0;
}

int SOLFEC_Get_Contacts(SOLFEC *sol, int *master, int *slave, double *mpnt, double *spnt, double *base,
                        double *gap, double *friction, double *restitution, double *cohesion,
                        double *R, double *U)
{
  int rc = dev_pull(sol);
  if (rc) return rc;
  size_t n = (size_t)sol->nc;
  if (master) memcpy(master, sol->c_master, sizeof(int)*n);
  if (slave) memcpy(slave, sol->c_slave, sizeof(int)*n);
  if (mpnt) memcpy(mpnt, sol->c_mpnt, sizeof(double)*3*n);
  if (spnt) memcpy(spnt, sol->c_spnt, sizeof(double)*3*n);
  if (base) memcpy(base, sol->c_base, sizeof(double)*9*n);
  if (gap) memcpy(gap, sol->c_gap, sizeof(double)*n);
  if (friction) memcpy(friction, sol->c_fri, sizeof(double)*n);
  if (restitution) memcpy(restitution, sol->c_rest, sizeof(double)*n);
  if (cohesion) memcpy(cohesion, sol->c_coh, sizeof(double)*n);
  if (R) memcpy(R, sol->c_R, sizeof(double)*3*n);
  if (U) memcpy(U, sol->c_U, sizeof(double)*3*n);
  return 0;
}

int SOLFEC_Set_RU(SOLFEC *sol, const double *R, const double *U)
{
  dev_edit(sol);
  size_t n = (size_t)sol->nc;
  if (R) memcpy(sol->c_R, R, sizeof(double)*3*n);
  if (U) memcpy(sol->c_U, U, sizeof(double)*3*n);
  return 0;
}

/* a SPRING_DASHPOT material in the domain (the Gauss-Seidel sweep here models
 * SIGNORINI_COULOMB contacts only) */
static int has_spring_dashpot(SOLFEC *sol)
{
  if (sol->defmat.model == SOLFEC_SPRING_DASHPOT) return 1;
  for (int k = 0; k < sol->nmat; k ++) if (sol->mat[k].model == SOLFEC_SPRING_DASHPOT) return 1;
  return 0;
}

/* a cohesive material in the domain: the contacts created at time 0 may carry
 * cohesion states, which the solves update */
static int has_cohesion(SOLFEC *sol)
{
  if (sol->defmat.cohesion > 0.0) return 1;
  for (int k = 0; k < sol->nmat; k ++) if (sol->mat[k].cohesion > 0.0) return 1;
  return 0;
}

/* the assembly of a step: in the host loop bodies and contacts go to the
 * device first; gs (may be NULL) asks for the learned warm start; upmin: the
 * diagonal blocks only (UPMIN, ldy.c:59, 644: NEWTON_SOLVER with locdyn OFF) */
static int step_assemble(SOLFEC *sol, sfgs_ctx *g, GAUSS_SEIDEL *gs, int upmin)
{
  double t0 = now_s();
  int rc = 0;
  if (!sol->devloop)
  {
    rc = SFGS_Bodies_Upload(g, sol->nb, sol->kind, sol->rot, sol->x0, sol->jinv, sol->minv, sol->velo, sol->velo0);
    if (!rc) rc = SFGS_Contacts_Upload(g, sol->nc, sol->c_master, sol->c_slave, sol->c_mpnt, sol->c_spnt, sol->c_base,
                                       sol->c_gap, sol->c_fri, sol->c_rest, sol->c_coh, sol->c_R, sol->c_U);
  }
  if (!rc) rc = upmin ? SFGS_Locdyn_Update_Min(g, sol->step, sol->dynamic)
                     : SFGS_Locdyn_Update_Begin(g, sol->step, sol->dynamic);
  sol->t_locdyn += now_s() - t0;
  sol->locdyn_current = rc == 0;
  if (!rc && sol->capture)
  {
    sol->captured = sol->nc <= sol->capture_cap ? sol->nc : 0;
    if (sol->captured) rc = SFGS_Features(g, sol->capture);
  }
  if (!rc && gs && gs->warmstart) rc = SFGS_Warm_Start(g);
  if (rc)
  {
    snprintf(sol->err, sizeof(sol->err), "LOCDYN on GPU failed: %s (%s)", SFGS_Error_String(rc), SFGS_Last_Error(g));
    return SOL_ERR_GPU;
  }
  return 0;
}

/* update_cohesion after a NEWTON step in a domain with SPRING_DASHPOT
 * materials (their contacts are solved as SIGNORINI_COULOMB ones, nts.c:1072-1081):
 * the states of the SPRING_DASHPOT contacts stay as they were (ldy.c:78).  The
 * device updates every contact; the kept states are written back.  Returns an
 * SFGS error code, or SOL_ERR_MEM */
static int cohesion_keep_spring_dashpot(SOLFEC *sol, sfgs_ctx *g)
{
  const size_t n = (size_t)sol->nc;
  if (n == 0) return SFGS_Update_Cohesion(g);
  int *ms = malloc(2*n*sizeof(int));
  double *before = malloc(n*sizeof(double)), *after = malloc(n*sizeof(double));
  int rc = 0;
  if (!ms || !before || !after)
  {
    snprintf(sol->err, sizeof(sol->err), "out of host memory for %zu contacts", n);
    rc = SOL_ERR_MEM;
  }
  else
  {
    /* the contact list is the device's in the device loop, the host's otherwise */
    int ndev = -1;
    rc = SFGS_Get_Sizes(g, &ndev, NULL, NULL);
    if (!rc && ndev != (int)n)
    {
      snprintf(sol->err, sizeof(sol->err), "device holds %d contacts, the domain %zu", ndev, n);
      rc = SFGS_E_STATE;
    }
    if (!rc && sol->devloop) rc = SFGS_Dom_Get_Contacts(g, ms, ms + n, NULL, NULL, NULL, NULL, NULL, NULL, NULL,
                                                        before, NULL, NULL);
    else if (!rc)
    {
      memcpy(ms, sol->c_master, n*sizeof(int));
      memcpy(ms + n, sol->c_slave, n*sizeof(int));
      memcpy(before, sol->c_coh, n*sizeof(double));
    }
    if (!rc) rc = SFGS_Update_Cohesion(g);
    if (!rc) rc = SFGS_Get_Cohesion(g, after);
    if (!rc)
    {
      for (size_t i = 0; i < n; i ++)
      {
        const MATPAIR *m = find_material(sol, sol->surf[ms[i]], sol->surf[ms[n + i]]);
        if (m->model == SOLFEC_SPRING_DASHPOT) after[i] = before[i];
      }
      rc = SFGS_Set_Cohesion(g, after);
    }
  }
  free(ms); free(before); free(after);
  return rc;
}

/* after the solve, in the host loop: R, U into the contact list and the
 * velocity update (the device loop does both in SFGS_Dom_Update_End) */
static int step_finish(SOLFEC *sol, sfgs_ctx *g)
{
  int rc = 0;
  if (has_cohesion(sol))
  {
    /* LOCDYN_Update_End (ldy.c:879-889): the cohesion states after the solve */
    if (has_spring_dashpot(sol)) rc = cohesion_keep_spring_dashpot(sol, g);
    else rc = SFGS_Update_Cohesion(g);
    if (!rc && !sol->devloop) rc = SFGS_Get_Cohesion(g, sol->c_coh);
    if (rc == SOL_ERR_MEM) return rc;
    if (rc)
    {
      snprintf(sol->err, sizeof(sol->err), "GPU cohesion update failed: %s (%s)", SFGS_Error_String(rc), SFGS_Last_Error(g));
      return SOL_ERR_GPU;
    }
  }
  if (sol->devloop) return 0;
  rc = SFGS_Get_RU(g, sol->c_R, sol->c_U);
  if (!rc) rc = SFGS_Body_Update(g, sol->step, sol->tmp_velo, sol->force);
  if (rc)
  {
    snprintf(sol->err, sizeof(sol->err), "GPU body update failed: %s (%s)", SFGS_Error_String(rc), SFGS_Last_Error(g));
    return SOL_ERR_GPU;
  }
  memcpy(sol->velo, sol->tmp_velo, sizeof(double)*6*(size_t)sol->nb);
  return 0;
}

/* upload, assemble, solve and update body velocities on the GPU */
int SOLFEC_Step_Solve(SOLFEC *sol, GAUSS_SEIDEL *gs)
{
  if (!sol || !gs) return SOL_ERR_ARG;
  if (has_spring_dashpot(sol))
  {
    snprintf(sol->err, sizeof(sol->err), "SPRING_DASHPOT materials need the PENALTY_SOLVER here");
    return SOL_ERR_ARG;
  }
  sfgs_ctx *g = SOLFEC_Gpu(sol);
  if (!g) return SOL_ERR_GPU;
  int rc = step_assemble(sol, g, gs, 0);
  if (rc) return rc;
  double t1 = now_s();
  int err = GAUSS_SEIDEL_Solve_Device(gs, g, sol->err, sizeof(sol->err));
  if (err == SOL_ERR_GPU || err == SOL_ERR_MEM) return err;
  rc = step_finish(sol, g);
  sol->t_consol += now_s() - t1;
  return rc ? rc : err;
}

/* the same with a NEWTON_SOLVER (sol.c:472): every contact is solved as a
 * SIGNORINI_COULOMB contact (nts.c:1072-1081) */
int SOLFEC_Step_Solve_Newton(SOLFEC *sol, NEWTON *ns)
{
  if (!sol || !ns) return SOL_ERR_ARG;
  sfgs_ctx *g = SOLFEC_Gpu(sol);
  if (!g) return SOL_ERR_GPU;
  int rc = step_assemble(sol, g, NULL, ns->par.bodyspace);
  if (rc) return rc;
  double t1 = now_s();
  rc = NEWTON_Solve_Device(ns, g, sol->err, sizeof(sol->err));
  if (!rc) rc = step_finish(sol, g);
  sol->t_consol += now_s() - t1;
  return rc;
}

int SOLFEC_Run_Newton(SOLFEC *sol, NEWTON *ns, double duration)
{
  if (!sol || !ns || !(duration > 0.0)) return SOL_ERR_ARG;
  double t0 = sol->time;
  while (sol->time < t0 + duration - 1e-12*sol->step)
  {
    int rc = SOLFEC_Step_Begin(sol);
    if (rc) return rc;
    rc = SOLFEC_Step_Solve_Newton(sol, ns);
    if (rc) return rc;
    rc = SOLFEC_Step_End(sol);
    if (rc) return rc;
    if (stop_file(sol)) break;   /* a STOP file created by the user ends the run (sol.c:804-805) */
  }
  return SOL_OK;
}

/* ---------------- LOCDYN dump formats ---------------- */

typedef struct { const double *point; const int *master, *slave; } DUMPKEY;

/* order of the neighbours of a dump line (dumpcmp, ldy.c:330-353): point, then
 * the sorted body pair (all rows here are contacts) */
static int dump_cmp(const DUMPKEY *k, int a, int b)
{
  for (int i = 0; i < 3; i ++)
  {
    if (k->point[3*a + i] < k->point[3*b + i]) return -1;
    if (k->point[3*a + i] > k->point[3*b + i]) return 1;
  }
  int am = k->master[a], as = k->slave[a], bm = k->master[b], bs = k->slave[b];
  int a0 = am < as ? am : as, a1 = am < as ? as : am, b0 = bm < bs ? bm : bs, b1 = bm < bs ? bs : bm;
  if (a0 != b0) return a0 < b0 ? -1 : 1;
  if (a1 != b1) return a1 < b1 ? -1 : 1;
  return 0;
}

#define G6 "%.6g"
#define V3 "(" G6 ", " G6 ", " G6 ")"
#define M9 "[" G6 ", " G6 ", " G6 ", " G6 ", " G6 ", " G6 ", " G6 ", " G6 ", " G6 "]"

/* LOCDYN_Dump (ldy.c:892-1002) on host arrays in contact order: one line per
 * row in list order (newest contact first, ldy.c:505-508), neighbours sorted
 * by dump_cmp with the blocks of a neighbour reached through both bodies
 * added together */
int SOLFEC_Locdyn_Write(const char *path, int n, const double *point, const int *master, const int *slave,
                        const double *base, const double *Wd, const double *B, const double *V, const double *U,
                        const double *R, const int *rowptr, const int *col, const double *W)
{
  FILE *f = fopen(path, "w");
  if (!f) return SOL_ERR_ARG;
  DUMPKEY key = {point, master, slave};
  int *nb = NULL, cap = 0;
  double *acc = NULL;
  for (int i = n - 1; i >= 0; i --)
  {
    const double *b = base + 9*i, *w = Wd + 9*i;
    fprintf(f, "CONTACT " V3 " (%d, %d) " M9 " " M9 " " V3 " " V3 " " V3 " " V3,
            point[3*i], point[3*i + 1], point[3*i + 2], master[i], slave[i],
            b[0], b[1], b[2], b[3], b[4], b[5], b[6], b[7], b[8], w[0], w[1], w[2], w[3], w[4], w[5], w[6], w[7], w[8],
            B[3*i], B[3*i + 1], B[3*i + 2], V[3*i], V[3*i + 1], V[3*i + 2],
            U[3*i], U[3*i + 1], U[3*i + 2], R[3*i], R[3*i + 1], R[3*i + 2]);
    /* distinct neighbours with summed blocks, insertion-sorted by dump_cmp */
    int m = 0, len = rowptr[i + 1] - rowptr[i];
    if (len > cap)
    {
      cap = len;
      if (grow(NULL, 0, &nb, cap, sizeof(int)) || grow(NULL, 0, &acc, 9LL*cap, sizeof(double)))
      {
        free(nb); free(acc); fclose(f);
        return SOL_ERR_MEM;
      }
    }
    for (int q = rowptr[i]; q < rowptr[i + 1]; q ++)
    {
      int j = col[q], at = 0;
      while (at < m && nb[at] != j) at ++;
      if (at == m)
      {
        int pos = m;
        while (pos > 0 && dump_cmp(&key, j, nb[pos - 1]) < 0) pos --;
        memmove(nb + pos + 1, nb + pos, sizeof(int)*(size_t)(m - pos));
        memmove(acc + 9*(pos + 1), acc + 9*pos, sizeof(double)*9*(size_t)(m - pos));
        nb[pos] = j;
        memset(acc + 9*pos, 0, 9*sizeof(double));
        at = pos;
        m ++;
      }
      for (int e = 0; e < 9; e ++) acc[9*at + e] += W[9*(size_t)q + e];
    }
    for (int t = 0; t < m; t ++)
    {
      int j = nb[t];
      const double *bj = base + 9*j, *a = acc + 9*t;
      fprintf(f, " => CONTACT " V3 " (%d, %d) " M9 " " M9, point[3*j], point[3*j + 1], point[3*j + 2], master[j], slave[j],
              bj[0], bj[1], bj[2], bj[3], bj[4], bj[5], bj[6], bj[7], bj[8], a[0], a[1], a[2], a[3], a[4], a[5], a[6], a[7], a[8]);
    }
    fprintf(f, "\n");
  }
  free(nb);
  free(acc);
  fclose(f);
  return 0;
}

/* LOCDYN_W_MatrixMarket (ldy.c:1005-1052): W as 3n x 3n coordinate entries,
 * rows numbered in list order */
int SOLFEC_W_MatrixMarket_Write(const char *path, int n, const double *Wd, const int *rowptr, const int *col,
                                const double *W)
{
  FILE *f = fopen(path, "w");
  if (!f) return SOL_ERR_ARG;
  long long nnz = 9LL*n + 9LL*rowptr[n];
  fprintf(f, "%%%%MatrixMarket matrix coordinate real general\n");
  fprintf(f, "%d  %d  %lld\n", 3*n, 3*n, nnz);
  for (int i = n - 1; i >= 0; i --)
  {
    int r = 3*(n - 1 - i) + 1;
    for (int a = 0; a < 3; a ++)
      for (int b = 0; b < 3; b ++) fprintf(f, "%d  %d  %.15g\n", r + a, r + b, Wd[9*i + 3*b + a]);
    for (int q = rowptr[i]; q < rowptr[i + 1]; q ++)
    {
      int c = 3*(n - 1 - col[q]) + 1;
      for (int a = 0; a < 3; a ++)
        for (int b = 0; b < 3; b ++) fprintf(f, "%d  %d  %.15g\n", r + a, c + b, W[9*(size_t)q + 3*b + a]);
    }
  }
  fclose(f);
  return 0;
}

/* the last assembled system of the domain, downloaded from the device */
static int locdyn_download(SOLFEC *sol, int **rowptr, int **col, double **W, double **Wd, double **B, double **V)
{
  sfgs_ctx *g = sol->gpu;
  int n = 0, ncol = 0;
  long long nnz = 0;
  if (!g || SFGS_Get_Sizes(g, &n, &nnz, &ncol) || n != sol->nc)
  {
    snprintf(sol->err, sizeof(sol->err), "no assembled local dynamics to dump");
    return SOL_ERR_ARG;
  }
  /* the callers free what was allocated, also after a failure */
  if (GROW(*rowptr, n + 1LL) || GROW(*col, nnz + 1) || GROW(*W, 9*(nnz + 1)) || GROW(*Wd, 9*(n + 1LL)) ||
      GROW(*B, 3*(n + 1LL)) || GROW(*V, 3*(n + 1LL))) return SOL_ERR_MEM;
  if (SFGS_Get_Locdyn(g, *rowptr, *col, NULL, *W, *Wd, NULL, *B, *V, NULL, NULL, NULL))
  {
    snprintf(sol->err, sizeof(sol->err), "LOCDYN download failed: %s", SFGS_Last_Error(g));
    return SOL_ERR_GPU;
  }
  return 0;
}

int SOLFEC_Locdyn_Dump(SOLFEC *sol, const char *path)
{
  int *rowptr = NULL, *col = NULL;
  double *W = NULL, *Wd = NULL, *B = NULL, *V = NULL;
  int rc = dev_pull(sol);
  if (!rc) rc = locdyn_download(sol, &rowptr, &col, &W, &Wd, &B, &V);
  if (!rc) rc = SOLFEC_Locdyn_Write(path, sol->nc, sol->c_point, sol->c_master, sol->c_slave, sol->c_base, Wd, B, V,
                                    sol->c_U, sol->c_R, rowptr, col, W);
  free(rowptr); free(col); free(W); free(Wd); free(B); free(V);
  return rc;
}

int SOLFEC_W_MatrixMarket(SOLFEC *sol, const char *path)
{
  int *rowptr = NULL, *col = NULL;
  double *W = NULL, *Wd = NULL, *B = NULL, *V = NULL;
  int rc = dev_pull(sol);
  if (!rc) rc = locdyn_download(sol, &rowptr, &col, &W, &Wd, &B, &V);
  if (!rc) rc = SOLFEC_W_MatrixMarket_Write(path, sol->nc, Wd, rowptr, col, W);
  free(rowptr); free(col); free(W); free(Wd); free(B); free(V);
  return rc;
}

/* ---------------- PENALTY_SOLVER ---------------- */

PENALTY* PENALTY_Create(int implicit)
{
  PENALTY *ps = calloc(1, sizeof(PENALTY));
  if (ps) ps->implicit = implicit ? 1 : 0;
  return ps;
}

void PENALTY_Destroy(PENALTY *ps)
{
  if (!ps) return;
  free(ps->kcp);
  free(ps);
}

static int penalty_failed(SOLFEC *sol, sfgs_ctx *g, int rc, const char *what)
{
  snprintf(sol->err, sizeof(sol->err), "%s failed: %s (%s)", what, SFGS_Error_String(rc), SFGS_Last_Error(g));
  return SOL_ERR_GPU;
}

/* PENALTY_Solve (pes.c:247-281): the spring-dashpot law per contact with the
 * material of its surface pair (pes.c:272), then the velocity update */
int SOLFEC_Step_Solve_Penalty(SOLFEC *sol, PENALTY *ps)
{
  if (!sol || !ps) return SOL_ERR_ARG;
  sfgs_ctx *g = SOLFEC_Gpu(sol);
  if (!g) return SOL_ERR_GPU;
  double t0 = now_s();
  int rc;
  sol->locdyn_current = 0;   /* the penalty update assembles no off-diagonal blocks to dump */
  if (sol->devloop)
  {
    rc = SFGS_Penalty_Solve(g, sol->step, ps->implicit, NULL, NULL, NULL);
    if (rc) return penalty_failed(sol, g, rc, "PENALTY_SOLVER on GPU");
    sol->t_consol += now_s() - t0;
    return 0;
  }
  const int n = sol->nc;
  if (n > ps->kcp_cap)
  {
    if (GROW(ps->kcp, 3LL*n)) return SOL_ERR_MEM;
    ps->kcp_cap = n;
  }
  for (int i = 0; i < n; i ++)
  {
    const MATPAIR *m = find_material(sol, sol->surf[sol->c_master[i]], sol->surf[sol->c_slave[i]]);
    ps->kcp[i] = m->spring;
    ps->kcp[(size_t)n + i] = m->dashpot;
    ps->kcp[2*(size_t)n + i] = m->hpow;
  }
  rc = SFGS_Bodies_Upload(g, sol->nb, sol->kind, sol->rot, sol->x0, sol->jinv, sol->minv, sol->velo, sol->velo0);
  if (!rc) rc = SFGS_Contacts_Upload(g, sol->nc, sol->c_master, sol->c_slave, sol->c_mpnt, sol->c_spnt, sol->c_base,
                                     sol->c_gap, sol->c_fri, sol->c_rest, sol->c_coh, sol->c_R, sol->c_U);
  if (!rc) rc = SFGS_Penalty_Solve(g, sol->step, ps->implicit, ps->kcp, ps->kcp + n, ps->kcp + 2*(size_t)n);
  if (!rc && has_cohesion(sol)) rc = SFGS_Get_Cohesion(g, sol->c_coh);   /* broken by the law itself */
  if (!rc) rc = SFGS_Get_RU(g, sol->c_R, sol->c_U);
  if (!rc) rc = SFGS_Body_Update(g, sol->step, sol->tmp_velo, sol->force);
  if (rc) return penalty_failed(sol, g, rc, "PENALTY_SOLVER on GPU");
  memcpy(sol->velo, sol->tmp_velo, sizeof(double)*6*(size_t)sol->nb);
  sol->t_consol += now_s() - t0;
  return 0;
}

int SOLFEC_Run_Penalty(SOLFEC *sol, PENALTY *ps, double duration)
{
  if (!sol || !ps || !(duration > 0.0)) return SOL_ERR_ARG;
  double t0 = sol->time;
  while (sol->time < t0 + duration - 1e-12*sol->step)
  {
    int rc = SOLFEC_Step_Begin(sol);
    if (rc) return rc;
    rc = SOLFEC_Step_Solve_Penalty(sol, ps);
    if (rc) return rc;
    rc = SOLFEC_Step_End(sol);
    if (rc) return rc;
    if (stop_file(sol)) break;   /* a STOP file created by the user ends the run (sol.c:804-805) */
  }
  return SOL_OK;
}

int SOLFEC_Run(SOLFEC *sol, GAUSS_SEIDEL *gs, double duration)
{
  if (!sol || !gs || !(duration > 0.0)) return SOL_ERR_ARG;
  double t0 = sol->time;
  while (sol->time < t0 + duration - 1e-12*sol->step)
  {
    int rc = SOLFEC_Step_Begin(sol);
    if (rc) return rc;
    rc = SOLFEC_Step_Solve(sol, gs);
    if (rc) return rc;
    rc = SOLFEC_Step_End(sol);
    if (rc) return rc;
    if (stop_file(sol)) break;   /* a STOP file created by the user ends the run (sol.c:804-805) */
  }
  return SOL_OK;
}
