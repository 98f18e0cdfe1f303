/* or_dom.c -- contact detection of spheres and planes, and the rigid time step.
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * OR_Sphere_Contact follows gjk_sphere_sphere (gjk.c:560-584), the
 * orientation rule of detect_sphere_sphere (goc.c:421-455: the normal is
 * built from the sphere whose centre compares lexicographically smaller under
 * POINTS_COMPARE, alg.h:35-52, which becomes the slave) and sphere_sphere_gap
 * (goc.c:50-59).  Coincident centres give no contact (the reference would
 * normalise a zero vector there).
 *
 * OR_Plane_Contact follows detect_convex_sphere (goc.c:308-348) with the
 * plane's half-space as the convex: the closest point of the plane to the
 * centre is its projection, the normal is the plane normal (the centre lies on
 * the positive side), gap = min(distance - r, 0); the plane is the slave.
 *
 * OR_Detect restates one DOM_Update_Begin contact pass (dom.c:3630) by brute
 * force, with no spatial grid: every given contact is re-evaluated and kept in
 * its order while its pair still touches (update_contact, dom.c:430-483); then
 * every sphere pair a < b whose pair is not kept and whose gap is below
 * GEOMETRIC_EPSILON is appended in ascending (a, b) order, then every
 * plane-sphere pair in ascending (plane, sphere) order (overlap_create,
 * dom.c:348-416, with the order of the pairs made deterministic).
 *
 * OR_Rig_Step_Begin / OR_Rig_Step_End follow BODY_Dynamic_Step_Begin
 * (bod.c:1313-1373) and BODY_Dynamic_Step_End (bod.c:1425-1492) for a rigid
 * body of scheme RIG_NEG under gravity only (rig_force gives m g and no
 * torque), with EXPMAP (alg.h:1138-1204): the Taylor series of sin(t)/t and
 * (1 - cos t)/t^2 through t^12 below (10 degrees)^2, libm sin and cos above.
 * Obstacles stay fixed (their velocity is zero).
 *
 * OR_Rig_Static_Begin / OR_Rig_Static_End follow BODY_Static_Step_Begin
 * (bod.c:1588-1614, with rig_static_force, bod.c:357-386) and
 * BODY_Static_Step_End (bod.c:1616-1645) for a rigid body under gravity only:
 * u(t+h) = h M^-1 f(t+h), the previous velocity not carried, the angular part
 * of f the gyroscopic term -W1 x J W1 with W1 = J^-1 exp[-hW] J W; at the end
 * x += h v and R = R exp[h W].  The previous velocity (velo0) is left as it
 * is.  Pinned in tests/test_oracle_dom.py by closed forms (free fall from any
 * velocity, spin about a principal axis stopped in one step, Euler's equations
 * of torque-free motion as h goes to 0, the end step against a Rodrigues
 * rotation).
 *
 * Parity: pinned in tests/test_oracle_dom.py by closed forms (two spheres on a
 * line, a sphere on a plane, the orientation rule, free fall, torque-free spin
 * of an isotropic body about a fixed axis, EXPMAP against a Rodrigues matrix
 * from numpy's sin and cos) and by the brute force itself, which has no grid
 * that could miss a pair. */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "oracle.h"
#include "or_la.h"

#define OR_GEOMETRIC_EPSILON 1e-6   /* alg.c:24 */
#define OR_DEG_10_SQ 3.0461741978671E-02

static int or_pntcmp(const double *a, const double *b)
{
  for (int k = 0; k < 3; k ++)
  {
    if (a[k] + OR_GEOMETRIC_EPSILON < b[k] - OR_GEOMETRIC_EPSILON) return -1;
    if (a[k] - OR_GEOMETRIC_EPSILON > b[k] + OR_GEOMETRIC_EPSILON) return 1;
  }
  return 0;
}

int OR_Sphere_Contact(const double *ca, double ra, const double *cb, double rb, double *out)
{
  /* out: swap (0: a is the slave, 1: b is), slave point 3, master point 3, normal 3, gap */
  double d[3], p[3], q[3];
  for (int k = 0; k < 3; k ++) d[k] = ca[k] - cb[k];
  double dl = sqrt(or_dot(d, d));
  if (dl == 0.0) return 0;
  double g, s = ra / dl, t = rb / dl;
  for (int k = 0; k < 3; k ++) { p[k] = ca[k] - s*d[k]; q[k] = cb[k] + t*d[k]; }
  if (dl > ra + rb) g = dl - (ra + rb);
  else
  {
    for (int k = 0; k < 3; k ++) { p[k] = 0.5*(p[k] + q[k]); q[k] = p[k]; }
    g = 0.0;
  }
  if (!(g < OR_GEOMETRIC_EPSILON)) return 0;
  int swap = or_pntcmp(ca, cb) > 0;
  const double *cs = swap ? cb : ca, *cm = swap ? ca : cb, *ps = swap ? q : p, *pm = swap ? p : q;
  double rs = swap ? rb : ra, rm = swap ? ra : rb, n[3];
  for (int k = 0; k < 3; k ++) n[k] = ps[k] - cs[k];
  double nl = sqrt(or_dot(n, n));
  for (int k = 0; k < 3; k ++) n[k] /= nl;
  double x[3] = {cm[0] - cs[0], cm[1] - cs[1], cm[2] - cs[2]}, dn = or_dot(x, n), e = rs + rm;
  out[0] = swap;
  for (int k = 0; k < 3; k ++) { out[1 + k] = ps[k]; out[4 + k] = pm[k]; out[7 + k] = n[k]; }
  out[10] = e > dn ? dn - e : 0.0;
  return 1;
}

int OR_Plane_Contact(const double *plane, const double *c, double r, double *out)
{
  /* out: slave (plane) point 3, master (sphere) point 3, normal 3, gap */
  const double *n = plane + 3;
  double x[3] = {c[0] - plane[0], c[1] - plane[1], c[2] - plane[2]}, dist = or_dot(x, n);
  if (!(dist < r + OR_GEOMETRIC_EPSILON)) return 0;
  for (int k = 0; k < 3; k ++)
  {
    out[k] = c[k] - dist*n[k];
    out[3 + k] = c[k] - r*n[k];
    out[6 + k] = n[k];
  }
  out[9] = fmin(dist - r, 0.0);
  return 1;
}

/* geometry of a (slave, master) pair: plane-sphere or sphere-sphere in the given orientation */
static int pair_geometry(int nb, const int *shape, const double *rad, const double *plane, const double *x,
                         int s, int m, double *G)
{
  if (s < 0 || s >= nb || m < 0 || m >= nb) return 0;
  if (shape[s] == 1 && shape[m] == 0) return OR_Plane_Contact(plane + 6*s, x + 3*m, rad[m], G);
  if (shape[s] == 0 && shape[m] == 0)
  {
    double o[11];
    if (!OR_Sphere_Contact(x + 3*s, rad[s], x + 3*m, rad[m], o)) return 0;
    /* an existing contact keeps its master and slave; the geometry of the pair
     * is the same whichever sphere is named first, up to the normal's sign */
    if (o[0] == 0.0) { memcpy(G, o + 1, 9*sizeof(double)); G[9] = o[10]; }
    else
    {
      for (int k = 0; k < 3; k ++) { G[k] = o[4 + k]; G[3 + k] = o[1 + k]; G[6 + k] = -o[7 + k]; }
      G[9] = o[10];
    }
    return 1;
  }
  return 0;
}

long long OR_Detect(int nb, const int *shape, const double *rad, const double *plane, const double *x,
                    int nold, const int *master_old, const int *slave_old, long long cap,
                    int *master, int *slave, int *old_index, double *geo)
{
  long long n = 0;
  char *keep = calloc((size_t)nold + 1, 1);
  int *key = malloc(sizeof(int)*(2*(size_t)nold + 1));
  /* kept contacts, in order */
  for (int i = 0; i < nold; i ++)
  {
    double G[10];
    if (!pair_geometry(nb, shape, rad, plane, x, slave_old[i], master_old[i], G)) continue;
    keep[i] = 1;
    if (n < cap)
    {
      master[n] = master_old[i]; slave[n] = slave_old[i]; old_index[n] = i;
      memcpy(geo + 10*n, G, sizeof(G));
    }
    n ++;
  }
  /* whether an unordered pair is among the kept ones: a linear scan suffices
   * for the sizes the oracle checks */
  int nk = 0;
  for (int i = 0; i < nold; i ++)
    if (keep[i])
    {
      key[2*nk] = master_old[i] < slave_old[i] ? master_old[i] : slave_old[i];
      key[2*nk + 1] = master_old[i] < slave_old[i] ? slave_old[i] : master_old[i];
      nk ++;
    }
  for (int a = 0; a < nb; a ++)
  {
    if (shape[a] != 0) continue;
    for (int b = a + 1; b < nb; b ++)
    {
      if (shape[b] != 0) continue;
      double o[11];
      if (!OR_Sphere_Contact(x + 3*a, rad[a], x + 3*b, rad[b], o)) continue;
      int known = 0;
      for (int k = 0; k < nk && !known; k ++) known = key[2*k] == a && key[2*k + 1] == b;
      if (known) continue;
      if (n < cap)
      {
        int sw = o[0] != 0.0;
        master[n] = sw ? a : b; slave[n] = sw ? b : a; old_index[n] = -1;
        memcpy(geo + 10*n, o + 1, 9*sizeof(double));
        geo[10*n + 9] = o[10];
      }
      n ++;
    }
  }
  for (int p = 0; p < nb; p ++)
  {
    if (shape[p] != 1) continue;
    for (int s = 0; s < nb; s ++)
    {
      if (shape[s] != 0) continue;
      double G[10];
      if (!OR_Plane_Contact(plane + 6*p, x + 3*s, rad[s], G)) continue;
      int known = 0, a = p < s ? p : s, b = p < s ? s : p;
      for (int k = 0; k < nk && !known; k ++) known = key[2*k] == a && key[2*k + 1] == b;
      if (known) continue;
      if (n < cap)
      {
        master[n] = s; slave[n] = p; old_index[n] = -1;
        memcpy(geo + 10*n, G, sizeof(G));
      }
      n ++;
    }
  }
  free(keep);
  free(key);
  return n;
}

void OR_Expmap(const double *O, double *R)
{
  double t2 = or_dot(O, O), a, b;
  if (t2 < OR_DEG_10_SQ)
  {
    /* sin(t)/t = sum (-t^2)^k / (2k+1)!, (1 - cos t)/t^2 = sum (-t^2)^k / (2k+2)!, k = 0..6 */
    double fa = 1.0, fb = 0.5, p = 1.0;
    a = 1.0; b = 0.5;
    for (int k = 1; k <= 6; k ++)
    {
      fa /= (double)((2*k)*(2*k + 1));       /* 1/(2k+1)! */
      fb /= (double)((2*k + 1)*(2*k + 2));   /* 1/(2k+2)! */
      p = -p*t2;                             /* (-t^2)^k */
      a += fa*p; b += fb*p;
    }
  }
  else
  {
    double t = sqrt(t2);
    a = sin(t)/t;
    b = (1.0 - cos(t))/t2;
  }
  double x = O[0], y = O[1], z = O[2];
  R[0] = 1.0 - b*(y*y + z*z); R[3] = b*x*y - a*z;           R[6] = b*x*z + a*y;
  R[1] = b*x*y + a*z;         R[4] = 1.0 - b*(x*x + z*z);   R[7] = b*y*z - a*x;
  R[2] = b*x*z - a*y;         R[5] = b*y*z + a*x;           R[8] = 1.0 - b*(x*x + y*y);
}

static void or_cross(const double *a, const double *b, double *c)
{
  double c0 = a[1]*b[2] - a[2]*b[1], c1 = a[2]*b[0] - a[0]*b[2], c2 = a[0]*b[1] - a[1]*b[0];
  c[0] = c0; c[1] = c1; c[2] = c2;
}

void OR_Rig_Step_Begin(int kind, double h, const double *gravity, double mass, double minv, const double *J,
                       const double *I, double *velo, double *velo0, double *R, double *x, double *A)
{
  double half = 0.5*h, *W = velo, *v = velo + 3;
  memcpy(velo0, velo, 6*sizeof(double));
  if (kind == OR_OBS) return;
  double O[3], DR[9], T[9], f[6], JW[3], W05[3], JW05[3], c[3];
  for (int k = 0; k < 3; k ++) O[k] = half*W[k];
  OR_Expmap(O, DR);
  or_mm(R, DR, T);
  memcpy(R, T, sizeof(T));                              /* R(t+h/2) = R(t) exp[(h/2) W(t)] */
  for (int k = 0; k < 3; k ++) x[k] += half*v[k];      /* x(t+h/2) = x(t) + (h/2) v(t) */
  f[0] = f[1] = f[2] = 0.0;                            /* body-frame torque: none under gravity */
  for (int k = 0; k < 3; k ++) f[3 + k] = mass*gravity[k];
  double DJW[3];
  or_mv(J, W, JW);
  or_mtv(DR, JW, DJW);
  for (int k = 0; k < 3; k ++) A[k] = DJW[k] + h*f[k];  /* exp[-(h/2)W] J W + h T */
  for (int k = 0; k < 3; k ++) O[k] = half*f[k] + DJW[k];
  or_mv(I, O, W05);                                    /* I [exp[-(h/2)W] J W + (h/2) T] */
  or_mv(J, W05, JW05);
  or_cross(W05, JW05, c);
  for (int k = 0; k < 3; k ++) f[k] -= c[k];           /* T - W05 x J W05 */
  double dW[3];
  or_mv(I, f, dW);
  for (int k = 0; k < 3; k ++) { W[k] += h*dW[k]; v[k] += h*(minv*f[3 + k]); }   /* u += h iM f */
}

void OR_Rig_Step_End(int kind, double h, const double *I, const double *force, double *velo, double *R, double *x,
                     double *A)
{
  if (kind == OR_OBS) return;
  double half = 0.5*h, *W = velo, *v = velo + 3, O[3], DR[9], T[9];
  for (int k = 0; k < 3; k ++) x[k] += half*v[k];      /* x(t+h) = x(t+h/2) + (h/2) v(t+h) */
  for (int k = 0; k < 3; k ++) O[k] = half*W[k];
  OR_Expmap(O, DR);
  or_mm(R, DR, T);
  memcpy(R, T, sizeof(T));                              /* R(t+h) = R(t+h/2) exp[(h/2) W(t+h)] */
  for (int k = 0; k < 3; k ++) A[k] += h*force[k];
  or_mtv(DR, A, O);
  or_mv(I, O, W);                                      /* W(t+h) = I exp[-(h/2)W](...) */
}

void OR_Rig_Static_Begin(int kind, double h, const double *gravity, double mass, double minv, const double *J,
                         const double *I, double *velo)
{
  if (kind == OR_OBS) return;
  double *W = velo, *v = velo + 3, O[3], DR[9], f[6], JW[3], DJW[3], W1[3], JW1[3], c[3], dW[3];
  for (int k = 0; k < 3; k ++) O[k] = h*W[k];
  OR_Expmap(O, DR);                                    /* exp[h W(t)] */
  f[0] = f[1] = f[2] = 0.0;                            /* body-frame torque: none under gravity */
  for (int k = 0; k < 3; k ++) f[3 + k] = mass*gravity[k];
  or_mv(J, W, JW);
  or_mtv(DR, JW, DJW);                                 /* exp[-hW] J W */
  for (int k = 0; k < 3; k ++) O[k] = h*f[k] + DJW[k];
  or_mv(I, O, W1);                                     /* W1 = I [exp[-hW] J W + h T] */
  or_mv(J, W1, JW1);
  or_cross(W1, JW1, c);
  for (int k = 0; k < 3; k ++) f[k] -= c[k];           /* T - W1 x J W1 */
  or_mv(I, f, dW);
  for (int k = 0; k < 3; k ++) { W[k] = h*dW[k]; v[k] = h*(minv*f[3 + k]); }   /* u = h iM f */
}

void OR_Rig_Static_End(int kind, double h, double *velo, double *R, double *x)
{
  if (kind == OR_OBS) return;
  double *W = velo, *v = velo + 3, O[3], DR[9], T[9];
  for (int k = 0; k < 3; k ++) x[k] += h*v[k];         /* x(t+h) = x(t) + h v(t+h) */
  for (int k = 0; k < 3; k ++) O[k] = h*W[k];
  OR_Expmap(O, DR);
  or_mm(R, DR, T);
  memcpy(R, T, sizeof(T));                             /* R(t+h) = R(t) exp[h W(t+h)] */
}
