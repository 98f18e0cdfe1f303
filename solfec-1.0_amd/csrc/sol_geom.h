/* sol_geom.h -- rigid-body geometry and time integration shared by the host
 * layer (gcc, C) and the device loop (hipcc, HIP C++).
 *
 * One source for both sides: compiled on both without floating-point
 * contraction (-ffp-contract=off), with only +, -, *, / and sqrt (correctly
 * rounded on both), the host and the device loops compute the same bits from
 * the same state.  The rotation exponential therefore uses its own series and
 * argument halving rather than the math library's sin and cos.
 *
 * Column-major 3x3 matrices throughout (alg.h:881-953). */
#ifndef SOL_GEOM_H
#define SOL_GEOM_H

#include <math.h>

#if defined(__HIPCC__)
#define SOL_GEOM_FN __host__ __device__ static inline
#else
#define SOL_GEOM_FN static inline
#endif

#define SOL_GEOMETRIC_EPSILON 1e-6   /* alg.c:24 */

SOL_GEOM_FN void geom_mm(const double *A, const double *B, double *C)   /* C = A B */
{
  double T[9];
  for (int c = 0; c < 3; c ++)
    for (int r = 0; r < 3; r ++)
      T[r + 3*c] = A[r]*B[3*c] + A[r + 3]*B[3*c + 1] + A[r + 6]*B[3*c + 2];
  for (int e = 0; e < 9; e ++) C[e] = T[e];
}

SOL_GEOM_FN void geom_mv(const double *A, const double *x, double *y)   /* y = A x */
{
  double t0 = A[0]*x[0] + A[3]*x[1] + A[6]*x[2];
  double t1 = A[1]*x[0] + A[4]*x[1] + A[7]*x[2];
  double t2 = A[2]*x[0] + A[5]*x[1] + A[8]*x[2];
  y[0] = t0; y[1] = t1; y[2] = t2;
}

SOL_GEOM_FN void geom_mtv(const double *A, const double *x, double *y)  /* y = A' x */
{
  double t0 = A[0]*x[0] + A[1]*x[1] + A[2]*x[2];
  double t1 = A[3]*x[0] + A[4]*x[1] + A[5]*x[2];
  double t2 = A[6]*x[0] + A[7]*x[1] + A[8]*x[2];
  y[0] = t0; y[1] = t1; y[2] = t2;
}

SOL_GEOM_FN void geom_cross(const double *a, const double *b, double *c)
{
  double c0 = a[1]*b[2] - a[2]*b[1], c1 = a[2]*b[0] - a[0]*b[2], c2 = a[0]*b[1] - a[1]*b[0];
  c[0] = c0; c[1] = c1; c[2] = c2;
}

/* sin(t)/t and (1 - cos t)/t^2.  For t <= 1/2 their Taylor series in nested
 * form, to the term below double precision (t^14 / 15!); above, sin and cos of
 * t / 2^m <= 1/2 from the series of sin and cos, then m angle doublings. */
SOL_GEOM_FN void geom_sinc(double t2, double *a, double *b)
{
  if (t2 <= 0.25)
  {
    *a = 1.0 - t2/6.0*(1.0 - t2/20.0*(1.0 - t2/42.0*(1.0 - t2/72.0*(1.0 - t2/110.0*(1.0 - t2/156.0*(1.0 - t2/210.0))))));
    *b = 0.5*(1.0 - t2/12.0*(1.0 - t2/30.0*(1.0 - t2/56.0*(1.0 - t2/90.0*(1.0 - t2/132.0*(1.0 - t2/182.0*(1.0 - t2/240.0)))))));
    return;
  }
  double t = sqrt(t2), x = t;
  int m = 0;
  while (x > 0.5 && m < 60) { x *= 0.5; m ++; }
  double x2 = x*x;
  double s = x*(1.0 - x2/6.0*(1.0 - x2/20.0*(1.0 - x2/42.0*(1.0 - x2/72.0*(1.0 - x2/110.0*(1.0 - x2/156.0*(1.0 - x2/210.0)))))));
  double c = 1.0 - x2*0.5*(1.0 - x2/12.0*(1.0 - x2/30.0*(1.0 - x2/56.0*(1.0 - x2/90.0*(1.0 - x2/132.0*(1.0 - x2/182.0*(1.0 - x2/240.0)))))));
  for (int k = 0; k < m; k ++)
  {
    double s2 = 2.0*s*c, c2 = (c - s)*(c + s);
    s = s2; c = c2;
  }
  *a = s/t;
  *b = (1.0 - c)/t2;
}

/* rotation exp([O]x) by the Rodrigues formula (EXPMAP, alg.h:1138-1204) */
SOL_GEOM_FN void geom_expmap(const double *O, double *R)
{
  double a, b;
  geom_sinc(O[0]*O[0] + O[1]*O[1] + O[2]*O[2], &a, &b);
  double x = O[0], y = O[1], z = O[2];
  R[0] = 1.0 - b*(y*y + z*z); R[3] = -a*z + b*x*y;          R[6] = a*y + b*x*z;
  R[1] = a*z + b*x*y;         R[4] = 1.0 - b*(x*x + z*z);   R[7] = -a*x + b*y*z;
  R[2] = -a*y + b*x*z;        R[5] = a*x + b*y*z;           R[8] = 1.0 - b*(x*x + y*y);
}

/* local frame with n as the last column (localbase, dom.c:133-151) */
SOL_GEOM_FN void geom_localbase(const double *n, double *loc)
{
  const double e0[3] = {1.0, 0.0, 0.0}, e1[3] = {0.0, 1.0, 0.0};
  loc[6] = n[0]; loc[7] = n[1]; loc[8] = n[2];
  geom_cross(e0, n, loc);
  double len = sqrt(loc[0]*loc[0] + loc[1]*loc[1] + loc[2]*loc[2]);
  if (len < SOL_GEOMETRIC_EPSILON)
  {
    geom_cross(e1, n, loc);
    len = sqrt(loc[0]*loc[0] + loc[1]*loc[1] + loc[2]*loc[2]);
  }
  loc[0] /= len; loc[1] /= len; loc[2] /= len;
  geom_cross(loc, n, loc + 3);
}

/* lexicographic comparison with the geometric tolerance (POINTS_COMPARE, alg.h:35-52) */
SOL_GEOM_FN int geom_pntcmp(const double *a, const double *b)
{
  for (int k = 0; k < 3; k ++)
  {
    if (a[k] + SOL_GEOMETRIC_EPSILON < b[k] - SOL_GEOMETRIC_EPSILON) return -1;
    if (a[k] - SOL_GEOMETRIC_EPSILON > b[k] + SOL_GEOMETRIC_EPSILON) return 1;
  }
  return 0;
}

/* referential image of a spatial point (BODY_Ref_Point): X = X0 + R'(p - x) */
SOL_GEOM_FN void geom_ref_point(const double *rot, const double *x, const double *x0, const double *p, double *X)
{
  double d[3] = {p[0] - x[0], p[1] - x[1], p[2] - x[2]}, t[3];
  geom_mtv(rot, d, t);
  for (int k = 0; k < 3; k ++) X[k] = x0[k] + t[k];
}

/* spatial image of a referential point (BODY_Cur_Point, bod.c:1706-1715): p = x + R(X - X0) */
SOL_GEOM_FN void geom_cur_point(const double *rot, const double *x, const double *x0, const double *X, double *p)
{
  double d[3] = {X[0] - x0[0], X[1] - x0[1], X[2] - x0[2]}, t[3];
  geom_mv(rot, d, t);
  for (int k = 0; k < 3; k ++) p[k] = x[k] + t[k];
}

/* sphere pair given as (slave a, master b): 1 when in contact (gap below
 * GEOMETRIC_EPSILON, goc.c:658-677 / gjk.c:560-584), with the points, the
 * normal from a towards b and the gap (sphere_sphere_gap, goc.c:50-59) */
SOL_GEOM_FN int geom_sphere_pair(const double *ca, double ra, const double *cb, double rb,
                                 double *spnt, double *mpnt, double *normal, double *gap)
{
  double d[3] = {ca[0] - cb[0], ca[1] - cb[1], ca[2] - cb[2]};
  double dl = sqrt(d[0]*d[0] + d[1]*d[1] + d[2]*d[2]), g;
  if (dl == 0.0) return 0;
  for (int k = 0; k < 3; k ++) { spnt[k] = ca[k] - ra/dl*d[k]; mpnt[k] = cb[k] + rb/dl*d[k]; }
  if (dl > ra + rb) g = dl - (ra + rb);
  else
  {
    for (int k = 0; k < 3; k ++) { spnt[k] = 0.5*(spnt[k] + mpnt[k]); mpnt[k] = spnt[k]; }
    g = 0.0;
  }
  if (g >= SOL_GEOMETRIC_EPSILON) return 0;
  double nl = 0.0;
  for (int k = 0; k < 3; k ++) { normal[k] = spnt[k] - ca[k]; nl += normal[k]*normal[k]; }
  nl = sqrt(nl);
  for (int k = 0; k < 3; k ++) normal[k] /= nl;
  double dn = (cb[0] - ca[0])*normal[0] + (cb[1] - ca[1])*normal[1] + (cb[2] - ca[2])*normal[2];
  double e = ra + rb;
  *gap = e > dn ? dn - e : 0.0;
  return 1;
}

/* principal frame Q (current) and semi-axes S of a convex body: an ellipsoid's
 * Q = R R0; a sphere is an ellipsoid with equal semi-axes in any frame */
SOL_GEOM_FN void geom_frame(int shape, const double *rot, const double *rot0, const double *sca, double rad,
                            double *Q, double *S)
{
  if (shape == 2)
  {
    geom_mm(rot, rot0, Q);
    S[0] = sca[0]; S[1] = sca[1]; S[2] = sca[2];
  }
  else
  {
    for (int e = 0; e < 9; e ++) Q[e] = (e % 4 == 0) ? 1.0 : 0.0;
    S[0] = S[1] = S[2] = rad;
  }
}

/* support function h(d) = |S Q' d| of the centred ellipsoid and the offset of
 * its support point, A d / h with A = Q S^2 Q' */
SOL_GEOM_FN double geom_ell_support(const double *Q, const double *S, const double *d, double *off)
{
  double t[3], u[3], v[3];
  geom_mtv(Q, d, t);
  for (int k = 0; k < 3; k ++) u[k] = S[k]*t[k];
  double h = sqrt(u[0]*u[0] + u[1]*u[1] + u[2]*u[2]);
  for (int k = 0; k < 3; k ++) v[k] = S[k]*u[k]/h;
  geom_mv(Q, v, off);
  return h;
}

/* pair of convex bodies with at least one ellipsoid, as (slave a, master b).
 * The normal n, from a towards b, maximises the separation along n,
 *   gap(n) = n.(c_b - c_a) - h_a(n) - h_b(n),
 * whose largest value is the distance of apart bodies and minus the
 * penetration depth of overlapping ones; Newton's method on the tangent plane
 * of the unit sphere finds it from n = (c_b - c_a)/|c_b - c_a|.  The contact
 * points are the support points c_a + h_a'(n) and c_b - h_b'(n), merged at
 * their midpoint when the bodies overlap, as for spheres.  The reference
 * (goc.c:457-492, gjk.c:651-836) takes the normal at the GJK closest points
 * instead; both give the common normal of touching bodies. */
SOL_GEOM_FN int geom_ellip_pair(const double *ca, const double *Qa, const double *Sa,
                                const double *cb, const double *Qb, const double *Sb,
                                double *spnt, double *mpnt, double *normal, double *gap)
{
  double Aa[9], Ab[9], T[9];
  for (int c = 0; c < 3; c ++)
    for (int r = 0; r < 3; r ++) T[r + 3*c] = Sa[r]*Sa[r]*Qa[c + 3*r];
  geom_mm(Qa, T, Aa);
  for (int c = 0; c < 3; c ++)
    for (int r = 0; r < 3; r ++) T[r + 3*c] = Sb[r]*Sb[r]*Qb[c + 3*r];
  geom_mm(Qb, T, Ab);
  double d[3] = {cb[0] - ca[0], cb[1] - ca[1], cb[2] - ca[2]};
  double dl = sqrt(d[0]*d[0] + d[1]*d[1] + d[2]*d[2]);
  if (dl == 0.0) return 0;
  double n[3] = {d[0]/dl, d[1]/dl, d[2]/dl}, oa[3], ob[3], ha = 0.0, hb = 0.0, g = 0.0;
  for (int it = 0; it < 50; it ++)
  {
    ha = geom_ell_support(Qa, Sa, n, oa);
    hb = geom_ell_support(Qb, Sb, n, ob);
    double grad[3], loc[9], H[9];
    for (int k = 0; k < 3; k ++) grad[k] = d[k] - oa[k] - ob[k];
    g = grad[0]*n[0] + grad[1]*n[1] + grad[2]*n[2];   /* = gap(n) */
    /* Euclidean Hessian: -(A_a - oa oa')/h_a - (A_b - ob ob')/h_b */
    for (int c = 0; c < 3; c ++)
      for (int r = 0; r < 3; r ++)
        H[r + 3*c] = -(Aa[r + 3*c] - oa[r]*oa[c])/ha - (Ab[r + 3*c] - ob[r]*ob[c])/hb;
    geom_localbase(n, loc);
    const double *e1 = loc, *e2 = loc + 3;
    double g1 = e1[0]*grad[0] + e1[1]*grad[1] + e1[2]*grad[2];
    double g2 = e2[0]*grad[0] + e2[1]*grad[1] + e2[2]*grad[2];
    if (sqrt(g1*g1 + g2*g2) <= 1e-14*(dl + ha + hb)) break;
    double He1[3], He2[3];
    geom_mv(H, e1, He1);
    geom_mv(H, e2, He2);
    /* Riemannian Hessian on the tangent plane: e_i' H e_j - (grad.n) delta_ij */
    double h11 = e1[0]*He1[0] + e1[1]*He1[1] + e1[2]*He1[2] - g;
    double h12 = e1[0]*He2[0] + e1[1]*He2[1] + e1[2]*He2[2];
    double h22 = e2[0]*He2[0] + e2[1]*He2[1] + e2[2]*He2[2] - g;
    double det = h11*h22 - h12*h12, s1, s2;
    if (h11 < 0.0 && det > 0.0) { s1 = -(h22*g1 - h12*g2)/det; s2 = -(h11*g2 - h12*g1)/det; }
    else { double al = 1.0/(dl + ha + hb); s1 = al*g1; s2 = al*g2; }   /* ascent step away from a maximum */
    double sl = sqrt(s1*s1 + s2*s2);
    if (sl > 0.5) { s1 *= 0.5/sl; s2 *= 0.5/sl; }
    double nl = 0.0;
    for (int k = 0; k < 3; k ++) { n[k] += s1*e1[k] + s2*e2[k]; nl += n[k]*n[k]; }
    nl = sqrt(nl);
    for (int k = 0; k < 3; k ++) n[k] /= nl;
  }
  if (g >= SOL_GEOMETRIC_EPSILON) return 0;
  ha = geom_ell_support(Qa, Sa, n, oa);
  hb = geom_ell_support(Qb, Sb, n, ob);
  g = n[0]*d[0] + n[1]*d[1] + n[2]*d[2] - ha - hb;
  for (int k = 0; k < 3; k ++) { spnt[k] = ca[k] + oa[k]; mpnt[k] = cb[k] - ob[k]; normal[k] = n[k]; }
  if (g < 0.0)
    for (int k = 0; k < 3; k ++) { spnt[k] = 0.5*(spnt[k] + mpnt[k]); mpnt[k] = spnt[k]; }
  *gap = g < 0.0 ? g : 0.0;
  return 1;
}

/* plane P = (point, unit normal) as slave and sphere (c, r) as master
 * (goc.c:308-348 with the plane as the convex) */
SOL_GEOM_FN int geom_plane_sphere(const double *P, const double *c, double r,
                                  double *spnt, double *mpnt, double *normal, double *gap)
{
  const double *n = P + 3;
  double dist = (c[0] - P[0])*n[0] + (c[1] - P[1])*n[1] + (c[2] - P[2])*n[2];
  if (dist >= r + SOL_GEOMETRIC_EPSILON) return 0;
  for (int k = 0; k < 3; k ++)
  {
    normal[k] = n[k];
    spnt[k] = c[k] - dist*n[k];
    mpnt[k] = c[k] - r*n[k];
  }
  *gap = dist - r < 0.0 ? dist - r : 0.0;
  return 1;
}

/* plane (slave) and ellipsoid (master, frame Q, semi-axes S): the ellipsoid's
 * support point against the plane normal (goc.c:350-404 with the plane as the
 * convex) */
SOL_GEOM_FN int geom_plane_ellip(const double *P, const double *c, const double *Q, const double *S,
                                 double *spnt, double *mpnt, double *normal, double *gap)
{
  const double *n = P + 3;
  double o[3];
  double h = geom_ell_support(Q, S, n, o);
  double dist = (c[0] - P[0])*n[0] + (c[1] - P[1])*n[1] + (c[2] - P[2])*n[2], g = dist - h;
  if (g >= SOL_GEOMETRIC_EPSILON) return 0;
  for (int k = 0; k < 3; k ++)
  {
    normal[k] = n[k];
    mpnt[k] = c[k] - o[k];
    spnt[k] = mpnt[k] - g*n[k];
  }
  *gap = g < 0.0 ? g : 0.0;
  return 1;
}

/* begin of a rigid time step, scheme RIG_NEG without point forces (gravity
 * only): bod.c:1313-1373.  W = velo[0..2] (spatial angular velocity in the
 * body frame), v = velo[3..5]; A carries exp[-(h/2)W] J W + h T to the end. */
SOL_GEOM_FN void geom_rig_step_begin(double h, const double *gravity, double mass, double minv,
                                     const double *J, const double *I, double *velo, double *velo0,
                                     double *R, double *x, double *A)
{
  double half = 0.5*h, *W = velo, *v = velo + 3;
  for (int k = 0; k < 6; k ++) velo0[k] = velo[k];
  double O[3] = {half*W[0], half*W[1], half*W[2]}, DR[9], JW[3], DJW[3], W05[3], T[3] = {0.0, 0.0, 0.0}, f[3];
  geom_expmap(O, DR);
  geom_mm(R, DR, R);                                /* R(t+h/2) = R(t) exp[(h/2) W(t)] */
  for (int k = 0; k < 3; k ++) x[k] += half*v[k];   /* x(t+h/2) = x(t) + (h/2) v(t) */
  for (int k = 0; k < 3; k ++) f[k] = mass*gravity[k];   /* no torque loads: T = 0 */
  geom_mv(J, W, JW);
  geom_mtv(DR, JW, DJW);                            /* exp[-(h/2)W] J W */
  for (int k = 0; k < 3; k ++) A[k] = DJW[k] + h*T[k];
  for (int k = 0; k < 3; k ++) O[k] = half*T[k] + DJW[k];
  geom_mv(I, O, W05);                               /* W05 = I [exp[-(h/2)W] J W + (h/2) T] */
  double JW05[3], c[3], dW[3];
  geom_mv(J, W05, JW05);
  geom_cross(W05, JW05, c);
  for (int k = 0; k < 3; k ++) T[k] -= c[k];        /* T - W05 x J W05 */
  geom_mv(I, T, dW);
  for (int k = 0; k < 3; k ++) { W[k] += h*dW[k]; v[k] += h*minv*f[k]; }   /* u += h iM f */
}

/* end of a rigid time step (bod.c:1425-1492), the velocity already updated by
 * the reactions; r[0..2] is the reactions' torque */
SOL_GEOM_FN void geom_rig_step_end(double h, const double *I, const double *r, double *velo,
                                   double *R, double *x, double *A)
{
  double half = 0.5*h, *W = velo, *v = velo + 3;
  for (int k = 0; k < 3; k ++) x[k] += half*v[k];
  double O[3] = {half*W[0], half*W[1], half*W[2]}, DR[9], t[3];
  geom_expmap(O, DR);
  geom_mm(R, DR, R);
  for (int k = 0; k < 3; k ++) A[k] += h*r[k];
  geom_mtv(DR, A, t);
  geom_mv(I, t, W);   /* W(t+h) = I exp[-(h/2)W1][exp[-(h/2)W(t)] J W(t) + h T] */
}

/* begin of a quasi-static rigid step under gravity only (bod.c:1588-1614 with
 * rig_static_force, bod.c:357-386): u(t+h) = h M^-1 f(t+h), where the
 * angular part of f is the gyroscopic term -W1 x J W1, W1 = I exp[-hW] J W
 * (no torque loads), and the linear part m g.  The previous velocity is not
 * carried, and velo0 is left alone. */
SOL_GEOM_FN void geom_rig_static_begin(double h, const double *gravity, double mass, double minv,
                                       const double *J, const double *I, double *velo)
{
  double *W = velo, *v = velo + 3;
  double O[3] = {h*W[0], h*W[1], h*W[2]}, DR[9], JW[3], DJW[3], W1[3], JW1[3], c[3], T[3], f[3], dW[3];
  geom_expmap(O, DR);                               /* exp[h W(t)] */
  for (int k = 0; k < 3; k ++) { T[k] = 0.0; f[k] = mass*gravity[k]; }
  geom_mv(J, W, JW);
  geom_mtv(DR, JW, DJW);                            /* exp[-hW] J W */
  for (int k = 0; k < 3; k ++) O[k] = h*T[k] + DJW[k];
  geom_mv(I, O, W1);                                /* W1 = I [exp[-hW] J W + h T] */
  geom_mv(J, W1, JW1);
  geom_cross(W1, JW1, c);
  for (int k = 0; k < 3; k ++) T[k] -= c[k];        /* T - W1 x J W1 */
  geom_mv(I, T, dW);
  for (int k = 0; k < 3; k ++) { W[k] = h*dW[k]; v[k] = h*(minv*f[k]); }   /* u = h iM f */
}

/* end of a quasi-static rigid step (bod.c:1616-1645), the velocity already
 * updated by the reactions: x(t+h) = x(t) + h v(t+h), R(t+h) = R(t) exp[h W(t+h)] */
SOL_GEOM_FN void geom_rig_static_end(double h, const double *velo, double *R, double *x)
{
  const double *W = velo, *v = velo + 3;
  for (int k = 0; k < 3; k ++) x[k] += h*v[k];
  double O[3] = {h*W[0], h*W[1], h*W[2]}, DR[9];
  geom_expmap(O, DR);
  geom_mm(R, DR, R);
}

/* grid cell coordinate of a position v (in cells from the grid's corner),
 * clamped to [0, dim - 1]; a non-finite position falls into cell 0, where the
 * distance tests reject it */
SOL_GEOM_FN int geom_cell_coord(double v, int dim)
{
  double f = floor(v);
  if (!(f >= 0.0)) return 0;
  if (f > (double)(dim - 1)) return dim - 1;
  return (int)f;
}

/* cells of the broad-phase grid over the box [lo, hi]: the cell size starts at
 * *cell and doubles until the cell count is at most cap (<= 2^31 - 1), so that
 * every dimension, the count and a cell index fit an int however far apart the
 * bodies are.  Counts are formed in double, where a product of large
 * dimensions cannot wrap.  Returns -1 for a box that is not finite. */
SOL_GEOM_FN int geom_grid_dims(const double *lo, const double *hi, long long cap, double *cell, int *dim,
                               long long *ncell)
{
  for (int k = 0; k < 3; k ++)
  {
    const double ext = hi[k] - lo[k];
    if (!(ext >= 0.0) || !(ext <= 1e300)) return -1;
  }
  for (;;)
  {
    double d[3], count = 1.0;
    for (int k = 0; k < 3; k ++)
    {
      d[k] = floor((hi[k] - lo[k])/(*cell)) + 1.0;
      count *= d[k];
    }
    if (count <= (double)cap)
    {
      for (int k = 0; k < 3; k ++) dim[k] = (int)d[k];
      *ncell = (long long)count;
      return 0;
    }
    *cell *= 2.0;
  }
}

/* key of an unordered body pair and its hash (splitmix64 finaliser) */
SOL_GEOM_FN unsigned long long geom_pair_key(int a, int b)
{
  unsigned lo = (unsigned)(a < b ? a : b), hi = (unsigned)(a < b ? b : a);
  return ((unsigned long long)lo << 32) | hi;
}

SOL_GEOM_FN unsigned long long geom_mix64(unsigned long long z)
{
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
  return z ^ (z >> 31);
}

#endif
