/* or_la.h -- small column-major 3x3 / 3-vector helpers for the oracle.
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 * Layout follows the reference convention (alg.h:881-953): M[r + 3*c]. */
#ifndef SFGS_OR_LA_H
#define SFGS_OR_LA_H

#include <math.h>

static inline double or_dot(const double *a, const double *b) { return a[0]*b[0] + a[1]*b[1] + a[2]*b[2]; }
static inline double or_dot2(const double *a, const double *b) { return a[0]*b[0] + a[1]*b[1]; }

/* y = M x */
static inline void or_mv(const double *M, const double *x, double *y)
{
  double y0 = M[0]*x[0] + M[3]*x[1] + M[6]*x[2];
  double y1 = M[1]*x[0] + M[4]*x[1] + M[7]*x[2];
  double y2 = M[2]*x[0] + M[5]*x[1] + M[8]*x[2];
  y[0] = y0; y[1] = y1; y[2] = y2;
}

/* y = c + M x */
static inline void or_mv_add(const double *c, const double *M, const double *x, double *y)
{
  double y0 = c[0] + (M[0]*x[0] + M[3]*x[1] + M[6]*x[2]);
  double y1 = c[1] + (M[1]*x[0] + M[4]*x[1] + M[7]*x[2]);
  double y2 = c[2] + (M[2]*x[0] + M[5]*x[1] + M[8]*x[2]);
  y[0] = y0; y[1] = y1; y[2] = y2;
}

/* y = M^T x */
static inline void or_mtv(const double *M, const double *x, double *y)
{
  double y0 = M[0]*x[0] + M[1]*x[1] + M[2]*x[2];
  double y1 = M[3]*x[0] + M[4]*x[1] + M[5]*x[2];
  double y2 = M[6]*x[0] + M[7]*x[1] + M[8]*x[2];
  y[0] = y0; y[1] = y1; y[2] = y2;
}

/* C = A B */
static inline void or_mm(const double *A, const double *B, double *C)
{
  double T[9];
  for (int c = 0; c < 3; c ++)
    for (int r = 0; r < 3; r ++)
      T[r + 3*c] = A[r]*B[3*c] + A[r + 3]*B[3*c + 1] + A[r + 6]*B[3*c + 2];
  for (int i = 0; i < 9; i ++) C[i] = T[i];
}

/* C = A^T B */
static inline void or_mtm(const double *A, const double *B, double *C)
{
  double T[9];
  for (int c = 0; c < 3; c ++)
    for (int r = 0; r < 3; r ++)
      T[r + 3*c] = A[3*r]*B[3*c] + A[3*r + 1]*B[3*c + 1] + A[3*r + 2]*B[3*c + 2];
  for (int i = 0; i < 9; i ++) C[i] = T[i];
}

/* C = A B^T */
static inline void or_mmt(const double *A, const double *B, double *C)
{
  double T[9];
  for (int c = 0; c < 3; c ++)
    for (int r = 0; r < 3; r ++)
      T[r + 3*c] = A[r]*B[c] + A[r + 3]*B[c + 3] + A[r + 6]*B[c + 6];
  for (int i = 0; i < 9; i ++) C[i] = T[i];
}

/* skew matrix S with S x = a cross x */
static inline void or_skew(const double *a, double *S)
{
  S[0] = 0.0;   S[3] = -a[2]; S[6] = a[1];
  S[1] = a[2];  S[4] = 0.0;   S[7] = -a[0];
  S[2] = -a[1]; S[5] = a[0];  S[8] = 0.0;
}

#endif
