/* or_ldy.c -- local dynamics (LOCDYN) assembly for rigid bodies.
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Follows:
 *   OR_Operator_H       bod.c:216-228 (rig_operator_H): H = [base' R skew(X-X0)' | base']
 *   inverse mass        bod.c:232-258 (rig_dynamic_inverse): blockdiag(J^-1, 1/m I)
 *   OR_Adjacency(_Fill) ldy.c:492-568 (LOCDYN_Insert): one block per shared body,
 *                       obstacles do not couple contacts (ldy.c:511, 539)
 *   V, B                ldy.c:364-397 (update_V_and_B) via bod.c:1817-1844
 *   W_ii, rho, A, E     ldy.c:654-736: W_ii = h (Hm iM Hm' + Hs iM Hs'), slave H negated
 *                       (ldy.c:685), obstacle H = 0 for contacts (bod.c:1852-1855),
 *                       rho = 1/lambda_max(W_ii) (dsyev, ldy.c:719-724),
 *                       A = inv(W_ii) (ldy.c:726-727), E = 0.5 sum B'AB
 *   W_ij                ldy.c:738-844: h H_i^(b) iM_b H_j^(b)' computed for the upper
 *                       triangle and transposed into the lower; here "upper" means
 *                       j > i by contact index instead of by pointer order.
 * OR_Eig3 stands for dsyev('N') (cyclic Jacobi rotations to convergence) and
 * OR_Inverse3 for dgetrf/dgetri (LU with partial pivoting).
 *
 * Parity: pinned in tests/test_oracle_ldy.py by the closed forms for a sphere on
 * a plane, W_ii = (h/m) diag(7/2, 7/2, 1), and a sphere pair, by a dense
 * brute-force h H iM H' over all bodies in numpy (every block, both triangles),
 * by numpy eigvalsh / inv for rho and A, and by W = W' and W >= 0. */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "oracle.h"
#include "or_la.h"

void OR_Operator_H(const OR_BODIES *bod, int b, const double *X, const double *base, double *H)
{
  const double *R = bod->rot + 9*b, *X0 = bod->x0 + 3*b;
  double A[3] = {X[0] - X0[0], X[1] - X0[1], X[2] - X0[2]}, S[9], T[9];
  or_skew(A, S);
  or_mmt(R, S, T);        /* T = R S' */
  or_mtm(base, T, H);     /* H_rot = base' T */
  for (int r = 0; r < 3; r ++)
    for (int c = 0; c < 3; c ++)
      H[9 + r + 3*c] = base[c + 3*r];  /* H_lin = base' */
}

/* y = H u for a 3x6 H (column-major) */
static void h_times_u(const double *H, const double *u, double *y)
{
  for (int r = 0; r < 3; r ++)
  {
    double s = 0.0;
    for (int k = 0; k < 6; k ++) s += H[r + 3*k] * u[k];
    y[r] = s;
  }
}

/* P = sign * H1 iM H2' for rigid body b (3x3, column-major) */
static void h_im_ht(const OR_BODIES *bod, int b, const double *H1, const double *H2, double sign, double *P)
{
  const double *J = bod->jinv + 9*b;
  double minv = bod->minv[b], G[18];
  /* G = iM H2' (6x3): rows 0..2 = J H2_rot', rows 3..5 = minv H2_lin' */
  for (int c = 0; c < 3; c ++)
  {
    for (int r = 0; r < 3; r ++)
    {
      double s = 0.0;
      for (int k = 0; k < 3; k ++) s += J[r + 3*k] * H2[c + 3*k];
      G[r + 6*c] = s;
      G[3 + r + 6*c] = minv * H2[9 + c + 3*r];
    }
  }
  for (int c = 0; c < 3; c ++)
    for (int r = 0; r < 3; r ++)
    {
      double s = 0.0;
      for (int k = 0; k < 6; k ++) s += H1[r + 3*k] * G[k + 6*c];
      P[r + 3*c] = sign * s;
    }
}

static int couples(const OR_BODIES *bod, int b) { return b >= 0 && bod->kind[b] != OR_OBS; }

/* body -> contacts incidence (ascending contact index) for coupling bodies */
static int incidence(const OR_BODIES *bod, const OR_CONTACTS *con, int **ptr_out, int **lst_out)
{
  int nb = bod->nb, nc = con->nc;
  int *ptr = calloc((size_t)nb + 1, sizeof(int)), *lst, *fill;
  if (!ptr) return -1;
  for (int i = 0; i < nc; i ++)
  {
    if (couples(bod, con->master[i])) ptr[con->master[i] + 1] ++;
    if (couples(bod, con->slave[i]) && con->slave[i] != con->master[i]) ptr[con->slave[i] + 1] ++;
  }
  for (int b = 0; b < nb; b ++) ptr[b + 1] += ptr[b];
  lst = malloc(sizeof(int) * (size_t)(ptr[nb] > 0 ? ptr[nb] : 1));
  fill = malloc(sizeof(int) * (size_t)(nb > 0 ? nb : 1));
  if (!lst || !fill) { free(ptr); free(lst); free(fill); return -1; }
  memcpy(fill, ptr, sizeof(int) * (size_t)nb);
  for (int i = 0; i < nc; i ++)
  {
    if (couples(bod, con->master[i])) lst[fill[con->master[i]] ++] = i;
    if (couples(bod, con->slave[i]) && con->slave[i] != con->master[i]) lst[fill[con->slave[i]] ++] = i;
  }
  free(fill);
  *ptr_out = ptr;
  *lst_out = lst;
  return 0;
}

long OR_Adjacency(const OR_BODIES *bod, const OR_CONTACTS *con, int *rowptr)
{
  int *ptr, *lst, nc = con->nc;
  if (incidence(bod, con, &ptr, &lst)) return -1;
  rowptr[0] = 0;
  for (int i = 0; i < nc; i ++)
  {
    int k = 0, m = con->master[i], s = con->slave[i];
    if (couples(bod, m)) k += ptr[m + 1] - ptr[m] - 1;
    if (couples(bod, s) && s != m) k += ptr[s + 1] - ptr[s] - 1;
    rowptr[i + 1] = rowptr[i] + k;
  }
  free(ptr);
  free(lst);
  return rowptr[nc];
}

void OR_Adjacency_Fill(const OR_BODIES *bod, const OR_CONTACTS *con, const int *rowptr, int *col, int *via)
{
  int *ptr, *lst, nc = con->nc;
  if (incidence(bod, con, &ptr, &lst)) return;
  for (int i = 0; i < nc; i ++)
  {
    int p = rowptr[i], side[2] = {con->master[i], con->slave[i]};
    for (int t = 0; t < 2; t ++)
    {
      int b = side[t];
      if (!couples(bod, b) || (t == 1 && b == side[0])) continue;
      for (int q = ptr[b]; q < ptr[b + 1]; q ++)
        if (lst[q] != i) { col[p] = lst[q]; via[p] = b; p ++; }
    }
  }
  free(ptr);
  free(lst);
}

/* cyclic Jacobi eigenvalues of a symmetric 3x3 (upper triangle is read) */
void OR_Eig3(const double *W, double *lambda)
{
  double a[3][3];
  for (int r = 0; r < 3; r ++)
    for (int c = r; c < 3; c ++)
      a[r][c] = a[c][r] = W[r + 3*c];
  for (int sweep = 0; sweep < 64; sweep ++)
  {
    double off = a[0][1]*a[0][1] + a[0][2]*a[0][2] + a[1][2]*a[1][2];
    double dia = a[0][0]*a[0][0] + a[1][1]*a[1][1] + a[2][2]*a[2][2];
    if (off <= 1e-34 * dia || off == 0.0) break;
    for (int p = 0; p < 2; p ++)
      for (int q = p + 1; q < 3; q ++)
      {
        if (a[p][q] == 0.0) continue;
        double theta = (a[q][q] - a[p][p]) / (2.0*a[p][q]);
        double t = (theta >= 0.0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta*theta + 1.0));
        double c = 1.0/sqrt(t*t + 1.0), s = t*c;
        for (int k = 0; k < 3; k ++)
        {
          double akp = a[k][p], akq = a[k][q];
          a[k][p] = c*akp - s*akq;
          a[k][q] = s*akp + c*akq;
        }
        for (int k = 0; k < 3; k ++)
        {
          double apk = a[p][k], aqk = a[q][k];
          a[p][k] = c*apk - s*aqk;
          a[q][k] = s*apk + c*aqk;
        }
      }
  }
  double e[3] = {a[0][0], a[1][1], a[2][2]};
  for (int i = 0; i < 2; i ++)
    for (int j = i + 1; j < 3; j ++)
      if (e[j] < e[i]) { double t = e[i]; e[i] = e[j]; e[j] = t; }
  lambda[0] = e[0]; lambda[1] = e[1]; lambda[2] = e[2];
}

int OR_Inverse3(const double *W, double *A)
{
  for (int c = 0; c < 3; c ++)
  {
    double M[9], x[3] = {c == 0, c == 1, c == 2};
    memcpy(M, W, sizeof(M));
    if (OR_Solve3(M, x)) return -1;
    A[3*c] = x[0]; A[3*c + 1] = x[1]; A[3*c + 2] = x[2];
  }
  return 0;
}

/* contact point of contact i on body b and the sign of its side */
static const double* side_point(const OR_CONTACTS *con, int i, int b, double *sign)
{
  if (b == con->master[i]) { *sign = 1.0; return con->mpnt + 3*i; }
  *sign = -1.0;
  return con->spnt + 3*i;
}

int OR_Locdyn_Assemble(const OR_BODIES *bod, const OR_CONTACTS *con, double step,
                       const int *rowptr, const int *col, const int *via,
                       double *W, double *Wd, double *A, double *B, double *V,
                       double *rho, double *free_energy)
{
  int nc = con->nc, fail = 0;
  double energy = 0.0;

  for (int i = 0; i < nc; i ++)
  {
    const double *base = con->base + 9*i;
    int m = con->master[i], s = con->slave[i];
    double Hm[18], Hs[18], x[3], P[9];

    /* V and B (ldy.c:364-397): master minus slave */
    OR_Operator_H(bod, m, con->mpnt + 3*i, base, Hm);
    h_times_u(Hm, bod->velo0 + 6*m, V + 3*i);
    h_times_u(Hm, bod->velo + 6*m, B + 3*i);
    if (s >= 0)
    {
      OR_Operator_H(bod, s, con->spnt + 3*i, base, Hs);
      h_times_u(Hs, bod->velo0 + 6*s, x);
      for (int k = 0; k < 3; k ++) V[3*i + k] -= x[k];
      h_times_u(Hs, bod->velo + 6*s, x);
      for (int k = 0; k < 3; k ++) B[3*i + k] -= x[k];
    }

    /* W_ii (ldy.c:654-714) */
    double *Wi = Wd + 9*i;
    for (int k = 0; k < 9; k ++) Wi[k] = 0.0;
    if (couples(bod, m))
    {
      h_im_ht(bod, m, Hm, Hm, 1.0, P);
      for (int k = 0; k < 9; k ++) Wi[k] += P[k];
    }
    if (s >= 0 && couples(bod, s))
    {
      h_im_ht(bod, s, Hs, Hs, 1.0, P);   /* (-Hs) iM (-Hs)' */
      for (int k = 0; k < 9; k ++) Wi[k] += P[k];
    }
    for (int k = 0; k < 9; k ++) Wi[k] *= step;

    double lam[3];
    OR_Eig3(Wi, lam);
    rho[i] = 1.0 / lam[2];
    if (OR_Inverse3(Wi, A + 9*i)) fail ++;
    or_mv(A + 9*i, B + 3*i, x);
    energy += or_dot(x, B + 3*i);
  }
  *free_energy = 0.5 * energy;

  /* off-diagonal blocks (ldy.c:738-844) */
  for (int i = 0; i < nc; i ++)
  {
    for (int p = rowptr[i]; p < rowptr[i + 1]; p ++)
    {
      int j = col[p], b = via[p], lo = i < j ? i : j, hi = i < j ? j : i;
      double sl, sh, Hl[18], Hh[18], P[9];
      const double *xl = side_point(con, lo, b, &sl), *xh = side_point(con, hi, b, &sh);
      OR_Operator_H(bod, b, xl, con->base + 9*lo, Hl);
      OR_Operator_H(bod, b, xh, con->base + 9*hi, Hh);
      h_im_ht(bod, b, Hl, Hh, sl*sh, P);   /* W_lo,hi / h */
      double *Wb = W + 9*(long)p;
      if (i == lo)
        for (int k = 0; k < 9; k ++) Wb[k] = step * P[k];
      else
        for (int r = 0; r < 3; r ++)
          for (int c = 0; c < 3; c ++)
            Wb[r + 3*c] = step * P[c + 3*r];   /* lower = transposed upper */
    }
  }
  return fail;
}
