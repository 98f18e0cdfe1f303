/* or_scf.c -- friction-cone projection and the contact residual C(U,R).
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * OR_SCF_Project   follows scf.c:136-174 (SCF_Project)
 * OR_SCF_Linearize follows scf.c:28-133  (SCF_Linearize)
 *
 * Parity: pinned in tests/test_oracle_scf.py by the textbook second-order-cone
 * projection (closed form of Fukushima-Luo-Tseng spectral factorisation:
 * idempotence, membership of the cone, polar-cone residual, orthogonality)
 * and by C = 0 at known exact solutions of the Signorini-Coulomb law; X, Y are
 * pinned by central finite differences of C. */
#include <math.h>
#include "oracle.h"
#include "or_la.h"

/* spectral factorisation of S with respect to the cone |x_T| <= fri x_N written
 * in the "minus" orientation used by scf.c: m = g1 u1 + g2 u2 */
static void spectral(double fri, const double *S, double *l1, double *l2, double *u1, double *u2, double *slen)
{
  double onefri2 = 1.0 + fri*fri;
  double sl = sqrt(or_dot2(S, S));
  *slen = sl;
  *l1 = -(S[2] + fri*sl) / onefri2;
  *l2 = (sl - fri*S[2]) / onefri2;
  if (sl != 0.0)
  {
    u2[0] = S[0]/sl; u2[1] = S[1]/sl; u2[2] = -fri;
  }
  else
  {
    u2[0] = 1.0; u2[1] = 0.0; u2[2] = -fri;
  }
  u1[0] = -fri*u2[0]; u1[1] = -fri*u2[1]; u1[2] = -1.0;
}

void OR_SCF_Project(double friction, double cohesion, double *S, double *R)
{
  double T[3] = {S[0], S[1], S[2] + cohesion}, l1, l2, u1[3], u2[3], sl, g1, g2, m[3];

  spectral(friction, T, &l1, &l2, u1, u2, &sl);
  g1 = l1 > 0.0 ? l1 : 0.0;
  g2 = l2 > 0.0 ? l2 : 0.0;
  for (int k = 0; k < 3; k ++) m[k] = g1*u1[k] + g2*u2[k];
  /* R = S - m (S taken without the cohesion shift, scf.c:171-173) */
  double s0 = S[0], s1 = S[1], s2 = S[2];
  R[0] = s0 - m[0];
  R[1] = s1 - m[1];
  R[2] = s2 - m[2];
}

void OR_SCF_Linearize(int dynamic, double step, double fri, double res, double coh, double gap,
                      const double *V, const double *U, const double *R, double UT, double omega,
                      double *C, double *X, double *Y)
{
  double eps2 = omega*omega, fri2 = fri*fri, onefri2 = 1.0 + fri2;
  double udash, ulen, F[3], S[3], l1, l2, u1[3], u2[3], slen, sq1, sq2, g1, g2, m[3];

  if (dynamic) udash = U[2] + res * (V[2] < 0.0 ? V[2] : 0.0);
  else udash = (gap > 0.0 ? gap : 0.0)/step + U[2];
  ulen = sqrt(or_dot2(U, U) + eps2);

  F[0] = U[0];
  F[1] = U[1];
  F[2] = UT >= 0.0 ? udash + fri*UT : udash + fri*(ulen - omega);

  S[0] = R[0] - F[0];
  S[1] = R[1] - F[1];
  S[2] = R[2] - F[2] + coh;

  spectral(fri, S, &l1, &l2, u1, u2, &slen);
  sq1 = sqrt(l1*l1 + 4.0*eps2);
  sq2 = sqrt(l2*l2 + 4.0*eps2*fri2);
  g1 = 0.5*(sq1 + l1);
  g2 = 0.5*(sq2 + l2);
  for (int k = 0; k < 3; k ++) m[k] = g1*u1[k] + g2*u2[k];
  for (int k = 0; k < 3; k ++) C[k] = F[k] + m[k];

  if (X)
  {
    double dF[9] = {1.0, 0.0, fri*U[0]/ulen, 0.0, 1.0, fri*U[1]/ulen, 0.0, 0.0, 1.0};
    double dg1 = 0.5*(1.0 + l1/sq1), sdot = or_dot2(S, S), T[9];

    if (slen != 0.0)
    {
      double dg2 = 0.5*(1.0 + l2/sq2);
      double a = 0.5*(1.0 + (l2 + fri*l1)/(sq2 + fri*sq1));
      double b = (fri2*dg1 + dg2)/onefri2;
      double c = (fri*(dg1 - dg2))/onefri2;
      double d = (dg1 + fri2*dg2)/onefri2;
      Y[0] = a + (b - a)*S[0]*S[0]/sdot;
      Y[1] = (b - a)*S[1]*S[0]/sdot;
      Y[2] = c*S[0]/slen;
      Y[3] = Y[1];
      Y[4] = a + (b - a)*S[1]*S[1]/sdot;
      Y[5] = c*S[1]/slen;
      Y[6] = Y[2];
      Y[7] = Y[5];
      Y[8] = d;
    }
    else
    {
      Y[1] = Y[2] = Y[3] = Y[5] = Y[6] = Y[7] = 0.0;
      Y[0] = Y[4] = Y[8] = dg1;
    }
    or_mm(Y, dF, T);
    for (int k = 0; k < 9; k ++) X[k] = dF[k] - T[k];  /* X = [I - dm/dS] dF/dU */
  }
}
