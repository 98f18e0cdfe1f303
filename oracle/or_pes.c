/* or_pes.c -- the spring-dashpot contact law of the penalty solver.
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * OR_Penalty_Contact follows PENALTY_Spring_Dashpot_Contact (pes.c:33-106) for a
 * contact that is not cohesive: the normal reaction from the spring
 * k (-g)^p and the dashpot c (critical when c < 0: 2 sqrt(h s / W_NN) with the
 * current stiffness s = k p (-g)^(p-1)), the gap g advanced by a quarter step of
 * the normal velocity in the IMPLICIT variant (pes.c:42-50) or taken as it is
 * in the EXPLICIT one (pes.c:51-57); a negative normal reaction means
 * separation, R = U = 0 (pes.c:59-64); the tangential reaction cancels the
 * tangential velocity through the 2x2 tangential block of W (INVERT2,
 * alg.h:995-1002) and is clamped to the Coulomb cone |R_T| <= mu |R_N|
 * (pes.c:84-101); U = B + W R.  The normal velocity uses the carried tangential
 * reaction, B_N + W_NT R_T (pes.c:40).  A cohesive contact (*coh > 0, the
 * cohesive force) is not separated by tension: its normal reaction is bounded
 * below by -*coh, and it loses its cohesion (*coh = 0) when that bound is hit
 * (mode I, pes.c:76-82) or when the tangential reaction is clamped to the cone
 * (mode II, pes.c:88-99).  coh may be NULL for a contact that is not cohesive.
 *
 * OR_Update_Cohesion follows update_cohesion (ldy.c:69-95): after a
 * Gauss-Seidel or Newton solve a cohesive contact loses its cohesion when
 * R_N + c < e (mode I) or |R_T| + e >= mu (R_N + c) (mode II), e = 1e-8 c
 * (COHESION_EPSILON, sps.h:34).
 *
 * Parity: pinned in tests/test_oracle_pes.py by closed forms (explicit and
 * implicit normal reaction by hand, separation, sticking with a diagonal W,
 * sliding on the cone, critical damping, the Hertz power, tension held and
 * broken by cohesion, both decohesion modes on either side of their bounds)
 * and by U = B + W R. */
#include <math.h>
#include <stddef.h>
#include "oracle.h"

int OR_Penalty_Contact(int implicit, double step, double gap, double spring, double dashpot, double hpow,
                       double friction, const double *W, const double *B, const double *V, double *U, double *R,
                       double *coh)
{
  int cohesive = coh && *coh > 0.0;
  const double BN = B[2] + W[2]*R[0] + W[5]*R[1];
  double g, s;
  if (implicit)
  {
    g = gap + 0.25*step*(BN - V[2]);
    if (g > 0.0) g = 0.0;
    s = spring*hpow*pow(-g, hpow - 1.0);
    if (dashpot < 0.0) dashpot = 2.0*sqrt(step*s/W[8]);
    R[2] = (spring*pow(-g, hpow) - 0.5*dashpot*(BN + V[2])) / (1.0 + (0.25*step*s + 0.5*dashpot)*W[8]);
  }
  else
  {
    g = gap < 0.0 ? gap : 0.0;
    s = spring*hpow*pow(-g, hpow - 1.0);
    if (dashpot < 0.0) dashpot = 2.0*sqrt(step*s/W[8]);
    R[2] = spring*pow(-g, hpow) - dashpot*V[2];
  }
  if (!cohesive && R[2] < 0.0)
  {
    R[0] = R[1] = R[2] = 0.0;
    U[0] = U[1] = U[2] = 0.0;
    return 0;
  }
  const double BT0 = B[0] + W[6]*R[2], BT1 = B[1] + W[7]*R[2];
  const double det = W[0]*W[4] - W[1]*W[3];
  if (det == 0.0) return -1;
  const double I0 = W[4]/det, I1 = -W[1]/det, I2 = -W[3]/det, I3 = W[0]/det;
  R[0] = -I0*BT0 - I2*BT1;
  R[1] = -I1*BT0 - I3*BT1;
  if (cohesive && R[2] < -*coh)
  {
    cohesive = 0;
    R[2] = -*coh;
    *coh = 0.0;
  }
  const double len = sqrt(R[0]*R[0] + R[1]*R[1]), lim = friction*fabs(R[2]);
  if (len > lim)
  {
    R[0] *= lim/len;
    R[1] *= lim/len;
    if (cohesive) *coh = 0.0;
  }
  for (int r = 0; r < 3; r ++) U[r] = B[r] + (W[r]*R[0] + W[r + 3]*R[1] + W[r + 6]*R[2]);
  return 0;
}

void OR_Update_Cohesion(int n, const double *R, const double *friction, double *coh)
{
  for (int i = 0; i < n; i ++)
  {
    const double c = coh[i];
    if (!(c > 0.0)) continue;
    const double e = 1e-8*c, *r = R + 3*(size_t)i;
    if (r[2] + c < e || sqrt(r[0]*r[0] + r[1]*r[1]) + e >= friction[i]*(r[2] + c)) coh[i] = 0.0;
  }
}
