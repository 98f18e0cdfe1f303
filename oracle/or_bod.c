/* or_bod.c -- rigid-body velocity update by constraint reactions.
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * OR_Body_Reaction follows rig_constraints_force, bod.c:418-435 with
 * rig_constraints_force_accum, bod.c:389-416: r_b = sum over contacts of
 * +H' R (master side) or -H' R (slave side); obstacles take no contact force
 * (bod.c:431).  OR_Body_Velocity_Update follows bod.c:1446-1447:
 * u_b += h iM_b r_b.
 *
 * Parity: pinned in tests/test_oracle_bod.py -- momentum balance for an
 * isolated pair (equal and opposite linear impulses), the closed form for a
 * sphere pushed along the normal at its surface point (no torque), and a dense
 * numpy h iM H' R evaluation. */
#include "oracle.h"
#include "or_la.h"

void OR_Body_Reaction(const OR_BODIES *bod, const OR_CONTACTS *con, const double *R, double *force)
{
  for (int b = 0; b < 6*bod->nb; b ++) force[b] = 0.0;
  for (int i = 0; i < con->nc; i ++)
  {
    int side[2] = {con->master[i], con->slave[i]};
    for (int t = 0; t < 2; t ++)
    {
      int b = side[t];
      if (b < 0 || bod->kind[b] == OR_OBS) continue;
      double H[18], sign = t == 0 ? 1.0 : -1.0;
      OR_Operator_H(bod, b, t == 0 ? con->mpnt + 3*i : con->spnt + 3*i, con->base + 9*i, H);
      for (int k = 0; k < 6; k ++)
      {
        double s = H[3*k]*R[3*i] + H[3*k + 1]*R[3*i + 1] + H[3*k + 2]*R[3*i + 2];
        force[6*b + k] += sign * s;
      }
    }
  }
}

void OR_Body_Velocity_Update(const OR_BODIES *bod, const double *force, double step, double *velo)
{
  for (int b = 0; b < bod->nb; b ++)
  {
    if (bod->kind[b] == OR_OBS) continue;
    double w[3];
    or_mv(bod->jinv + 9*b, force + 6*b, w);
    for (int k = 0; k < 3; k ++)
    {
      velo[6*b + k] += step * w[k];
      velo[6*b + 3 + k] += step * bod->minv[b] * force[6*b + 3 + k];
    }
  }
}
