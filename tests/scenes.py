"""Scene helpers shared by the tests: seeded lattice packings through the host
layer (scene generation and contact detection run on the CPU; no GPU)."""
import numpy as np

import solfec


def lattice(n=125, L=None, seed=1, friction=0.3, restitution=0.0, steps_begin=1, jitter=0.005, cohesion=0.0):
    side = int(np.ceil(round(n ** (1.0 / 3.0), 9)))
    L = 0.1 * side if L is None else L
    s = solfec.SOLFEC("DYNAMIC", 1e-3, "/tmp/solfec_test")
    solfec.SURFACE_MATERIAL(s, friction=friction, restitution=restitution, cohesion=cohesion)
    solfec.GRAVITY(s, (0.0, 0.0, -9.81))
    solfec.LATTICE(s, n, L, 0.045, 0.055, jitter, 2000.0, seed)
    for _ in range(steps_begin):
        s.step_begin()
    return s
