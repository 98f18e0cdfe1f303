"""Scene helpers shared by the tests: seeded lattice packings through the host
layer (scene generation and contact detection run on the CPU; no GPU)."""
import numpy as np

import solfec


def lattice(n=125, L=None, seed=1, friction=0.3, restitution=0.0, steps_begin=1, jitter=0.005, cohesion=0.0,
            analysis="DYNAMIC"):
    side = int(np.ceil(round(n ** (1.0 / 3.0), 9)))
    L = 0.1 * side if L is None else L
    s = solfec.SOLFEC(analysis, 1e-3, "/tmp/solfec_test")
    solfec.SURFACE_MATERIAL(s, friction=friction, restitution=restitution, cohesion=cohesion)
    solfec.GRAVITY(s, (0.0, 0.0, -9.81))
    solfec.LATTICE(s, n, L, 0.045, 0.055, jitter, 2000.0, seed)
    for _ in range(steps_begin):
        s.step_begin()
    return s


def batch_scenes(nscenes, seed, lo=500, hi=50000):
    """Independent scenes with contact counts log-uniform in [lo, hi] and mu
    uniform in [0.1, 0.8]: each the 1000-sphere generator (jittered lattice of
    spacing 0.1 m, radii U[0.045, 0.055], density 2000, 6 planes) sized for its
    count (about 1.52 contacts per sphere at the first step), at its first step.
    Returns the list of (bodies, contacts) and the frictions."""
    rng = np.random.default_rng(seed)
    want = np.exp(rng.uniform(np.log(lo), np.log(hi), nscenes))
    mus = rng.uniform(0.1, 0.8, nscenes)
    out = []
    for k in range(nscenes):
        n = max(8, int(round(want[k] / 1.52)))
        s = lattice(n, seed=int(seed * 1000 + k), friction=float(mus[k]))
        out.append((s.bodies(), s.contacts()))
    return out, mus
