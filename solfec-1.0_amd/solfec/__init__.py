"""Python surface of the MI355X contact-solve build of Solfec-1.0.

Names, keywords, defaults and members follow the reference input language
(lng.c): ``SOLFEC`` (lng.c:1719), ``SURFACE_MATERIAL`` (lng.c:2128),
``BULK_MATERIAL`` (lng.c:2389), ``SPHERE`` (lng.c:1454), ``BODY`` (lng.c:2936),
``GRAVITY`` (lng.c:7232), ``INITIAL_VELOCITY`` (lng.c:7784),
``GAUSS_SEIDEL_SOLVER`` (lng.c:4173-4641) and ``RUN`` (lng.c:8973-9025).
Only rigid spheres, fixed half-space obstacles (``PLANE``, an addition: the
reference builds boxes from convex obstacles) and the Signorini-Coulomb model
are supported: the scope of this build is the contact solve, which runs on the
GPU through libsfgs.so.  ``RUN`` raises ``RuntimeError`` with the reference's
error text where the reference would THROW.
"""
import ctypes as C

import numpy as np

from ._lib import SFGS_NEWTON_REPORT, SFGS_REPORT, f64, ptr, solfec_lib
from .device import (DIAGSOLVERS, FAILURES, GS_ERRORS, Device, SfgsError, make_param, default_diag,  # noqa: F401
                     newton_param)

__all__ = ["SOLFEC", "SURFACE_MATERIAL", "BULK_MATERIAL", "SPHERE", "ELLIP", "PLANE", "BODY", "GRAVITY",
           "INITIAL_VELOCITY", "GAUSS_SEIDEL_SOLVER", "PENALTY_SOLVER", "NEWTON_SOLVER", "RUN", "LATTICE", "ELLIP_LATTICE", "LOCDYN_DUMP",
           "W_MATRIX_MARKET", "Device", "SfgsError"]

_CALLBACK_T = C.CFUNCTYPE(C.c_int, C.c_void_p)


class SOLFEC:
    """SOLFEC(analysis, step, output) with analysis 'DYNAMIC' or 'QUASI_STATIC' (lng.c:1719-1767).

    QUASI_STATIC integrates the rigid bodies with the static scheme (bod.c:1588-1645:
    u(t+h) = h M^-1 f(t+h), no carried velocity) and assembles the local dynamics with
    the static gap shift max(gap, 0)/h and no restitution."""

    def __init__(self, analysis, step, output):
        if analysis not in ("DYNAMIC", "QUASI_STATIC"):
            raise ValueError("Invalid analysis kind")
        if not step > 0:
            raise ValueError("'step' must be positive")
        self.lib = solfec_lib()
        self.analysis = analysis
        self.sol = self.lib.SOLFEC_Create(1 if analysis == "DYNAMIC" else 0, float(step), str(output).encode())
        if not self.sol:
            raise RuntimeError("SOLFEC_Create failed")
        self.step = float(step)
        self._keep = []

    def __del__(self):
        try:
            if self.sol:
                self.lib.SOLFEC_Destroy(self.sol)
                self.sol = None
        except Exception:
            pass

    @property
    def time(self):
        return self.lib.SOLFEC_Time(self.sol)

    @property
    def nbod(self):
        return self.lib.SOLFEC_Nbod(self.sol)

    @property
    def ncon(self):
        return self.lib.SOLFEC_Ncon(self.sol)

    def set_velocity(self, body, velo6):
        """Velocity [W v] (6 values) of the body with index `body` (SOLFEC_Set_Velocity)."""
        u = f64(velo6, 6)
        if self.lib.SOLFEC_Set_Velocity(self.sol, int(body), ptr(u)):
            raise ValueError("invalid body")

    def write_state(self, path):
        """State file for restart (SOLFEC_Write_State): the body and constraint records
        of BODY_Write_State / write_constraint, XDR-encoded."""
        if self.lib.SOLFEC_Write_State(self.sol, str(path).encode()):
            raise RuntimeError(self.lib.SOLFEC_Last_Error(self.sol).decode())

    def read_state(self, path):
        """Restart from a state file written by write_state for the same scene."""
        if self.lib.SOLFEC_Read_State(self.sol, str(path).encode()):
            raise RuntimeError(self.lib.SOLFEC_Last_Error(self.sol).decode())

    def device_loop(self, on=True):
        """Keep bodies and contacts on the GPU across steps (SOLFEC_Device_Loop):
        the half steps, contact update and detection run on the device; reading
        bodies() or contacts() downloads the current state."""
        rc = self.lib.SOLFEC_Device_Loop(self.sol, 1 if on else 0)
        if rc:
            raise RuntimeError(self.lib.SOLFEC_Last_Error(self.sol).decode())

    def timing(self, label):
        """Accumulated wall-clock seconds of a phase (TIMING, lng.c:10892)."""
        return self.lib.SOLFEC_Timing(self.sol, label.encode())

    # ---- state export used by tests and the benchmark -----------------
    def bodies(self):
        nb = self.nbod
        out = dict(kind=np.zeros(nb, np.int32), rot=np.zeros(9 * nb), x=np.zeros(3 * nb), x0=np.zeros(3 * nb),
                   jinv=np.zeros(9 * nb), minv=np.zeros(nb), velo=np.zeros(6 * nb), velo0=np.zeros(6 * nb),
                   radius=np.zeros(nb))
        if self.lib.SOLFEC_Get_Bodies(self.sol, *[ptr(out[k]) for k in
                                                  ("kind", "rot", "x", "x0", "jinv", "minv", "velo", "velo0",
                                                   "radius")]):
            raise RuntimeError(self.lib.SOLFEC_Last_Error(self.sol).decode())
        return out

    def contacts(self):
        n = self.ncon
        out = dict(master=np.zeros(n, np.int32), slave=np.zeros(n, np.int32), mpnt=np.zeros(3 * n),
                   spnt=np.zeros(3 * n), base=np.zeros(9 * n), gap=np.zeros(n), friction=np.zeros(n),
                   restitution=np.zeros(n), cohesion=np.zeros(n), R=np.zeros(3 * n), U=np.zeros(3 * n))
        if self.lib.SOLFEC_Get_Contacts(self.sol, *[ptr(out[k]) for k in
                                                    ("master", "slave", "mpnt", "spnt", "base", "gap", "friction",
                                                     "restitution", "cohesion", "R", "U")]):
            raise RuntimeError(self.lib.SOLFEC_Last_Error(self.sol).decode())
        return out

    def set_RU(self, R=None, U=None):
        n = self.ncon
        self.lib.SOLFEC_Set_RU(self.sol, ptr(None if R is None else f64(R, 3 * n)),
                               ptr(None if U is None else f64(U, 3 * n)))

    def step_begin(self):
        return self.lib.SOLFEC_Step_Begin(self.sol)

    def step_solve(self, solver):
        if isinstance(solver, NEWTON_SOLVER):
            return self.lib.SOLFEC_Step_Solve_Newton(self.sol, solver.ns)
        if isinstance(solver, PENALTY_SOLVER):
            return self.lib.SOLFEC_Step_Solve_Penalty(self.sol, solver.ps)
        return self.lib.SOLFEC_Step_Solve(self.sol, solver.gs)

    def step_end(self):
        return self.lib.SOLFEC_Step_End(self.sol)


class SURFACE_MATERIAL:
    """SURFACE_MATERIAL(solfec, surf1=None, surf2=None, model='SIGNORINI_COULOMB', label=None,
    friction=0.0, cohesion=0.0, restitution=0.0, spring=0.0, dashpot=0.0, hpow=1.0) (lng.c:2128-2190).

    model 'SPRING_DASHPOT' contacts are solved by the PENALTY_SOLVER; a
    GAUSS_SEIDEL_SOLVER step refuses a domain that has one.  spring, dashpot
    (negative: critical damping) and hpow (Hertz power) enter the
    PENALTY_SOLVER's law for every contact of the surface pair (pes.c:272)."""

    MODELS = {"SIGNORINI_COULOMB": 0, "SPRING_DASHPOT": 1}

    def __init__(self, solfec, surf1=None, surf2=None, model="SIGNORINI_COULOMB", label=None,
                 friction=0.0, cohesion=0.0, restitution=0.0, spring=0.0, dashpot=0.0, hpow=1.0):
        if model not in self.MODELS:
            raise ValueError("Unknown SURFACE_MATERIAL model")
        for v, k in ((friction, "friction"), (cohesion, "cohesion"), (spring, "spring"), (hpow, "hpow")):
            if not v >= 0:
                raise ValueError(f"'{k}' must be non-negative")
        if not 0 <= restitution <= 1:
            raise ValueError("'restitution' must be in [0, 1]")
        s1 = -1 if surf1 is None else int(surf1)
        s2 = -1 if surf2 is None else int(surf2)
        if (s1 < 0) != (s2 < 0):
            raise ValueError("both surf1 and surf2 must be given")
        rc = solfec.lib.SOLFEC_Surface_Material_Ex(solfec.sol, s1, s2, self.MODELS[model], float(friction),
                                                   float(restitution), float(cohesion), float(spring),
                                                   float(dashpot), float(hpow))
        if rc:
            raise ValueError("invalid surface material parameters")
        self.friction, self.cohesion, self.restitution = friction, cohesion, restitution
        self.spring, self.dashpot, self.hpow = spring, dashpot, hpow
        self.model, self.label = model, label


class BULK_MATERIAL:
    """BULK_MATERIAL(solfec, model='KIRCHHOFF', label=None, young=1E6, poisson=0.25, density=1E3)."""

    def __init__(self, solfec, model="KIRCHHOFF", label=None, young=1e6, poisson=0.25, density=1e3):
        if not density > 0:
            raise ValueError("'density' must be positive")
        self.model, self.label, self.young, self.poisson, self.density = model, label, young, poisson, density


class SPHERE:
    """SPHERE(center, radius, volid, surfid) (lng.c:1454)."""

    def __init__(self, center, radius, volid, surfid):
        if not radius > 0:
            raise ValueError("'radius' must be positive")
        self.center, self.radius, self.volid, self.surfid = tuple(map(float, center)), float(radius), volid, surfid


class ELLIP:
    """ELLIP(center, radii, volid, surfid) (lng.c:1560-1600); this build adds
    rotation: a column-major 3x3 rotation of the principal axes (default identity)."""

    def __init__(self, center, radii, volid, surfid, rotation=None):
        radii = tuple(map(float, radii))
        if len(radii) != 3 or not all(r > 0 for r in radii):
            raise ValueError("'radii' must be three positive numbers")
        self.center, self.radii, self.volid, self.surfid = tuple(map(float, center)), radii, volid, surfid
        self.rotation = None if rotation is None else np.asarray(rotation, np.float64).reshape(9)


class PLANE:
    """PLANE(point, normal, surfid): fixed half-space {x : (x - point).normal >= 0}, obstacles only."""

    def __init__(self, point, normal, surfid):
        self.point, self.normal, self.surfid = tuple(map(float, point)), tuple(map(float, normal)), surfid


class BODY:
    """BODY(solfec, kind, shape, material, label=None) with kind 'RIGID' (SPHERE, ELLIP) or 'OBSTACLE' (PLANE)."""

    def __init__(self, solfec, kind, shape, material, label=None):
        lib = solfec.lib
        if kind == "RIGID":
            if isinstance(shape, SPHERE):
                c = f64(shape.center, 3)
                self.id = lib.SOLFEC_Sphere(solfec.sol, ptr(c), shape.radius, float(material.density),
                                            int(shape.surfid))
            elif isinstance(shape, ELLIP):
                c, r = f64(shape.center, 3), f64(shape.radii, 3)
                rot = None if shape.rotation is None else f64(shape.rotation, 9)
                self.id = lib.SOLFEC_Ellipsoid(solfec.sol, ptr(c), ptr(r), ptr(rot), float(material.density),
                                               int(shape.surfid))
            else:
                raise ValueError("this build supports RIGID spheres and ellipsoids only")
        elif kind == "OBSTACLE":
            if not isinstance(shape, PLANE):
                raise ValueError("this build supports PLANE obstacles only")
            p, n = f64(shape.point, 3), f64(shape.normal, 3)
            self.id = lib.SOLFEC_Plane(solfec.sol, ptr(p), ptr(n), int(shape.surfid))
        else:
            raise ValueError("Invalid body kind")
        if self.id < 0:
            raise ValueError("invalid body")
        self.solfec, self.kind, self.label = solfec, kind, label


def GRAVITY(solfec, vector):
    """GRAVITY(solfec, vector) with a constant vector (lng.c:7232)."""
    g = tuple(map(float, vector))
    solfec.lib.SOLFEC_Gravity(solfec.sol, *g)


def INITIAL_VELOCITY(body, linear, angular):
    """INITIAL_VELOCITY(body, linear, angular) (lng.c:7784); angular is spatial for spheres."""
    u = f64(list(angular) + list(linear), 6)
    if body.solfec.lib.SOLFEC_Set_Velocity(body.solfec.sol, body.id, ptr(u)):
        raise ValueError("invalid body")


def ELLIP_LATTICE(solfec, n, L, amin, amax, jitter, density, seed):
    """Seeded jittered cubic lattice of n rigid ellipsoids (semi-axes U[amin, amax], random
    orientations) in [0, L]^3 with 6 bounding planes."""
    if solfec.lib.SOLFEC_Ellipsoid_Lattice_Scene(solfec.sol, int(n), float(L), float(amin), float(amax),
                                                 float(jitter), float(density), int(seed)):
        raise ValueError("invalid lattice parameters")


def LATTICE(solfec, n, L, rmin, rmax, jitter, density, seed):
    """Seeded jittered cubic lattice of n rigid spheres in [0, L]^3 with 6 bounding planes."""
    if solfec.lib.SOLFEC_Lattice_Scene(solfec.sol, int(n), float(L), float(rmin), float(rmax), float(jitter),
                                       float(density), int(seed)):
        raise ValueError("invalid lattice parameters")


class NEWTON_SOLVER:
    """NEWTON_SOLVER(meritval=1E-8, maxiter=1000, locdyn='ON', linver='GMRES', linmaxiter=10,
    maxmatvec=10000, epsilon=0.25, delta=0.0, theta=0.25, omega=1E-10, gsflag='ON', reldelta='OFF')
    (lng.c:4740-4857): a projected semismooth Newton solve of all contacts at once
    (nts.c) on the GPU; linver 'GMRES' (flexible GMRES step) or 'DIAG' (per-row
    steps); gsflag 'ON' runs a Gauss-Seidel solve from the initial reactions when
    the Newton solve ends above meritval.  locdyn 'OFF' forms the W products
    through the bodies, H (h M^-1 H' x) (nts.c:396-431); a solve that ends
    above meritval with gsflag 'ON' then raises, where the reference aborts
    (nts.c:1370-1374).

    Read-only members: merhist, mvhist, iters, itershist (set 'ON' to collect).
    Read/write members: meritval, maxiter, locdyn, linver, linmaxiter, maxmatvec,
    epsilon, delta, theta, omega, gsflag."""

    _LINVER = {"GMRES": 0, "DIAG": 1}
    _RELDELTA = {"OFF": 0, "avgWii": 1, "minWii": 2, "maxWii": 3}
    _RO = ("merhist", "mvhist", "iters")

    def __init__(self, meritval=1e-8, maxiter=1000, locdyn=None, linver=None, linmaxiter=10, maxmatvec=10000,
                 epsilon=0.25, delta=0.0, theta=0.25, omega=1e-10, gsflag=None, reldelta=None):
        for v, k in ((meritval, "meritval"), (maxiter, "maxiter"), (linmaxiter, "linmaxiter"),
                     (maxmatvec, "maxmatvec"), (epsilon, "epsilon"), (omega, "omega")):
            if not v > 0:
                raise ValueError(f"'{k}' must be positive")
        if not delta >= 0:
            raise ValueError("'delta' must be non-negative")
        if not 0 < theta <= 1:
            raise ValueError("'theta' must be in (0, 1]")
        lib = solfec_lib()
        object.__setattr__(self, "lib", lib)
        ns = lib.NEWTON_Create(float(meritval), int(maxiter))
        if not ns:
            raise ValueError("invalid NEWTON_SOLVER parameters")
        object.__setattr__(self, "ns", ns)
        for k, v in (("linmaxiter", linmaxiter), ("maxmatvec", maxmatvec), ("epsilon", epsilon), ("delta", delta),
                     ("theta", theta), ("omega", omega)):
            lib.NEWTON_Set(ns, k.encode(), float(v))
        if locdyn is not None:
            self.locdyn = locdyn
        if linver is not None:
            self.linver = linver
        if gsflag is not None:
            self.gsflag = gsflag
        if reldelta is not None:
            if reldelta not in self._RELDELTA:
                raise ValueError("Invalid reldelta value: neither of {OFF, avgWii, minWii, maxWii} ")
            lib.NEWTON_Set(ns, b"reldelta", float(self._RELDELTA[reldelta]))

    def __del__(self):
        try:
            if self.ns:
                self.lib.NEWTON_Destroy(self.ns)
                object.__setattr__(self, "ns", None)
        except Exception:
            pass

    def _get(self, m):
        return self.lib.NEWTON_Get(self.ns, m.encode())

    def __getattr__(self, name):
        if name in ("ns", "lib"):
            raise AttributeError(name)
        lib, ns = self.lib, self.ns
        if name in ("meritval", "epsilon", "delta", "theta", "omega"):
            return self._get(name)
        if name in ("maxiter", "linmaxiter", "maxmatvec", "iters"):
            return int(self._get(name))
        if name == "locdyn":
            return "ON" if int(self._get("locdyn")) else "OFF"
        if name == "linver":
            return "DIAG" if int(self._get("linver")) == 1 else "GMRES"
        if name == "gsflag":
            return "ON" if int(self._get("gsflag")) else "OFF"
        if name in ("merhist", "mvhist"):
            n = int(self._get("iters"))
            a, b = np.zeros(max(n, 1)), np.zeros(max(n, 1), np.int32)
            lib.NEWTON_History(ns, ptr(a), ptr(b))
            return [float(x) for x in a[:n]] if name == "merhist" else [float(x) for x in b[:n]]
        if name == "itershist":
            n = lib.NEWTON_Itershist(ns, None, 0)
            if n <= 0:
                return None
            out = np.zeros(n, np.int32)
            lib.NEWTON_Itershist(ns, ptr(out), n)
            return [int(x) for x in out]
        raise AttributeError(name)

    def __setattr__(self, name, value):
        lib, ns = self.lib, self.ns
        if name in self._RO:
            raise ValueError("Writing to a read-only member")
        if name in ("meritval", "maxiter", "linmaxiter", "maxmatvec", "epsilon", "omega"):
            if not value > 0:
                raise ValueError(f"'{name}' must be positive")
            lib.NEWTON_Set(ns, name.encode(), float(value))
        elif name == "delta":
            if not value >= 0:
                raise ValueError("'delta' must be non-negative")
            lib.NEWTON_Set(ns, b"delta", float(value))
        elif name == "theta":
            if not 0 < value <= 1:
                raise ValueError("'theta' must be in (0, 1]")
            lib.NEWTON_Set(ns, b"theta", float(value))
        elif name == "locdyn":
            if value not in ("ON", "OFF"):
                raise ValueError("Invalid locdyn value: neither ON nor OFF")
            lib.NEWTON_Set(ns, b"locdyn", 1.0 if value == "ON" else 0.0)
        elif name == "linver":
            if value not in self._LINVER:
                raise ValueError("Invalid linver value: neither GMRES nor DIAG")
            lib.NEWTON_Set(ns, b"linver", float(self._LINVER[value]))
        elif name == "gsflag":
            if value not in ("ON", "OFF"):
                raise ValueError("Invalid gsflag value: neither ON nor OFF")
            lib.NEWTON_Set(ns, b"gsflag", 1.0 if value == "ON" else 0.0)
        elif name == "itershist":
            if value == "ON":
                lib.NEWTON_Set(ns, b"itershist", 1.0)
        else:
            raise AttributeError(name)

    def report(self):
        rep = SFGS_NEWTON_REPORT()
        self.lib.NEWTON_Report(self.ns, C.byref(rep))
        return rep


class PENALTY_SOLVER:
    """PENALTY_SOLVER(variant='IMPLICIT') (lng.c:4659-4698): every contact is a
    spring and a dashpot (pes.c:33-106) with the constants of its surface
    material; variant 'IMPLICIT' or 'EXPLICIT'."""

    def __init__(self, variant="IMPLICIT"):
        if variant not in ("IMPLICIT", "EXPLICIT"):
            raise ValueError("Invalid variant")
        self.lib = solfec_lib()
        self.variant = variant
        self.ps = self.lib.PENALTY_Create(1 if variant == "IMPLICIT" else 0)
        if not self.ps:
            raise MemoryError("PENALTY_Create failed")

    def __del__(self):
        try:
            if self.ps:
                self.lib.PENALTY_Destroy(self.ps)
                self.ps = None
        except Exception:
            pass


class GAUSS_SEIDEL_SOLVER:
    """GAUSS_SEIDEL_SOLVER(epsilon, maxiter, meritval=1.0, failure='CONTINUE', diagepsilon=None,
    diagmaxiter=None, diagsolver='SEMISMOOTH_NEWTON', data=None, callback=None) (lng.c:4243-4349).

    Read-only members: failure, error, iters, rerhist, merhist (and itershist).
    Read/write members: epsilon, maxiter, meritval, diagepsilon, diagmaxiter, diagsolver,
    reverse ('ON'/'OFF'), variant, innerloops, itershist (set 'ON'); and, for this
    build, precision ('FP64'/'MIXED'), refine (FP64 refinement sweeps of MIXED)
    and warmstart ('PREVIOUS' = previous-step R as the reference, 'LEARNED' =
    previous-step R corrected by the network loaded into the domain's device,
    see solfec.learned).
    """

    _RO = ("failure", "error", "iters", "rerhist", "merhist")

    def __init__(self, epsilon, maxiter, meritval=1.0, failure="CONTINUE", diagepsilon=None, diagmaxiter=None,
                 diagsolver="SEMISMOOTH_NEWTON", data=None, callback=None):
        for v, k in ((epsilon, "epsilon"), (maxiter, "maxiter"), (meritval, "meritval")):
            if not v > 0:
                raise ValueError(f"'{k}' must be positive")
        if diagepsilon is not None and not diagepsilon > 0:
            raise ValueError("'diagepsilon' must be positive")
        if diagmaxiter is not None and not diagmaxiter > 0:
            raise ValueError("'diagmaxiter' must be positive")
        if failure not in FAILURES:
            raise ValueError("Invalid failure action")
        if diagsolver not in DIAGSOLVERS:
            raise ValueError("Invalid diagonal solver")
        if failure == "CALLBACK" and callback is None:
            raise ValueError("A callback routine must be given with CALLBACK failure action")
        if callback is not None and not callable(callback):
            raise TypeError("'callback' must be callable")
        object.__setattr__(self, "_data", data)
        object.__setattr__(self, "_callback", callback)

        def _cb(_):
            args = (self,) + (tuple(data) if isinstance(data, tuple) else (() if data is None else (data,)))
            return int(callback(*args))

        object.__setattr__(self, "_cfun", _CALLBACK_T(_cb) if callback is not None else None)
        lib = solfec_lib()
        object.__setattr__(self, "lib", lib)
        de = -1.0 if diagepsilon is None else float(diagepsilon)
        dm = -1 if diagmaxiter is None else int(diagmaxiter)
        gs = lib.GAUSS_SEIDEL_Create(float(epsilon), int(maxiter), float(meritval), FAILURES[failure], de, dm,
                                     DIAGSOLVERS[diagsolver], None,
                                     C.cast(self._cfun, C.c_void_p) if self._cfun else None)
        if not gs:
            raise ValueError("invalid GAUSS_SEIDEL_SOLVER parameters")
        object.__setattr__(self, "gs", gs)

    def __del__(self):
        try:
            if self.gs:
                self.lib.GAUSS_SEIDEL_Destroy(self.gs)
                object.__setattr__(self, "gs", None)
        except Exception:
            pass

    def _get(self, m):
        return self.lib.GAUSS_SEIDEL_Get(self.gs, m.encode())

    def __getattr__(self, name):
        if name in ("gs", "lib"):
            raise AttributeError(name)
        lib, gs = self.lib, self.gs
        if name == "failure":
            return lib.GAUSS_SEIDEL_Failure(gs).decode()
        if name == "error":
            return lib.GAUSS_SEIDEL_Error(gs).decode()
        if name == "iters":
            return int(self._get("iters"))
        if name in ("rerhist", "merhist"):
            n = int(self._get("iters"))
            a, b = np.zeros(max(n, 1)), np.zeros(max(n, 1))
            lib.GAUSS_SEIDEL_History(gs, ptr(a), ptr(b))
            return list((a if name == "rerhist" else b)[:n])
        if name == "itershist":
            n = lib.GAUSS_SEIDEL_Itershist(gs, None, 0)
            if n <= 0:
                return None
            out = np.zeros(n, np.int32)
            lib.GAUSS_SEIDEL_Itershist(gs, ptr(out), n)
            return [int(x) for x in out]
        if name in ("epsilon", "meritval", "diagepsilon"):
            return self._get(name)
        if name in ("maxiter", "diagmaxiter", "innerloops", "refine"):
            return int(self._get(name))
        if name == "diagsolver":
            return lib.GAUSS_SEIDEL_Diagsolver(gs).decode()
        if name == "reverse":
            return lib.GAUSS_SEIDEL_Reverse(gs).decode()
        if name == "variant":
            return lib.GAUSS_SEIDEL_Variant(gs).decode()
        if name == "precision":
            return "MIXED" if int(self._get("precision")) == 1 else "FP64"
        if name == "warmstart":
            return "LEARNED" if int(self._get("warmstart")) == 1 else "PREVIOUS"
        raise AttributeError(name)

    def __setattr__(self, name, value):
        lib, gs = self.lib, self.gs
        if name in self._RO:
            raise ValueError("Writing to a read-only member")
        if name in ("epsilon", "maxiter", "meritval", "diagepsilon", "diagmaxiter"):
            if not value > 0:
                raise ValueError(f"'{name}' must be positive")
            lib.GAUSS_SEIDEL_Set(gs, name.encode(), float(value))
        elif name == "innerloops":
            if not value >= 1:
                raise ValueError("'innerloops' must be >= 1")
            lib.GAUSS_SEIDEL_Set(gs, b"innerloops", float(value))
        elif name == "diagsolver":
            if value not in DIAGSOLVERS:
                raise ValueError("Invalid diagonal solver")
            lib.GAUSS_SEIDEL_Set(gs, b"diagsolver", float(DIAGSOLVERS[value]))
        elif name == "reverse":
            if value not in ("ON", "OFF"):
                raise ValueError("Invalid reverse switch (ON/OFF accepted)")
            lib.GAUSS_SEIDEL_Set(gs, b"reverse", 1.0 if value == "ON" else 0.0)
        elif name == "variant":
            v = {"FULL": 0, "MIDDLE_JACOBI": 1, "BOUNDARY_JACOBI": 2}
            if value not in v:
                raise ValueError("Invalid variant")
            lib.GAUSS_SEIDEL_Set(gs, b"variant", float(v[value]))
        elif name == "itershist":
            if value == "ON":
                lib.GAUSS_SEIDEL_Set(gs, b"itershist", 1.0)
        elif name == "precision":
            if value not in ("FP64", "MIXED"):
                raise ValueError("Invalid precision (FP64/MIXED accepted)")
            lib.GAUSS_SEIDEL_Set(gs, b"precision", 1.0 if value == "MIXED" else 0.0)
        elif name == "refine":
            if not value >= 0:
                raise ValueError("'refine' must be >= 0")
            lib.GAUSS_SEIDEL_Set(gs, b"refine", float(value))
        elif name == "warmstart":
            if value not in ("PREVIOUS", "LEARNED"):
                raise ValueError("Invalid warm start (PREVIOUS/LEARNED accepted)")
            lib.GAUSS_SEIDEL_Set(gs, b"warmstart", 1.0 if value == "LEARNED" else 0.0)
        else:
            raise AttributeError(name)

    def report(self):
        rep = SFGS_REPORT()
        self.lib.GAUSS_SEIDEL_Report(self.gs, C.byref(rep))
        return rep


def LOCDYN_DUMP(solfec, path):
    """LOCDYN_DUMP(solfec, path) (lng.c, ldy.c:892-1002): text dump of the last assembled step."""
    if solfec.lib.SOLFEC_Locdyn_Dump(solfec.sol, str(path).encode()):
        raise RuntimeError(solfec.lib.SOLFEC_Last_Error(solfec.sol).decode())


def W_MATRIX_MARKET(solfec, path):
    """W of the last assembled step in MatrixMarket format (LOCDYN_W_MatrixMarket, ldy.c:1005-1052)."""
    if solfec.lib.SOLFEC_W_MatrixMarket(solfec.sol, str(path).encode()):
        raise RuntimeError(solfec.lib.SOLFEC_Last_Error(solfec.sol).decode())


def RUN(solfec, solver, duration):
    """RUN(solfec, solver, duration) (lng.c:8973-9025) with a GAUSS_SEIDEL_SOLVER,
    a PENALTY_SOLVER or a NEWTON_SOLVER (sol.c:469-473)."""
    if not isinstance(solfec, SOLFEC):
        raise TypeError("'solfec' must be a SOLFEC object")
    if not isinstance(solver, (GAUSS_SEIDEL_SOLVER, PENALTY_SOLVER, NEWTON_SOLVER)):
        raise TypeError("'solver' must be a constraint solver object")
    if not duration > 0:
        raise ValueError("'duration' must be positive")
    if isinstance(solver, NEWTON_SOLVER):
        rc = solfec.lib.SOLFEC_Run_Newton(solfec.sol, solver.ns, float(duration))
    elif isinstance(solver, PENALTY_SOLVER):
        rc = solfec.lib.SOLFEC_Run_Penalty(solfec.sol, solver.ps, float(duration))
    else:
        rc = solfec.lib.SOLFEC_Run(solfec.sol, solver.gs, float(duration))
    if rc:
        msg = solfec.lib.SOLFEC_Error_String(rc).decode()
        detail = solfec.lib.SOLFEC_Last_Error(solfec.sol).decode()
        raise RuntimeError(msg + (f": {detail}" if detail and rc in (4, 6, 7) else ""))
