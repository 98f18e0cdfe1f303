"""ctypes front end of the CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker.  The product never imports it.
Build with ``make -C oracle``.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_i, _d, _p = C.c_int, C.c_double, C.c_void_p
_libpath = os.path.join(_HERE, "liboracle.so")


class OR_BODIES(C.Structure):
    _fields_ = [("nb", _i), ("kind", _p), ("rot", _p), ("x0", _p), ("jinv", _p), ("minv", _p),
                ("velo", _p), ("velo0", _p)]


class OR_CONTACTS(C.Structure):
    _fields_ = [("nc", _i), ("master", _p), ("slave", _p), ("mpnt", _p), ("spnt", _p), ("base", _p),
                ("gap", _p), ("friction", _p), ("restitution", _p), ("cohesion", _p)]


class OR_SYSTEM(C.Structure):
    _fields_ = [("n", _i), ("rowptr", _p), ("col", _p), ("W", _p), ("Wd", _p), ("A", _p), ("B", _p),
                ("V", _p), ("rho", _p), ("mu", _p), ("rest", _p), ("coh", _p), ("gap", _p),
                ("free_energy", _d), ("step", _d), ("dynamic", _i)]


class OR_GSPARAM(C.Structure):
    _fields_ = [("epsilon", _d), ("maxiter", _i), ("meritval", _d), ("failure", _i), ("diagepsilon", _d),
                ("diagmaxiter", _i), ("diagsolver", _i), ("reverse", _i)]


class OR_REPORT(C.Structure):
    _fields_ = [("iters", _i), ("error", _i), ("rel_error", _d), ("merit", _d), ("diag_fail", C.c_long),
                ("newton_iters", C.c_long), ("row_updates", C.c_long)]


class OR_NTPARAM(C.Structure):
    _fields_ = [("meritval", _d), ("maxiter", _i), ("linver", _i), ("linmaxiter", _i), ("maxmatvec", _i),
                ("epsilon", _d), ("delta", _d), ("theta", _d), ("omega", _d), ("gsflag", _i), ("reldelta", _i)]


class OR_NTREPORT(C.Structure):
    _fields_ = [("iters", _i), ("matvecs", _i), ("converged", _i), ("gs_fallback", _i), ("gs_iters", _i),
                ("singular", _i), ("merit", _d), ("W_norm", _d)]


_lib = None


def use_native_build():
    """Build the oracle for this host's own CPU (-O3 -march=native, make target
    `native`, into oracle/_native/) and load that build instead of the default
    one; returns the compiler flags used, or None when the build failed (the
    default build is then kept).  Called by the CPU-baseline leg of bench.py
    before anything else loads the oracle."""
    import subprocess
    global _lib, _libpath
    try:
        subprocess.run(["make", "-s", "-C", _HERE, "native"], check=True, capture_output=True, timeout=300)
    except Exception:
        return None
    _libpath, _lib = os.path.join(_HERE, "_native", "liboracle.so"), None
    return "-O3 -march=native"


def lib():
    global _lib
    if _lib is None:
        path = _libpath
        if not os.path.exists(path):
            raise RuntimeError(f"{path} missing: run `make -C oracle`")
        L = C.CDLL(path)
        protos = {
            "OR_Adjacency": (C.c_long, [C.POINTER(OR_BODIES), C.POINTER(OR_CONTACTS), _p]),
            "OR_Adjacency_Fill": (None, [C.POINTER(OR_BODIES), C.POINTER(OR_CONTACTS), _p, _p, _p]),
            "OR_Locdyn_Assemble": (_i, [C.POINTER(OR_BODIES), C.POINTER(OR_CONTACTS), _d, _p, _p, _p,
                                        _p, _p, _p, _p, _p, _p, _p]),
            "OR_Operator_H": (None, [C.POINTER(OR_BODIES), _i, _p, _p, _p]),
            "OR_Eig3": (None, [_p, _p]),
            "OR_Inverse3": (_i, [_p, _p]),
            "OR_Diag_Solve": (_i, [_i, _d, _i, _i, _d, _d, _d, _d, _d, _d, _p, _p, _p, _p, _p]),
            "OR_Solve3": (_i, [_p, _p]),
            "OR_SCF_Project": (None, [_d, _d, _p, _p]),
            "OR_SCF_Linearize": (None, [_i, _d, _d, _d, _d, _d, _p, _p, _p, _d, _d, _p, _p, _p]),
            "OR_Merit": (_d, [C.POINTER(OR_SYSTEM), _p, _p, _i, _p]),
            "OR_GS_Solve": (_i, [C.POINTER(OR_SYSTEM), C.POINTER(OR_GSPARAM), _p, _p, _p, _p, _p,
                                 C.POINTER(OR_REPORT)]),
            "OR_GS_Sweep": (_d, [C.POINTER(OR_SYSTEM), C.POINTER(OR_GSPARAM), _p, _i, _p, _p, _p]),
            "OR_Hash32": (C.c_uint, [C.c_uint]),
            "OR_Colour": (_i, [_i, _p, _p, _p]),
            "OR_Colour_Order": (None, [_i, _p, _i, _p, _p]),
            "OR_Body_Reaction": (None, [C.POINTER(OR_BODIES), C.POINTER(OR_CONTACTS), _p, _p]),
            "OR_Body_Velocity_Update": (None, [C.POINTER(OR_BODIES), _p, _d, _p]),
            "OR_Sphere_Contact": (_i, [_p, _d, _p, _d, _p]),
            "OR_Plane_Contact": (_i, [_p, _p, _d, _p]),
            "OR_Detect": (C.c_longlong, [_i, _p, _p, _p, _p, _i, _p, _p, C.c_longlong, _p, _p, _p, _p]),
            "OR_Expmap": (None, [_p, _p]),
            "OR_Penalty_Contact": (_i, [_i, _d, _d, _d, _d, _d, _d, _p, _p, _p, _p, _p, _p]),
            "OR_Update_Cohesion": (None, [_i, _p, _p, _p]),
            "OR_Newton_Solve": (_i, [C.POINTER(OR_SYSTEM), C.POINTER(OR_NTPARAM), _p, _p, _p, _p, _p,
                                     C.POINTER(OR_NTREPORT)]),
            "OR_FlexGMRES_Dense": (_i, [_i, _p, _p, _p, _p, _i, _i, _i, _d, _p, _p]),
            "OR_Rig_Step_Begin": (None, [_i, _d, _p, _d, _d, _p, _p, _p, _p, _p, _p, _p]),
            "OR_Rig_Step_End": (None, [_i, _d, _p, _p, _p, _p, _p, _p]),
            "OR_Rig_Static_Begin": (None, [_i, _d, _p, _d, _d, _p, _p, _p]),
            "OR_Rig_Static_End": (None, [_i, _d, _p, _p, _p]),
        }
        for fn, (res, args) in protos.items():
            f = getattr(L, fn)
            f.restype, f.argtypes = res, args
        _lib = L
    return _lib


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _f(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _n(a):
    return np.ascontiguousarray(a, dtype=np.int32)


class Bodies:
    """Keeps the arrays alive while the C struct points at them."""

    def __init__(self, kind, rot, x0, jinv, minv, velo, velo0):
        self.a = dict(kind=_n(kind), rot=_f(rot), x0=_f(x0), jinv=_f(jinv), minv=_f(minv), velo=_f(velo),
                      velo0=_f(velo0))
        self.s = OR_BODIES(len(self.a["kind"]), *[_ptr(self.a[k]) for k in
                                                  ("kind", "rot", "x0", "jinv", "minv", "velo", "velo0")])


class Contacts:
    def __init__(self, master, slave, mpnt, spnt, base, gap, friction, restitution, cohesion):
        self.a = dict(master=_n(master), slave=_n(slave), mpnt=_f(mpnt), spnt=_f(spnt), base=_f(base), gap=_f(gap),
                      friction=_f(friction), restitution=_f(restitution), cohesion=_f(cohesion))
        self.s = OR_CONTACTS(len(self.a["master"]), *[_ptr(self.a[k]) for k in
                                                      ("master", "slave", "mpnt", "spnt", "base", "gap", "friction",
                                                       "restitution", "cohesion")])


def bodies_from(d):
    return Bodies(d["kind"], d["rot"], d["x0"], d["jinv"], d["minv"], d["velo"], d["velo0"])


def contacts_from(d):
    return Contacts(d["master"], d["slave"], d["mpnt"], d["spnt"], d["base"], d["gap"], d["friction"],
                    d["restitution"], d["cohesion"])


def assemble(bod, con, step):
    """LOCDYN on the CPU: returns a dict of natural-order arrays (see oracle.h)."""
    L = lib()
    n = con.s.nc
    rowptr = np.zeros(n + 1, np.int32)
    nnz = L.OR_Adjacency(C.byref(bod.s), C.byref(con.s), _ptr(rowptr))
    col = np.zeros(max(nnz, 1), np.int32)
    via = np.zeros(max(nnz, 1), np.int32)
    L.OR_Adjacency_Fill(C.byref(bod.s), C.byref(con.s), _ptr(rowptr), _ptr(col), _ptr(via))
    W = np.zeros(9 * max(nnz, 1))
    Wd, A, B, V, rho = np.zeros(9 * n), np.zeros(9 * n), np.zeros(3 * n), np.zeros(3 * n), np.zeros(n)
    E = C.c_double(0.0)
    fail = L.OR_Locdyn_Assemble(C.byref(bod.s), C.byref(con.s), step, _ptr(rowptr), _ptr(col), _ptr(via),
                                _ptr(W), _ptr(Wd), _ptr(A), _ptr(B), _ptr(V), _ptr(rho), C.byref(E))
    return dict(n=n, nnz=int(nnz), rowptr=rowptr, col=col[:nnz], via=via[:nnz], W=W[:9 * nnz], Wd=Wd, A=A, B=B,
                V=V, rho=rho, free_energy=E.value, singular=fail, step=step,
                mu=con.a["friction"], rest=con.a["restitution"], coh=con.a["cohesion"], gap=con.a["gap"])


class System:
    """OR_SYSTEM view over a dict produced by assemble() (or loaded from a fixture)."""

    def __init__(self, d, dynamic=1):
        self.d = {k: (np.ascontiguousarray(v) if isinstance(v, np.ndarray) else v) for k, v in d.items()}
        d = self.d
        for k in ("W", "Wd", "A", "B", "V", "rho", "mu", "rest", "coh", "gap"):
            d[k] = _f(d[k])
        d["rowptr"], d["col"] = _n(d["rowptr"]), _n(d["col"])
        Wp = d["W"] if d["W"].size else np.zeros(9)
        cp = d["col"] if d["col"].size else np.zeros(1, np.int32)
        self._keep = (Wp, cp)
        self.s = OR_SYSTEM(int(d["n"]), _ptr(d["rowptr"]), _ptr(cp), _ptr(Wp), _ptr(d["Wd"]), _ptr(d["A"]),
                           _ptr(d["B"]), _ptr(d["V"]), _ptr(d["rho"]), _ptr(d["mu"]), _ptr(d["rest"]),
                           _ptr(d["coh"]), _ptr(d["gap"]), float(d["free_energy"]), float(d["step"]), int(dynamic))


def default_diag(epsilon, maxiter, meritval=1.0):
    de = max(0.01 * min(epsilon, min(meritval, 1e-4)), np.finfo(np.float64).eps)
    return de, max(100, maxiter // 100)


def gs_param(epsilon, maxiter, meritval=1.0, failure=0, diagepsilon=None, diagmaxiter=None, diagsolver=2,
             reverse=0):
    de, dm = default_diag(epsilon, maxiter, meritval)
    return OR_GSPARAM(epsilon, maxiter, meritval, failure, de if diagepsilon is None else diagepsilon,
                      dm if diagmaxiter is None else diagmaxiter, diagsolver, reverse)


def colour(rowptr, col):
    n = len(rowptr) - 1
    rowptr, col = _n(rowptr), _n(col if len(col) else np.zeros(1))
    out = np.zeros(max(n, 1), np.int32)
    nc = lib().OR_Colour(n, _ptr(rowptr), _ptr(col), _ptr(out))
    return nc, out[:n]


def colour_order(colours, ncolours):
    n = len(colours)
    order = np.zeros(max(n, 1), np.int32)
    colptr = np.zeros(ncolours + 1, np.int32)
    lib().OR_Colour_Order(n, _ptr(_n(colours)), ncolours, _ptr(order), _ptr(colptr))
    return order[:n], colptr


def list_order(n):
    """The reference's LOCDYN list order: newest contact first (ldy.c:505-508)."""
    return np.arange(n - 1, -1, -1, dtype=np.int32)


def gs_solve(sysobj, param, order, R0, U0, history=True):
    n = sysobj.s.n
    R, U = _f(R0).copy(), _f(U0).copy()
    rer = np.zeros(param.maxiter) if history else None
    mer = np.zeros(param.maxiter) if history else None
    rep = OR_REPORT()
    order = _n(order)
    rc = lib().OR_GS_Solve(C.byref(sysobj.s), C.byref(param), _ptr(order), _ptr(R), _ptr(U), _ptr(rer), _ptr(mer),
                           C.byref(rep))
    if history:
        rer, mer = rer[:rep.iters], mer[:rep.iters]
    return rc, rep, R.reshape(n, 3), U.reshape(n, 3), rer, mer


def merit(sysobj, R, U, update_U=1):
    U = _f(U).copy()
    cm = np.zeros(sysobj.s.n)
    m = lib().OR_Merit(C.byref(sysobj.s), _ptr(_f(R)), _ptr(U), update_U, _ptr(cm))
    return m, U.reshape(-1, 3), cm


def diag_solve(solver, diagepsilon, diagmaxiter, W, B, V, U, R, friction, restitution=0.0, cohesion=0.0,
               gap=0.0, rho=None, step=1e-3, dynamic=1):
    W, B, V = _f(W), _f(B), _f(V)
    U, R = _f(U).copy(), _f(R).copy()
    if rho is None:
        rho = 1.0 / np.linalg.eigvalsh(W.reshape(3, 3).T).max()
    it = lib().OR_Diag_Solve(solver, diagepsilon, diagmaxiter, dynamic, step, friction, restitution, cohesion, gap,
                             rho, _ptr(W), _ptr(B), _ptr(V), _ptr(U), _ptr(R))
    return it, R, U


def body_update(bod, con, R, step):
    nb = bod.s.nb
    force = np.zeros(6 * nb)
    R = _f(R)
    lib().OR_Body_Reaction(C.byref(bod.s), C.byref(con.s), _ptr(R), _ptr(force))
    velo = bod.a["velo"].copy()
    lib().OR_Body_Velocity_Update(C.byref(bod.s), _ptr(force), step, _ptr(velo))
    return velo.reshape(nb, 6), force.reshape(nb, 6)


def sphere_contact(ca, ra, cb, rb):
    """(swap, slave point, master point, normal, gap) or None (goc.c:421-455)."""
    out = np.zeros(11)
    if not lib().OR_Sphere_Contact(_ptr(_f(ca)), ra, _ptr(_f(cb)), rb, _ptr(out)):
        return None
    return bool(out[0]), out[1:4], out[4:7], out[7:10], out[10]


def plane_contact(plane, c, r):
    """(plane point, sphere point, normal, gap) or None (goc.c:308-348)."""
    out = np.zeros(10)
    if not lib().OR_Plane_Contact(_ptr(_f(plane)), _ptr(_f(c)), r, _ptr(out)):
        return None
    return out[0:3], out[3:6], out[6:9], out[9]


def detect(shape, rad, plane, x, master_old=(), slave_old=()):
    """One contact pass by brute force: dict of master, slave, old (index of the kept
    contact or -1), spnt, mpnt (spatial), normal, gap."""
    shape, rad, plane, x = _n(shape), _f(rad), _f(plane), _f(x)
    mo, so = _n(np.asarray(master_old, np.int32)), _n(np.asarray(slave_old, np.int32))
    nb, nold = len(shape), len(mo)
    cap = 4 * nb + nold + 16
    while True:
        m, s, o = np.zeros(cap, np.int32), np.zeros(cap, np.int32), np.zeros(cap, np.int32)
        g = np.zeros(10 * cap)
        n = lib().OR_Detect(nb, _ptr(shape), _ptr(rad), _ptr(plane), _ptr(x), nold,
                            _ptr(mo if nold else np.zeros(1, np.int32)), _ptr(so if nold else np.zeros(1, np.int32)),
                            cap, _ptr(m), _ptr(s), _ptr(o), _ptr(g))
        if n <= cap:
            break
        cap = int(n)
    g = g[:10 * n].reshape(n, 10)
    return dict(master=m[:n], slave=s[:n], old=o[:n], spnt=g[:, 0:3], mpnt=g[:, 3:6], normal=g[:, 6:9], gap=g[:, 9])


def expmap(O):
    R = np.zeros(9)
    lib().OR_Expmap(_ptr(_f(O)), _ptr(R))
    return R


def rig_step_begin(kind, h, gravity, mass, minv, J, I, velo, rot, x):
    """BODY_Dynamic_Step_Begin of one rigid body: returns velo, velo0, rot, x, aux."""
    velo, rot, x = _f(velo).copy(), _f(rot).copy(), _f(x).copy()
    velo0, aux = np.zeros(6), np.zeros(3)
    lib().OR_Rig_Step_Begin(int(kind), h, _ptr(_f(gravity)), mass, minv, _ptr(_f(J)), _ptr(_f(I)), _ptr(velo),
                            _ptr(velo0), _ptr(rot), _ptr(x), _ptr(aux))
    return velo, velo0, rot, x, aux


def rig_step_end(kind, h, I, force, velo, rot, x, aux):
    """BODY_Dynamic_Step_End of one rigid body (velocity already updated): velo, rot, x, aux."""
    velo, rot, x, aux = _f(velo).copy(), _f(rot).copy(), _f(x).copy(), _f(aux).copy()
    lib().OR_Rig_Step_End(int(kind), h, _ptr(_f(I)), _ptr(_f(force)), _ptr(velo), _ptr(rot), _ptr(x), _ptr(aux))
    return velo, rot, x, aux


def rig_static_begin(kind, h, gravity, mass, minv, J, I, velo):
    """BODY_Static_Step_Begin of one rigid body: returns the velocity h M^-1 f(t+h)."""
    velo = _f(velo).copy()
    lib().OR_Rig_Static_Begin(int(kind), h, _ptr(_f(gravity)), mass, minv, _ptr(_f(J)), _ptr(_f(I)), _ptr(velo))
    return velo


def rig_static_end(kind, h, velo, rot, x):
    """BODY_Static_Step_End of one rigid body (velocity already updated): rot, x."""
    rot, x = _f(rot).copy(), _f(x).copy()
    lib().OR_Rig_Static_End(int(kind), h, _ptr(_f(velo)), _ptr(rot), _ptr(x))
    return rot, x


def penalty_contact(implicit, step, gap, spring, dashpot, hpow, friction, W, B, V, R0=(0.0, 0.0, 0.0),
                    U0=(0.0, 0.0, 0.0), cohesion=None):
    """PENALTY_Spring_Dashpot_Contact (pes.c:33-106) from the carried R0, U0: returns rc, R, U, or
    rc, R, U, cohesion after the law when a cohesion state is given."""
    R = _f(np.asarray(R0, float)).copy()
    U = _f(np.asarray(U0, float)).copy()
    c = None if cohesion is None else np.array([float(cohesion)])
    rc = lib().OR_Penalty_Contact(int(implicit), step, gap, spring, dashpot, hpow, friction, _ptr(_f(W)), _ptr(_f(B)),
                                  _ptr(_f(V)), _ptr(U), _ptr(R), None if c is None else _ptr(c))
    return (rc, R, U) if c is None else (rc, R, U, float(c[0]))


def update_cohesion(R, friction, cohesion):
    """update_cohesion (ldy.c:69-95): the cohesion states after a solve with reactions R (3 per contact)."""
    c = _f(np.asarray(cohesion, float)).copy()
    lib().OR_Update_Cohesion(len(c), _ptr(_f(R)), _ptr(_f(friction)), _ptr(c))
    return c


def nt_param(meritval=1e-8, maxiter=1000, linver="GMRES", linmaxiter=10, maxmatvec=10000, epsilon=0.25, delta=0.0,
             theta=0.25, omega=1e-10, gsflag="ON", reldelta="OFF"):
    """NEWTON_SOLVER parameters with the Python defaults of lng.c:4740-4857."""
    return OR_NTPARAM(meritval, maxiter, {"GMRES": 0, "DIAG": 1}[linver], linmaxiter, maxmatvec, epsilon, delta,
                      theta, omega, 1 if gsflag == "ON" else 0,
                      {"OFF": 0, "avgWii": 1, "minWii": 2, "maxWii": 3}[reldelta])


def newton_solve(sysobj, param, R0, U0=None, order=None):
    """NEWTON_Solve (nts.c:1333-1379): returns rep, R, U, merhist, mvhist."""
    n = sysobj.s.n
    R = _f(R0).copy()
    U = np.zeros(3 * n) if U0 is None else _f(U0).copy()
    mer, mv = np.zeros(param.maxiter), np.zeros(param.maxiter, np.int32)
    rep = OR_NTREPORT()
    order = None if order is None else _n(order)
    lib().OR_Newton_Solve(C.byref(sysobj.s), C.byref(param), _ptr(order), _ptr(R), _ptr(U), _ptr(mer), _ptr(mv),
                          C.byref(rep))
    return rep, R.reshape(n, 3), U.reshape(n, 3), mer[:rep.iters], mv[:rep.iters]


def flexgmres_dense(A, b, x0=None, P=None, kdim=20, maxiter=10, miniter=1, atol=0.0):
    """The Newton solve's flexible GMRES on a dense operator: returns iters, x, rnorm, matvecs."""
    A = _f(np.asarray(A, float))
    n = len(b)
    x = np.zeros(n) if x0 is None else _f(np.asarray(x0, float)).copy()
    rn, mv = C.c_double(0.0), C.c_int(0)
    Pp = None if P is None else _f(np.asarray(P, float))
    it = lib().OR_FlexGMRES_Dense(n, _ptr(A), _ptr(Pp), _ptr(_f(np.asarray(b, float))), _ptr(x), kdim, maxiter,
                                  miniter, atol, C.byref(rn), C.byref(mv))
    return it, x, rn.value, mv.value
