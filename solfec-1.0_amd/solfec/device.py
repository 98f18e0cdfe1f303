"""Thin Python handle over one libsfgs.so context (one GPU W/B/R store).

Every method calls straight into the HIP library; there is no CPU path.
"""
import ctypes as C

import numpy as np

from ._lib import (SFGS_BATCH_TUNE, SFGS_NEWTON_PARAM, SFGS_NEWTON_REPORT, SFGS_PARAM, SFGS_REPORT,
                   SFGS_SCENE_REPORT, f64, i32, ptr, sfgs)

FP64, MIXED = 0, 1
DIAGSOLVERS = {"PROJECTED_GRADIENT": 0, "DE_SAXCE_FENG": 1, "SEMISMOOTH_NEWTON": 2, "PROJECTED_NEWTON": 3}
FAILURES = {"CONTINUE": 0, "EXIT": 1, "CALLBACK": 2}
GS_ERRORS = {0: "OK", 1: "DIVERGED", 2: "DIAGONAL_DIVERGED", 3: "DIAGONAL_FAILED"}


class SfgsError(RuntimeError):
    pass


def newton_param(meritval=1e-8, maxiter=1000, linver="GMRES", linmaxiter=10, maxmatvec=10000, epsilon=0.25,
                 delta=0.0, theta=0.25, omega=1e-10, gsflag="ON", reldelta="OFF"):
    """SFGS_NEWTON_PARAM with the defaults of NEWTON_SOLVER (lng.c:4740-4857)."""
    return SFGS_NEWTON_PARAM(meritval, maxiter, {"GMRES": 0, "DIAG": 1}[linver], linmaxiter, maxmatvec, epsilon,
                             delta, theta, omega, 1 if gsflag == "ON" else 0,
                             {"OFF": 0, "avgWii": 1, "minWii": 2, "maxWii": 3}[reldelta])


def default_diag(epsilon, maxiter, meritval=1.0):
    """diagepsilon / diagmaxiter defaults of GAUSS_SEIDEL_SOLVER (lng.c:4336-4342)."""
    de = 0.01 * min(epsilon, min(meritval, 1e-4))
    de = max(de, np.finfo(np.float64).eps)
    return de, max(100, maxiter // 100)


def make_param(epsilon, maxiter, meritval=1.0, failure=0, diagepsilon=None, diagmaxiter=None,
               diagsolver=2, reverse=0, precision=FP64, refine=2, profile=0):
    de, dm = default_diag(epsilon, maxiter, meritval)
    p = SFGS_PARAM()
    p.epsilon = epsilon
    p.maxiter = maxiter
    p.meritval = meritval
    p.failure = failure
    p.diagepsilon = de if diagepsilon is None else diagepsilon
    p.diagmaxiter = dm if diagmaxiter is None else diagmaxiter
    p.diagsolver = diagsolver
    p.reverse = reverse
    p.precision = precision
    p.refine = refine
    p.profile = profile
    return p


class Device:
    """One device context (``SFGS_Create``)."""

    def __init__(self, device=0):
        self.lib = sfgs()
        self.ctx = self.lib.SFGS_Create(device)
        if not self.ctx:
            raise SfgsError("SFGS_Create failed: no HIP device")

    def close(self):
        if self.ctx:
            self.lib.SFGS_Destroy(self.ctx)
            self.ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc, what):
        if rc < 0:
            msg = self.lib.SFGS_Last_Error(self.ctx).decode()
            raise SfgsError(f"{what}: {self.lib.SFGS_Error_String(rc).decode()} ({msg})")
        return rc

    def set_layout(self, lanes_per_row):
        self._check(self.lib.SFGS_Set_Layout(self.ctx, int(lanes_per_row)), "SFGS_Set_Layout")

    def set_row_order(self, order):
        """SFGS_Set_Row_Order: rows of a colour by 0 decreasing length, 1 contact index, 2 master body
        block, 3 (default) Z-order of the master body's position."""
        self._check(self.lib.SFGS_Set_Row_Order(self.ctx, int(order)), "SFGS_Set_Row_Order")

    def set_sweep_block(self, waves_per_block):
        """SFGS_Set_Sweep_Block: wavefronts per workgroup of the per-colour sweep launches (1, 2, 4;
        0, the default: 2 for FP64 solves, 4 for MIXED)."""
        self._check(self.lib.SFGS_Set_Sweep_Block(self.ctx, int(waves_per_block)), "SFGS_Set_Sweep_Block")

    def set_schedule(self, schedule):
        """0: one launch per colour; 1: persistent dataflow sweeps (same iterates)."""
        self._check(self.lib.SFGS_Set_Schedule(self.ctx, int(schedule)), "SFGS_Set_Schedule")

    # ---- inputs -------------------------------------------------------
    def bodies(self, kind, rot, x0, jinv, minv, velo, velo0):
        nb = len(kind)
        a = [i32(kind, nb), f64(rot, 9 * nb), f64(x0, 3 * nb), f64(jinv, 9 * nb), f64(minv, nb),
             f64(velo, 6 * nb), f64(velo0, 6 * nb)]
        self._check(self.lib.SFGS_Bodies_Upload(self.ctx, nb, *[ptr(x) for x in a]), "SFGS_Bodies_Upload")

    def contacts(self, master, slave, mpnt, spnt, base, gap, friction, restitution, cohesion, R, U):
        n = len(master)
        a = [i32(master, n), i32(slave, n), f64(mpnt, 3 * n), f64(spnt, 3 * n), f64(base, 9 * n), f64(gap, n),
             f64(friction, n), f64(restitution, n), f64(cohesion, n), f64(R, 3 * n), f64(U, 3 * n)]
        self._check(self.lib.SFGS_Contacts_Upload(self.ctx, n, *[ptr(x) for x in a]), "SFGS_Contacts_Upload")

    def assemble(self, step, dynamic=1, diagonal_only=False):
        """SFGS_Locdyn_Update_Begin, or SFGS_Locdyn_Update_Min (UPMIN: diagonal blocks only)."""
        if diagonal_only:
            self._check(self.lib.SFGS_Locdyn_Update_Min(self.ctx, step, dynamic), "SFGS_Locdyn_Update_Min")
        else:
            self._check(self.lib.SFGS_Locdyn_Update_Begin(self.ctx, step, dynamic), "SFGS_Locdyn_Update_Begin")

    def newton(self, param):
        """SFGS_Newton_Solve: returns the report, merhist and mvhist."""
        mer, mv = np.zeros(param.maxiter), np.zeros(param.maxiter, np.int32)
        rep = SFGS_NEWTON_REPORT()
        self._check(self.lib.SFGS_Newton_Solve(self.ctx, C.byref(param), ptr(mer), ptr(mv), C.byref(rep)),
                    "SFGS_Newton_Solve")
        return rep, mer[:rep.iters], mv[:rep.iters]

    def penalty(self, step, implicit=1, spring=None, dashpot=None, hpow=None):
        """SFGS_Penalty_Solve: the spring-dashpot law per uploaded contact; the
        three constants per contact, or all None for the material table."""
        n = self.sizes()[0]
        if spring is None and dashpot is None and hpow is None:
            rc = self.lib.SFGS_Penalty_Solve(self.ctx, float(step), int(implicit), None, None, None)
        else:
            a = [f64(spring, n), f64(dashpot, n), f64(hpow, n)]
            rc = self.lib.SFGS_Penalty_Solve(self.ctx, float(step), int(implicit), *[ptr(x) for x in a])
        self._check(rc, "SFGS_Penalty_Solve")

    def system(self, rowptr, col, W, Wd, A, B, V, rho, friction, restitution, cohesion, gap,
               free_energy, step, R, U, dynamic=1):
        n = len(rowptr) - 1
        rowptr = i32(rowptr, n + 1)
        nnz = int(rowptr[-1])
        a = [rowptr, i32(col, nnz), f64(W, 9 * nnz), f64(Wd, 9 * n), f64(A, 9 * n), f64(B, 3 * n),
             f64(V, 3 * n), f64(rho, n), f64(friction, n), f64(restitution, n), f64(cohesion, n), f64(gap, n)]
        Rr, Ur = f64(R, 3 * n), f64(U, 3 * n)
        rc = self.lib.SFGS_System_Upload(self.ctx, n, *[ptr(x) for x in a], float(free_energy), float(step),
                                         int(dynamic), ptr(Rr), ptr(Ur))
        self._check(rc, "SFGS_System_Upload")

    # ---- solve --------------------------------------------------------
    def solve(self, param, history=True):
        rep = SFGS_REPORT()
        rer = np.zeros(param.maxiter) if history else None
        mer = np.zeros(param.maxiter) if history else None
        self._check(self.lib.SFGS_Solve(self.ctx, C.byref(param), ptr(rer), ptr(mer), C.byref(rep)), "SFGS_Solve")
        if history:
            rer, mer = rer[:rep.iters].copy(), mer[:rep.iters].copy()
        return rep, rer, mer

    def bodyspace_product(self, x):
        """SFGS_Bodyspace_Product: z = W x through the bodies (locdyn 'OFF', nts.c:396-431),
        x and z natural, 3 per contact."""
        n = self.sizes()[0]
        x = np.ascontiguousarray(x, np.float64).reshape(3 * n)
        z = np.zeros(3 * n)
        self._check(self.lib.SFGS_Bodyspace_Product(self.ctx, ptr(x), ptr(z)), "SFGS_Bodyspace_Product")
        return z.reshape(n, 3)

    def merit(self, update_U=1, conmerit=False):
        m = C.c_double(0.0)
        cm = np.zeros(self.sizes()[0]) if conmerit else None
        self._check(self.lib.SFGS_Merit(self.ctx, update_U, C.byref(m), ptr(cm)), "SFGS_Merit")
        return (m.value, cm) if conmerit else m.value

    # ---- outputs ------------------------------------------------------
    def sizes(self):
        n, nnz, nc = C.c_int(0), C.c_longlong(0), C.c_int(0)
        self._check(self.lib.SFGS_Get_Sizes(self.ctx, C.byref(n), C.byref(nnz), C.byref(nc)), "SFGS_Get_Sizes")
        return n.value, nnz.value, nc.value

    def get_RU(self):
        n = self.sizes()[0]
        R, U = np.zeros(3 * n), np.zeros(3 * n)
        self._check(self.lib.SFGS_Get_RU(self.ctx, ptr(R), ptr(U)), "SFGS_Get_RU")
        return R.reshape(n, 3), U.reshape(n, 3)

    def update_cohesion(self):
        """SFGS_Update_Cohesion: update_cohesion of LOCDYN_Update_End (ldy.c:69-95) after a solve."""
        self._check(self.lib.SFGS_Update_Cohesion(self.ctx), "SFGS_Update_Cohesion")

    def cohesion(self):
        """The contacts' cohesion states, natural order."""
        c = np.zeros(self.sizes()[0])
        self._check(self.lib.SFGS_Get_Cohesion(self.ctx, ptr(c)), "SFGS_Get_Cohesion")
        return c

    def locdyn(self):
        n, nnz, _ = self.sizes()
        out = dict(rowptr=np.zeros(n + 1, np.int32), col=np.zeros(nnz, np.int32), via=np.zeros(nnz, np.int32),
                   W=np.zeros(9 * nnz), Wd=np.zeros(9 * n), A=np.zeros(9 * n), B=np.zeros(3 * n),
                   V=np.zeros(3 * n), rho=np.zeros(n), colour=np.zeros(n, np.int32))
        e = C.c_double(0.0)
        rc = self.lib.SFGS_Get_Locdyn(self.ctx, ptr(out["rowptr"]), ptr(out["col"]), ptr(out["via"]),
                                      ptr(out["W"]), ptr(out["Wd"]), ptr(out["A"]), ptr(out["B"]),
                                      ptr(out["V"]), ptr(out["rho"]), C.byref(e), ptr(out["colour"]))
        self._check(rc, "SFGS_Get_Locdyn")
        out["free_energy"] = e.value
        return out

    def colour(self, rowptr, col):
        n = len(rowptr) - 1
        rowptr = i32(rowptr, n + 1)
        col = i32(col, int(rowptr[-1]))
        out = np.zeros(max(n, 1), np.int32)
        nc = self._check(self.lib.SFGS_Colour(self.ctx, n, ptr(rowptr), ptr(col), ptr(out)), "SFGS_Colour")
        return nc, out[:n]

    # ---- learned warm start --------------------------------------------
    def features(self):
        n = self.sizes()[0]
        out = np.zeros(18 * max(n, 1), np.float32)
        self._check(self.lib.SFGS_Features(self.ctx, ptr(out)), "SFGS_Features")
        return out[:18 * n].reshape(n, 18)

    def warm_start(self):
        self._check(self.lib.SFGS_Warm_Start(self.ctx), "SFGS_Warm_Start")

    def warm_start_mode(self, new_only=0, scale=1.0):
        self._check(self.lib.SFGS_Warm_Start_Mode(self.ctx, int(new_only), float(scale)), "SFGS_Warm_Start_Mode")

    def mlp_set(self, params, norm):
        params = np.ascontiguousarray(params, np.float32)
        norm = np.ascontiguousarray(norm, np.float32)
        self._check(self.lib.SFGS_Mlp_Set(self.ctx, ptr(params), params.size, ptr(norm)), "SFGS_Mlp_Set")

    def mlp_get(self):
        nparam = C.c_longlong(0)
        self._check(self.lib.SFGS_Mlp_Get(self.ctx, None, C.byref(nparam)), "SFGS_Mlp_Get")
        out = np.zeros(nparam.value, np.float32)
        self._check(self.lib.SFGS_Mlp_Get(self.ctx, ptr(out), None), "SFGS_Mlp_Get")
        return out

    def mlp_forward(self, X):
        X = np.ascontiguousarray(X, np.float32)
        out = np.zeros((len(X), 3), np.float32)
        self._check(self.lib.SFGS_Mlp_Forward(self.ctx, ptr(X), len(X), ptr(out)), "SFGS_Mlp_Forward")
        return out

    def mlp_grad(self, X, Y):
        X = np.ascontiguousarray(X, np.float32)
        Y = np.ascontiguousarray(Y, np.float32)
        g = np.zeros(self.mlp_nparam(), np.float32)
        loss = C.c_float(0.0)
        self._check(self.lib.SFGS_Mlp_Grad(self.ctx, ptr(X), ptr(Y), len(X), ptr(g), C.byref(loss)), "SFGS_Mlp_Grad")
        return loss.value, g

    def mlp_nparam(self):
        nparam = C.c_longlong(0)
        self._check(self.lib.SFGS_Mlp_Get(self.ctx, None, C.byref(nparam)), "SFGS_Mlp_Get")
        return nparam.value

    def mlp_train(self, X, Y, epochs, batch, lr, seed=0):
        X = np.ascontiguousarray(X, np.float32)
        Y = np.ascontiguousarray(Y, np.float32)
        hist = np.zeros(epochs, np.float32)
        rc = self.lib.SFGS_Mlp_Train(self.ctx, ptr(X), ptr(Y), len(X), int(epochs), int(batch), float(lr), int(seed),
                                     ptr(hist))
        self._check(rc, "SFGS_Mlp_Train")
        return hist

    # ---- batched solve of independent scenes ----------------------------
    def batch_scenes(self, body_scene, nscenes=None):
        body_scene = np.ascontiguousarray(body_scene, np.int32)
        ns = int(body_scene.max()) + 1 if nscenes is None else int(nscenes)
        self._check(self.lib.SFGS_Batch_Scenes(self.ctx, ns, ptr(body_scene) if ns > 1 else None),
                    "SFGS_Batch_Scenes")
        return ns

    def batch_solve(self, param, nscenes, tune=None):
        """Solve every scene; returns (per-scene records as a structured array, SFGS_REPORT)."""
        reps = (SFGS_SCENE_REPORT * max(nscenes, 1))()
        tot = SFGS_REPORT()
        t = None
        if tune is not None:
            t = SFGS_BATCH_TUNE(*[int(v) for v in (tuple(tune) + (0,))[:4]])
        rc = self.lib.SFGS_Batch_Solve(self.ctx, C.byref(param), C.byref(t) if t is not None else None, reps,
                                       C.byref(tot))
        self._check(rc, "SFGS_Batch_Solve")
        out = np.zeros(nscenes, dtype=[("rows", "i4"), ("groups", "i4"), ("iters", "i4"), ("error_code", "i4"),
                                       ("fail_rows", "i4"), ("aborted", "i4"), ("error", "f8"), ("merit", "f8"),
                                       ("energy", "f8")])
        for k in range(nscenes):
            r = reps[k]
            out[k] = (r.rows, r.groups, r.iters, r.error_code, r.fail_rows, r.aborted, r.error, r.merit, r.energy)
        return out, tot

    def batch_autotune(self, param, trial_iters=20):
        best = SFGS_BATCH_TUNE()
        ms = C.c_double(0.0)
        self._check(self.lib.SFGS_Batch_Autotune(self.ctx, C.byref(param), int(trial_iters), C.byref(best),
                                                 C.byref(ms)), "SFGS_Batch_Autotune")
        return (best.lanes_per_row, best.waves_per_group, best.rows_per_group, best.schedule), ms.value

    def body_update(self, step, nb, download=True):
        if not download:
            self._check(self.lib.SFGS_Body_Update(self.ctx, step, None, None), "SFGS_Body_Update")
            return None, None
        velo, force = np.zeros(6 * nb), np.zeros(6 * nb)
        self._check(self.lib.SFGS_Body_Update(self.ctx, step, ptr(velo), ptr(force)), "SFGS_Body_Update")
        return velo.reshape(nb, 6), force.reshape(nb, 6)
