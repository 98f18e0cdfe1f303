"""NEWTON_SOLVER on the GPU (SFGS_Newton_Solve) against the oracle's restatement
of nts.c and ext/krylov/flexgmres.c on the same assembled systems: a fixed
number of Newton steps must give the same reactions, merit and product
histories (the step is FP64 throughout; the device sums dot products per block
and W x in slice order, the oracle sequentially, so iterates agree to rounding:
1e-8 relative to max|R| for GMRES steps, 1e-10 for DIAG steps); a converged
solve satisfies the contact law; the Gauss-Seidel fallback matches the oracle's;
RUN with a NEWTON_SOLVER gives the same bits in the host and the device loop
and carries a resting sphere's weight."""
import numpy as np
import pytest

import oracle
import solfec
from scenes import lattice
from test_oracle_scf import residual

pytestmark = pytest.mark.gpu

H = 1e-3


@pytest.fixture(scope="module")
def dev():
    d = solfec.Device(0)
    yield d
    d.close()


def upload_system(dev, O, R0):
    n = O["n"]
    dev.system(O["rowptr"], O["col"], O["W"], O["Wd"], O["A"], O["B"], O["V"], O["rho"], O["mu"], O["rest"],
               O["coh"], O["gap"], O["free_energy"], O["step"], R0, np.zeros(3 * n))


def assembled(n, seed):
    s = lattice(n, seed=seed)
    b, c = s.bodies(), s.contacts()
    O = oracle.assemble(oracle.bodies_from(b), oracle.contacts_from(c), H)
    return O, oracle.System(O), b, c


@pytest.mark.parametrize("linver,steps,rtol", [("GMRES", 4, 1e-8), ("DIAG", 12, 1e-10)])
@pytest.mark.parametrize("n,seed", [(216, 3), (1000, 1)])
def test_newton_steps_match_oracle(dev, n, seed, linver, steps, rtol):
    print(f"[newton] n={n} seed={seed} {linver} {steps} steps", flush=True)
    O, S, b, c = assembled(n, seed)
    R0 = np.random.default_rng(seed).normal(scale=1e-2, size=3 * O["n"])
    kw = dict(linver=linver, maxiter=steps, meritval=1e-30, gsflag="OFF")
    upload_system(dev, O, R0)
    rep, mer, mv = dev.newton(solfec.newton_param(**kw))
    Rg, Ug = dev.get_RU()
    orep, Ro, Uo, omer, omv = oracle.newton_solve(S, oracle.nt_param(**kw), R0)
    dR = np.abs(Rg - Ro).max() / np.abs(Ro).max()
    print(f"[newton] merit {rep.merit:.3e} / {orep.merit:.3e}, products {rep.matvecs} / {orep.matvecs},"
          f" max |dR|/max|R| {dR:.2e}", flush=True)
    assert rep.iters == orep.iters == steps and not rep.gs_fallback and rep.singular == 0
    np.testing.assert_array_equal(mv, omv)
    np.testing.assert_allclose(mer, omer, rtol=1e3 * rtol)
    assert dR <= rtol
    assert np.abs(Ug - Uo).max() <= rtol * 1e3 * np.abs(Uo).max()


@pytest.mark.parametrize("linver", ["GMRES", "DIAG"])
def test_newton_converges_to_the_contact_law(dev, linver):
    O, S, b, c = assembled(1000, 1)
    n = O["n"]
    meritval = 1e-10 if linver == "GMRES" else 1e-8
    upload_system(dev, O, np.zeros(3 * n))
    rep, mer, mv = dev.newton(solfec.newton_param(linver=linver, maxiter=2000, meritval=meritval, gsflag="OFF"))
    R, U = dev.get_RU()
    print(f"[newton] {linver}: {rep.iters} steps, {rep.matvecs} products, merit {rep.merit:.3e}", flush=True)
    assert rep.converged and rep.merit <= meritval and mer[-1] == rep.merit
    Cv = np.array([residual(U[i], R[i], O["mu"][i]) for i in range(n)])
    assert np.abs(Cv).max() <= 1e-4 * max(np.abs(R).max(), np.abs(U).max())
    m, Um, _ = oracle.merit(S, R.ravel(), U.ravel(), 1)
    np.testing.assert_allclose(Um, U, rtol=0, atol=1e-10 * np.abs(U).max())


def test_newton_fallback_matches_oracle(dev):
    """Three Newton steps above meritval, then R reset and the Gauss-Seidel
    fallback (three sweeps here: maxiter of the Newton solver).  Three sweeps
    from a random R0 amplify rounding: on the CPU a 1e-12 relative change of R0
    moves the oracle's own R by 1.5e-6 relative, so R is compared at 1e-3 and
    the merit, a smooth function of R, at 1e-6."""
    O, S, b, c = assembled(512, 6)
    n = O["n"]
    R0 = np.random.default_rng(2).normal(scale=1e-2, size=3 * n)
    kw = dict(maxiter=3, meritval=1e-12)
    upload_system(dev, O, R0)
    rep, mer, mv = dev.newton(solfec.newton_param(**kw))
    Rg, Ug = dev.get_RU()
    colours = dev.locdyn()["colour"]
    order, _ = oracle.colour_order(colours, int(colours.max()) + 1)
    orep, Ro, Uo, omer, omv = oracle.newton_solve(S, oracle.nt_param(**kw), R0, order=order)
    print(f"[newton] fallback: {rep.gs_iters} / {orep.gs_iters} sweeps, merit {rep.merit:.3e} / {orep.merit:.3e}",
          flush=True)
    assert rep.gs_fallback == orep.gs_fallback == 1 and rep.gs_iters == orep.gs_iters == 3
    assert np.abs(Rg - Ro).max() <= 1e-3 * np.abs(Ro).max()
    assert rep.merit == pytest.approx(orep.merit, rel=1e-6)


def newton_scene(seed):
    s = solfec.SOLFEC("DYNAMIC", H, "/tmp/solfec_test")
    solfec.SURFACE_MATERIAL(s, friction=0.3)
    solfec.GRAVITY(s, (0.0, 0.0, -9.81))
    solfec.LATTICE(s, 343, 0.7, 0.045, 0.055, 0.005, 2000.0, seed)
    return s


def test_newton_run_device_loop_equals_host_loop():
    ns = solfec.NEWTON_SOLVER(meritval=1e-8, maxiter=200)
    ns.itershist = "ON"
    host, dvc = newton_scene(5), newton_scene(5)
    dvc.device_loop(True)
    solfec.RUN(host, ns, 0.01)
    solfec.RUN(dvc, ns, 0.01)
    bh, bd = host.bodies(), dvc.bodies()
    for k in ("x", "rot", "velo"):
        assert np.array_equal(bh[k], bd[k]), k
    ch, cd = host.contacts(), dvc.contacts()
    assert len(ch["master"]) == len(cd["master"]) > 100
    for k in ("R", "U"):
        assert np.array_equal(ch[k], cd[k]), k
    hist = ns.itershist
    print(f"[newton] RUN: iterations per step {hist}", flush=True)
    assert len(hist) == 20 and hist[:10] == hist[10:]


@pytest.mark.parametrize("linver", ["GMRES", "DIAG"])
def test_sphere_resting_on_a_plane(linver):
    s = solfec.SOLFEC("DYNAMIC", H, "/tmp/solfec_test")
    solfec.SURFACE_MATERIAL(s, friction=0.5)
    solfec.GRAVITY(s, (0.0, 0.0, -9.81))
    mat = solfec.BULK_MATERIAL(s, density=2000.0)
    r = 0.05
    solfec.BODY(s, "OBSTACLE", solfec.PLANE((0, 0, 0), (0, 0, 1), 0), mat)
    solfec.BODY(s, "RIGID", solfec.SPHERE((0.3, 0.2, r), r, 1, 1), mat)
    ns = solfec.NEWTON_SOLVER(meritval=1e-14, maxiter=100, linver=linver, theta=1.0)
    solfec.RUN(s, ns, 0.02)
    b, c = s.bodies(), s.contacts()
    m = 2000.0 * 4.0 / 3.0 * np.pi * r ** 3
    np.testing.assert_allclose(b["velo"][6:12], 0.0, atol=1e-9)
    np.testing.assert_allclose(c["R"], [0.0, 0.0, m * 9.81], rtol=1e-6, atol=1e-9)
    assert ns.report().gs_fallback == 0


@pytest.mark.parametrize("solver", ["newton", "penalty"])
@pytest.mark.parametrize("loop", ["host", "device"])
def test_other_solvers_run_empty_domains(solver, loop):
    """Planes only, and then nothing at all: the NEWTON and PENALTY solvers run
    steps without contacts."""
    for planes in (1, 0):
        s = solfec.SOLFEC("DYNAMIC", H, "/tmp/solfec_test")
        solfec.SURFACE_MATERIAL(s, model="SPRING_DASHPOT" if solver == "penalty" else "SIGNORINI_COULOMB",
                                spring=1e5)
        mat = solfec.BULK_MATERIAL(s, density=1000.0)
        if planes:
            solfec.BODY(s, "OBSTACLE", solfec.PLANE((0, 0, 0), (0, 0, 1), 0), mat)
        s.device_loop(loop == "device")
        sv = solfec.NEWTON_SOLVER() if solver == "newton" else solfec.PENALTY_SOLVER()
        solfec.RUN(s, sv, 0.003)
        assert s.ncon == 0 and s.time == pytest.approx(0.003)
        if solver == "newton":
            assert sv.iters == 0 and sv.report().converged == 1


@pytest.mark.parametrize("linver", ["GMRES", "DIAG"])
def test_newton_body_space_equals_assembled(dev, linver):
    """locdyn OFF forms W x and U = B + W R through the bodies, H (h M^-1 H' x)
    (nts.c:396-431); the same Newton steps on the assembled blocks give the
    same iterates up to the rounding of the products.  Three steps: inexact
    GMRES steps let rounding differences grow over a long solve (at 60 steps
    the two merits differ by orders of magnitude while both decrease), so the
    operator is compared where the paths still coincide."""
    print(f"[newton] body space against assembled, {linver}", flush=True)
    s = lattice(343, seed=5)
    b, c = s.bodies(), s.contacts()

    def run(bodyspace):
        dev.bodies(b["kind"], b["rot"], b["x0"], b["jinv"], b["minv"], b["velo"], b["velo0"])
        dev.contacts(c["master"], c["slave"], c["mpnt"], c["spnt"], c["base"], c["gap"], c["friction"],
                     c["restitution"], c["cohesion"], c["R"], c["U"])
        dev.assemble(H)
        p = solfec.newton_param(meritval=1e-14, maxiter=3, linver=linver, gsflag="OFF")
        p.bodyspace = bodyspace
        rep, mer, mv = dev.newton(p)
        R, U = dev.get_RU()
        return rep, np.asarray(mer), R, U
    ra, ma, Ra, Ua = run(0)
    rb, mb, Rb, Ub = run(1)
    k = min(ra.iters, rb.iters)
    err = np.abs(Rb - Ra).max() / np.abs(Ra).max()
    print(f"[newton] iters {ra.iters} / {rb.iters}, merit {ra.merit:.3e} / {rb.merit:.3e}, max |dR|/max|R| {err:.2e}",
          flush=True)
    assert ra.iters == rb.iters == 3
    assert np.allclose(ma, mb, rtol=1e-8, atol=0.0)
    assert err < 1e-8
    assert np.abs(Ub - Ua).max() <= 1e-8 * np.abs(Ua).max()


def test_newton_body_space_run_device_loop_equals_host_loop():
    ns = solfec.NEWTON_SOLVER(meritval=1e-8, maxiter=200, locdyn="OFF", gsflag="OFF")
    ref = solfec.NEWTON_SOLVER(meritval=1e-8, maxiter=200, gsflag="OFF")
    host, dvc, asm = newton_scene(5), newton_scene(5), newton_scene(5)
    dvc.device_loop(True)
    solfec.RUN(host, ns, 0.01)
    solfec.RUN(dvc, ns, 0.01)
    solfec.RUN(asm, ref, 0.01)
    bh, bd, ba = host.bodies(), dvc.bodies(), asm.bodies()
    for k in ("x", "rot", "velo"):
        assert np.array_equal(bh[k], bd[k]), k
    ch, cd = host.contacts(), dvc.contacts()
    for k in ("R", "U"):
        assert np.array_equal(ch[k], cd[k]), k
    d = np.abs(bh["velo"] - ba["velo"]).max()
    print(f"[newton] body space against assembled after 10 steps: max |d velo| {d:.2e}"
          f" of {np.abs(ba['velo']).max():.2e}", flush=True)
    assert d <= 1e-3 * np.abs(ba["velo"]).max()


def _oracle_W_times(O, x):
    """W x from the oracle's assembled blocks: W_ii and the off-diagonal blocks
    (column-major 3x3, natural CSR)."""
    n = O["n"]
    Wd = O["Wd"].reshape(n, 3, 3).transpose(0, 2, 1)          # [i, row, col]
    z = np.einsum("irc,ic->ir", Wd, x)
    rows = np.repeat(np.arange(n), np.diff(O["rowptr"]))
    Wb = O["W"].reshape(-1, 3, 3).transpose(0, 2, 1)
    np.add.at(z, rows, np.einsum("krc,kc->kr", Wb, x[O["col"]]))
    return z


def _ellipsoid_scene(n, seed):
    s = solfec.SOLFEC("DYNAMIC", H, "/tmp/solfec_test")
    solfec.SURFACE_MATERIAL(s, friction=0.3)
    solfec.GRAVITY(s, (0.0, 0.0, -9.81))
    solfec.ELLIP_LATTICE(s, n, 0.6, 0.03, 0.06, 0.005, 2000.0, seed)
    s.step_begin()
    return s


@pytest.mark.parametrize("scene", ["spheres", "cohesive spheres", "ellipsoids"])
def test_bodyspace_product_matches_oracle_W(dev, scene):
    """The operator of locdyn 'OFF' (nts.c:396-431), h H M^-1 H' x formed through
    the bodies (k_bs_bodies, k_bs_rows), against the oracle's assembled W
    (ldy.c:654-844, pinned by impulse responses and closed forms) times the same
    x: within 1e-12 of max|z|, obstacles included."""
    print(f"[newton] body-space product against the oracle's W, {scene}", flush=True)
    s = _ellipsoid_scene(343, 4) if scene == "ellipsoids" else lattice(343, seed=5)
    b, c = s.bodies(), s.contacts()
    if scene == "cohesive spheres":
        c = dict(c)
        c["cohesion"] = np.where(np.arange(len(c["master"])) % 2 == 0, 0.05, 0.0)
    n = len(c["master"])
    dev.bodies(b["kind"], b["rot"], b["x0"], b["jinv"], b["minv"], b["velo"], b["velo0"])
    dev.contacts(c["master"], c["slave"], c["mpnt"], c["spnt"], c["base"], c["gap"], c["friction"],
                 c["restitution"], c["cohesion"], c["R"], c["U"])
    dev.assemble(H)
    O = oracle.assemble(oracle.bodies_from(b), oracle.contacts_from(c), H)
    rng = np.random.default_rng(11)
    for k in range(3):
        x = rng.normal(size=(n, 3)) * (10.0 ** rng.uniform(-2, 2, size=(n, 1)))
        z = dev.bodyspace_product(x)
        zo = _oracle_W_times(O, x)
        err = np.abs(z - zo).max() / np.abs(zo).max()
        print(f"[newton] {n} contacts, {O['nnz']} blocks: max |z - W x| / max |W x| {err:.2e}", flush=True)
        assert err < 1e-12


@pytest.mark.parametrize("linver", ["GMRES", "DIAG"])
def test_upmin_assembly_and_body_space_newton(dev, linver):
    """SFGS_Locdyn_Update_Min (UPMIN, ldy.c:59, 644), the update the reference
    runs for locdyn OFF: the diagonal blocks, A, rho, B, V and the free energy
    equal the full assembly's bit for bit (the same per-row kernel), no
    off-diagonal blocks or adjacency are formed, the solves that need them
    refuse, and three body-space Newton steps on it give what they give on the
    full assembly: the same reactions bit for bit with DIAG steps (per-row
    updates), to rounding with GMRES (the dot products run in another row
    order)."""
    print(f"[newton] UPMIN assembly, body-space {linver}", flush=True)
    s = lattice(343, seed=5)
    b, c = s.bodies(), s.contacts()

    def run(diagonal_only):
        dev.bodies(b["kind"], b["rot"], b["x0"], b["jinv"], b["minv"], b["velo"], b["velo0"])
        dev.contacts(c["master"], c["slave"], c["mpnt"], c["spnt"], c["base"], c["gap"], c["friction"],
                     c["restitution"], c["cohesion"], c["R"], c["U"])
        dev.assemble(H, diagonal_only=diagonal_only)
        L = dev.locdyn()
        sizes = dev.sizes()
        if diagonal_only:
            with pytest.raises(solfec.device.SfgsError):
                dev.solve(solfec.make_param(1e-6, 10))
            with pytest.raises(solfec.device.SfgsError):
                dev.newton(solfec.newton_param(meritval=1e-14, maxiter=3, linver=linver, gsflag="OFF"))
        p = solfec.newton_param(meritval=1e-14, maxiter=3, linver=linver, gsflag="OFF")
        p.bodyspace = 1
        rep, mer, _ = dev.newton(p)
        R, U = dev.get_RU()
        return L, sizes, rep, np.asarray(mer), R
    Lf, sf, rf, mf, Rf = run(False)
    Lm, sm, rm, mm, Rm = run(True)
    print(f"[newton] full: {sf[1]} blocks, {sf[2]} colours; UPMIN: {sm[1]} blocks, {sm[2]} colours; merits "
          f"{mf[-1]:.6e} / {mm[-1]:.6e}, max |dR| {np.abs(Rf - Rm).max():.2e}", flush=True)
    assert sm[1] == 0 and sm[2] == 1 and sf[1] > 0
    for k in ("Wd", "A", "B", "V", "rho"):
        assert np.array_equal(Lf[k], Lm[k]), k
    assert Lf["free_energy"] == Lm["free_energy"]
    assert rf.iters == rm.iters == 3
    if linver == "DIAG":
        assert np.array_equal(Rf, Rm)
    else:
        assert np.abs(Rf - Rm).max() <= 1e-10 * np.abs(Rf).max()
        np.testing.assert_allclose(mm, mf, rtol=1e-8, atol=0.0)
