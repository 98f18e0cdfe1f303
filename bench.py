"""Benchmark of the contact-solve hot path on one MI355X.

Workload (the configuration the headline metric is quoted on): 200,000 rigid
spheres, radii U[0.045, 0.055] m, density 2000 kg/m^3, on a seeded jittered
cubic lattice (+-0.005 m) in a 6x6x6 m box of 6 fixed planes, gravity 9.81,
Coulomb mu = 0.5, restitution 0, h = 1e-3 s, settled for 500 steps through the
product RUN loop (GAUSS_SEIDEL_SOLVER(1e-4, 1000) while settling).  The
settled state of the next step (bodies, detected contacts, previous-step R and
U as warm start) is then uploaded once and stays resident in HBM.

One timed step = LOCDYN assembly on the device + GAUSS_SEIDEL_SOLVER with
epsilon = 1e-6, meritval = 1e-6, maxiter = 5000 + body velocity update, all
on the GPU; each step starts from the same resident state.  Every library
call returns after synchronising its stream, so the host clock brackets
completed device work.

Printed line: value = contact-block updates/s (row updates of the GS sweeps per
second of whole-step time), ms_per_step, per mode the merit reached, the first
iteration with merit <= 1e-6 (none, on this packing: the reference GS stalls
at merit ~1.5e-5) and the time to the merit the plain FP64 solve ends with;
the roofline of the sweep kernel (algorithmic bytes per row s(9k+26)+4k+4,
k = off-diagonal blocks, of every sweep of the timed steps over those solves'
device time, HIP events on the library's stream); and the CPU oracle, built
-march=native on this host, timed on a bounded sample of the same system.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for p in (os.path.join(ROOT, "solfec-1.0_amd"), ROOT):
    if p not in sys.path:
        sys.path.insert(0, p)

import numpy as np  # noqa: E402

METRIC = "ms/step to merit<=1e-6 & contact-block updates/s: learned+mixed vs plain GPU GS"
HBM_PEAK_GBS = 8000.0


def log(msg):
    print(msg, file=sys.stderr, flush=True)


PHASES = ("TIMINT", "CONUPD", "CONDET", "LOCDYN", "CONSOL")


def scene(n, L, seed, friction, shape="sphere"):
    import solfec
    s = solfec.SOLFEC("DYNAMIC", 1e-3, "/tmp/solfec_bench")
    solfec.SURFACE_MATERIAL(s, friction=friction, restitution=0.0)
    solfec.GRAVITY(s, (0.0, 0.0, -9.81))
    if shape == "ellipsoid":   # semi-axes U[0.03, 0.06], random orientations
        solfec.ELLIP_LATTICE(s, n, L, 0.03, 0.06, 0.005, 2000.0, seed)
    else:
        solfec.LATTICE(s, n, L, 0.045, 0.055, 0.005, 2000.0, seed)
    return s


def settle(n, L, steps, seed, friction, shape="sphere", device_loop=True):
    """The settling run through the product's RUN loop: by default the device loop
    (bodies and contacts stay on the GPU; the same bits as the host loop)."""
    import solfec
    s = scene(n, L, seed, friction, shape)
    s.device_loop(device_loop)
    gs = solfec.GAUSS_SEIDEL_SOLVER(1e-4, 1000)
    t0 = time.perf_counter()
    for k in range(steps):
        rc = s.step_begin() or s.step_solve(gs) or s.step_end()
        if rc:
            raise RuntimeError(f"settling step {k} failed: {s.lib.SOLFEC_Last_Error(s.sol).decode()}")
        if (k + 1) % 25 == 0 or k == 0:
            log(f"[settle] step {k + 1}/{steps} contacts={s.ncon} gs_iters={gs.iters} "
                f"elapsed={time.perf_counter() - t0:.1f}s")
    log("[settle] phases (s): " + ", ".join(f"{p} {s.timing(p):.2f}" for p in PHASES))
    s.step_begin()   # contacts of the next step; R, U carried from the last solve
    s.device_loop(False)
    return s


def bench_run(args):
    """The whole time step (RUN loop) with bodies and contacts on the device
    against the host loop: config 1 (1000 spheres, 200 steps, GS 1e-4 / 1000)
    and the first steps of the config-2 settle.  Both loops compute the same bits;
    the final states are compared."""
    import solfec
    out = {}
    cases = [("config1", 1000, 1.0, 200, 0.3), ("config2_settle", args.spheres, args.box, args.run_steps, args.friction)]
    want = [w.strip() for w in args.run_configs.split(",")]
    for name, n, L, steps, fr in [c for c in cases if c[0] in want]:
        res = {}
        states = []
        for loop in ("host", "device"):
            s = scene(n, L, args.seed if name != "config1" else 1, fr)
            s.device_loop(loop == "device")
            gs = solfec.GAUSS_SEIDEL_SOLVER(1e-4, 1000)
            t0 = time.perf_counter()
            for k in range(steps):
                rc = s.step_begin() or s.step_solve(gs) or s.step_end()
                if rc:
                    raise RuntimeError(f"{name} {loop} step {k}: {s.lib.SOLFEC_Last_Error(s.sol).decode()}")
            dt = time.perf_counter() - t0
            ph = {p: round(1e3 * s.timing(p) / steps, 3) for p in PHASES}
            res[loop] = dict(ms_per_step=1e3 * dt / steps, contacts=s.ncon, phases_ms_per_step=ph)
            log(f"[run] {name} {loop}: {1e3 * dt / steps:.2f} ms/step, {s.ncon} contacts, phases {ph}")
            b = s.bodies()
            states.append((b["x"], b["velo"], s.contacts()["R"]))
        same = all(np.array_equal(a, b) for a, b in zip(*states))
        res["identical_state"] = bool(same)
        res["speedup"] = res["host"]["ms_per_step"] / res["device"]["ms_per_step"]
        out[name] = res
    head = out[want[-1]]["device"]["ms_per_step"]
    line = {"metric": "RUN loop ms/step: device-resident bodies and contacts vs host loop", "value":
            head, "unit": "ms/step", "n_gpus": 1, "steps": args.run_steps,
            "warmup": 0, "ms_per_step": head, "higher_is_better": False,
            "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
            "config": {"workload": f"config-1 run (1000 spheres, 200 steps) and the first {args.run_steps} steps of "
                                   f"the {args.spheres}-sphere settle, GS 1e-4 / 1000", **out}}
    print(json.dumps(line), flush=True)


def settled_state(args):
    """Bodies and contacts of the settled packing: settled here, or read from
    --snapshot; written to --save-snapshot when asked."""
    if args.snapshot:
        z = np.load(args.snapshot)
        b = {k[2:]: z[k] for k in z.files if k.startswith("b_")}
        c = {k[2:]: z[k] for k in z.files if k.startswith("c_")}
        return b, c
    s = settle(args.spheres, args.box, args.settle, args.seed, args.friction, args.shape,
               device_loop=not args.host_loop)
    b, c = s.bodies(), s.contacts()
    if args.save_snapshot:
        np.savez(args.save_snapshot, **{"b_" + k: v for k, v in b.items()}, **{"c_" + k: v for k, v in c.items()})
    return b, c


def bench_consumers(args):
    """The other consumers of the local dynamics on the settled packing of the
    Gauss-Seidel workload: NEWTON_SOLVER (GMRES and DIAG steps) to merit 1e-6
    against the FP64 Gauss-Seidel solve of the same system, and the
    PENALTY_SOLVER's per-contact law (one fused kernel per call)."""
    import solfec
    h = 1e-3
    b, c = settled_state(args)
    dev = solfec.Device(0)
    dev.bodies(b["kind"], b["rot"], b["x0"], b["jinv"], b["minv"], b["velo"], b["velo0"])

    def fresh(assemble=True):
        dev.contacts(c["master"], c["slave"], c["mpnt"], c["spnt"], c["base"], c["gap"], c["friction"],
                     c["restitution"], c["cohesion"], c["R"], c["U"])
        if assemble:
            dev.assemble(h)

    out = {}
    fresh()
    n, nnz, ncol = dev.sizes()
    t0 = time.perf_counter()
    rep, _, _ = dev.solve(solfec.make_param(1e-6, args.maxiter, meritval=1e-6), history=False)
    dt = time.perf_counter() - t0
    out["gauss_seidel_fp64"] = dict(ms=1e3 * dt, iters=rep.iters, merit=rep.merit, converged=bool(rep.merit <= 1e-6))
    log(f"[consumers] GS FP64: {1e3 * dt:.1f} ms, {rep.iters} iterations, merit {rep.merit:.3e}")
    for lv, maxiter in (("GMRES", args.nt_maxiter), ("DIAG", 4 * args.nt_maxiter)):
        fresh()
        t0 = time.perf_counter()
        rep, mer, mv = dev.newton(solfec.newton_param(meritval=1e-6, maxiter=maxiter, linver=lv, gsflag="OFF"))
        dt = time.perf_counter() - t0
        out["newton_" + lv.lower()] = dict(ms=1e3 * dt, steps=rep.iters, products=rep.matvecs, merit=rep.merit,
                                           converged=bool(rep.converged),
                                           merit_every_10=[float(x) for x in mer[9::10][:30]])
        log(f"[consumers] NEWTON {lv}: {1e3 * dt:.1f} ms, {rep.iters} steps, {rep.matvecs} products, "
            f"merit {rep.merit:.3e}")
    # body space (locdyn OFF), DIAG steps, the assembly included: the full update
    # against UPMIN (ldy.c:59, 644), which forms no off-diagonal blocks
    for tag, diag_only in (("full", False), ("upmin", True)):
        fresh(assemble=False)
        t0 = time.perf_counter()
        dev.assemble(h, diagonal_only=diag_only)
        ta = time.perf_counter()
        p = solfec.newton_param(meritval=1e-6, maxiter=4 * args.nt_maxiter, linver="DIAG", gsflag="OFF")
        p.bodyspace = 1
        rep, _, _ = dev.newton(p)
        t1 = time.perf_counter()
        out["newton_diag_bodyspace_" + tag] = dict(assembly_ms=1e3 * (ta - t0), solve_ms=1e3 * (t1 - ta),
                                                   ms=1e3 * (t1 - t0), steps=rep.iters, merit=rep.merit)
        log(f"[consumers] NEWTON DIAG body space, {tag} assembly: {1e3 * (ta - t0):.1f} + {1e3 * (t1 - ta):.1f} ms, "
            f"{rep.iters} steps, merit {rep.merit:.3e}")
    # penalty: a linear spring with critical damping on every contact
    fresh(assemble=False)
    k = np.full(n, 1e6)
    dev.penalty(h, 1, k, np.full(n, -1.0), np.ones(n))
    reps = 20
    t0 = time.perf_counter()
    for _ in range(reps):
        dev.penalty(h, 1, k, np.full(n, -1.0), np.ones(n))
    dt = (time.perf_counter() - t0) / reps
    R, _ = dev.get_RU()
    out["penalty_implicit"] = dict(ms_per_call=1e3 * dt, contacts_per_s=n / dt, active=int((R[:, 2] > 0).sum()),
                                   bytes_per_contact=248, bytes_per_body=276)
    log(f"[consumers] PENALTY: {1e3 * dt:.3f} ms per call ({n} contacts, {(R[:, 2] > 0).sum()} compressed)")
    head = min(out["newton_gmres"]["ms"], out["newton_diag"]["ms"])
    line = {"metric": "NEWTON_SOLVER ms to merit<=1e-6 and PENALTY_SOLVER calls on the settled packing",
            "value": head, "unit": "ms", "n_gpus": 1, "steps": 1, "warmup": 0, "ms_per_step": head,
            "higher_is_better": False, "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
            "config": {"workload": f"{args.spheres}-{args.shape} packing settled {args.settle} steps: NEWTON_SOLVER "
                                   "(GMRES, DIAG; gsflag OFF) against GAUSS_SEIDEL FP64 to merit 1e-6, and "
                                   "PENALTY_SOLVER IMPLICIT (k 1e6, critical damping)",
                       "contacts": n, "offdiag_blocks": nnz, "colours": ncol, **out}}
    print(json.dumps(line), flush=True)
    dev.close()


def host_cpu():
    """Model name and logical CPU count of this host (lscpu, else /proc/cpuinfo)."""
    import subprocess
    model, count = None, os.cpu_count()
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=20).stdout
        for line in out.splitlines():
            k, _, v = line.partition(":")
            if k.strip() == "Model name":
                model = v.strip()
            elif k.strip() == "CPU(s)" and v.strip().isdigit():
                count = int(v.strip())
    except Exception:
        pass
    if model is None and os.path.exists("/proc/cpuinfo"):
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    return model, count


def cpu_baseline(b, c, colours, n, seconds, maxiter):
    """The oracle (C restatement, 1 core) on the same system: GS iterations of
    sweep + merit, as the GPU solve does, for about `seconds` of CPU time.  The
    local solver runs with the GPU solve's diagepsilon and diagmaxiter (those of
    epsilon = meritval = 1e-6); only the outer test is switched off, so that
    exactly the planned number of iterations runs.  The oracle is compiled on
    this host with -O3 -march=native first (the shipped build is kept if that
    fails, and the line says which ran)."""
    import oracle
    flags = oracle.use_native_build() or "-O3 -march=x86-64-v3 (prebuilt; native build failed)"
    model, ncpu = host_cpu()
    O = oracle.assemble(oracle.bodies_from(b), oracle.contacts_from(c), 1e-3)
    sysobj = oracle.System(O)
    order, _ = oracle.colour_order(colours, int(colours.max()) + 1)
    R0, U0 = c["R"].copy(), c["U"].copy()
    de, dm = oracle.default_diag(1e-6, maxiter, 1e-6)
    p = oracle.gs_param(1e-300, 1, meritval=1e-300, diagepsilon=de, diagmaxiter=dm)
    t0 = time.perf_counter()
    oracle.gs_solve(sysobj, p, order, R0, U0, history=False)
    t1 = time.perf_counter() - t0
    k = max(1, int(seconds / max(t1, 1e-6)))
    p = oracle.gs_param(1e-300, k, meritval=1e-300, diagepsilon=de, diagmaxiter=dm)
    t0 = time.perf_counter()
    rc, rep, _, _, _, _ = oracle.gs_solve(sysobj, p, order, R0, U0, history=False)
    dt = time.perf_counter() - t0
    if rep.iters != k:
        log(f"[bench] cpu baseline: the oracle stopped after {rep.iters} of {k} iterations "
            f"(relative error {rep.rel_error}, merit {rep.merit})")
    # the reference's own row order (LOCDYN list order, newest contact first), half the sample
    kl = max(1, k // 2)
    p = oracle.gs_param(1e-300, kl, meritval=1e-300, diagepsilon=de, diagmaxiter=dm)
    t0 = time.perf_counter()
    _, repl, _, _, _, _ = oracle.gs_solve(sysobj, p, oracle.list_order(n), R0, U0, history=False)
    dtl = time.perf_counter() - t0
    return dict(value=rep.iters * n / dt, unit="contact-block updates/s", cores=1, kind="port",
                sample=f"{rep.iters} GS iterations (colour-major sweep + merit) of the same {n}-contact "
                       f"system, the oracle on 1 host core, {dt:.1f} s",
                host_cpu=model, host_logical_cpus=ncpu, compiler_flags=flags,
                list_order=dict(value=repl.iters * n / dtl, iterations=int(repl.iters), seconds=round(dtl, 2)))


def stream_copy_rate(mib=1024, reps=10):
    """On-device copy rate of this GPU, for comparison with the nominal HBM peak:
    hipMemcpyAsync device to device of `mib` MiB, `reps` times, timed with HIP
    events; bytes counted = read + write."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so")
    hip.hipMemcpyAsync.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int, C.c_void_p]
    hip.hipEventRecord.argtypes = [C.c_void_p, C.c_void_p]
    hip.hipEventSynchronize.argtypes = [C.c_void_p]
    hip.hipEventElapsedTime.argtypes = [C.POINTER(C.c_float), C.c_void_p, C.c_void_p]
    hip.hipEventDestroy.argtypes = [C.c_void_p]
    hip.hipFree.argtypes = [C.c_void_p]
    nbytes = mib << 20
    src, dst, e0, e1 = C.c_void_p(), C.c_void_p(), C.c_void_p(), C.c_void_p()
    if hip.hipMalloc(C.byref(src), C.c_size_t(nbytes)) or hip.hipMalloc(C.byref(dst), C.c_size_t(nbytes)):
        raise RuntimeError("stream copy: hipMalloc failed")
    assert hip.hipEventCreate(C.byref(e0)) == 0 and hip.hipEventCreate(C.byref(e1)) == 0
    assert hip.hipMemcpyAsync(dst, src, nbytes, 3, None) == 0       # hipMemcpyDeviceToDevice, warm-up
    hip.hipEventRecord(e0, None)
    for _ in range(reps):
        hip.hipMemcpyAsync(dst, src, nbytes, 3, None)
    hip.hipEventRecord(e1, None)
    hip.hipEventSynchronize(e1)
    ms = C.c_float()
    hip.hipEventElapsedTime(C.byref(ms), e0, e1)
    for e in (e0, e1):
        hip.hipEventDestroy(e)
    hip.hipFree(src)
    hip.hipFree(dst)
    return 2.0 * nbytes * reps / (ms.value * 1e-3) / 1e9


def measured_traffic(kernel, args):
    """HBM bytes per colour of the headline sweep (a sweep's bytes over its
    colours) from the committed rocprofv3 --pmc summary of this workload and shape
    (FETCH_SIZE and WRITE_SIZE of every launch, scaled by a device copy of known
    size), or None when none matches."""
    import re
    path = os.path.join(ROOT, "profiles", "k_sweep_config2_traffic.md")
    if args.schedule == 1 or args.shape != "sphere" or not os.path.exists(path):
        return None
    key = f"{kernel} lanes={args.lanes} order={args.row_order}"
    for line in open(path):
        m = re.match(r"\|\s*`([^`]+)`\s*\|\s*([0-9.e+]+)\s*\|", line)
        if m and m.group(1) == key:
            return float(m.group(2))
    return None


def copy_calibration(mib):
    """Device-to-device copies of a known size, for scaling HBM traffic counters."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so")
    nbytes = mib << 20
    src, dst = C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(src), C.c_size_t(nbytes)) == 0 and hip.hipMalloc(C.byref(dst), C.c_size_t(nbytes)) == 0
    hip.hipMemset(src, 1, C.c_size_t(nbytes))
    for _ in range(3):
        hip.hipMemcpy(dst, src, C.c_size_t(nbytes), 3)   # hipMemcpyDeviceToDevice
    hip.hipDeviceSynchronize()
    hip.hipFree(src)
    hip.hipFree(dst)
    print(json.dumps({"copy_calibration_bytes": nbytes, "copies": 3}), flush=True)


def batch_scenes(nscenes, seed, lo=500, hi=50000):
    """Independent scenes with contact counts log-uniform in [lo, hi] and mu uniform in
    [0.1, 0.8]: each the config-1 generator (jittered lattice of spacing 0.1 m, radii
    U[0.045, 0.055], density 2000, 6 planes) sized for its count (about 1.52 contacts
    per sphere at the first step), at its first step."""
    import solfec
    rng = np.random.default_rng(seed)
    want = np.exp(rng.uniform(np.log(lo), np.log(hi), nscenes))
    mus = rng.uniform(0.1, 0.8, nscenes)
    out = []
    for k in range(nscenes):
        n = max(8, int(round(want[k] / 1.52)))
        side = int(np.ceil(round(n ** (1.0 / 3.0), 9)))
        s = solfec.SOLFEC("DYNAMIC", 1e-3, "/tmp/solfec_bench")
        solfec.SURFACE_MATERIAL(s, friction=float(mus[k]), restitution=0.0)
        solfec.GRAVITY(s, (0.0, 0.0, -9.81))
        solfec.LATTICE(s, n, 0.1 * side, 0.045, 0.055, 0.005, 2000.0, int(seed * 1000 + k))
        s.step_begin()
        out.append((s.bodies(), s.contacts()))
    return out


def bench_batch(args):
    """Config 5: many independent scenes in one persistent launch (learned warm
    start + mixed precision), against the same batch in FP64 without the network."""
    import solfec
    from solfec import batch, learned
    h = 1e-3
    t0 = time.perf_counter()
    scenes = batch_scenes(args.scenes, args.seed)
    B, Cn, bs, cs = batch.concat(scenes)
    ncon = len(cs)
    log(f"[batch] {args.scenes} scenes, {len(bs)} bodies, {ncon} contacts, built in {time.perf_counter() - t0:.1f}s")
    dev = solfec.Device(0)
    dev.set_layout(args.lanes)
    ns = batch.upload(dev, B, Cn, bs, h)
    have_model = os.path.exists(learned.DEFAULT_MODEL)
    if have_model:
        learned.install(dev)
        dev.warm_start_mode(args.ws_new_only, args.ws_scale)
    eps, maxiter = 1e-6, args.maxiter
    pm = solfec.make_param(eps, maxiter, meritval=1e-6, precision=solfec.device.MIXED, refine=2)
    pf = solfec.make_param(eps, maxiter, meritval=1e-6, precision=solfec.device.FP64)
    tune, tms = dev.batch_autotune(pm, trial_iters=20)
    log(f"[batch] autotuned shape {tune} ({tms:.2f} ms for 20 iterations)")

    def run(par, learned_ws, steps, warmup):
        batch.upload(dev, B, Cn, bs, h)   # every run starts from the scenes' own R0
        def step():
            dev.assemble(h)   # resident inputs; R restarts from the uploaded R0 (only get_RU writes it back)
            if learned_ws:
                dev.warm_start()
            return dev.batch_solve(par, ns)
        for _ in range(warmup):
            step()
        t = time.perf_counter()
        for _ in range(steps):
            reps, tot = step()
        dt = (time.perf_counter() - t) / steps
        R, _ = dev.get_RU()
        return dict(ms_per_step=1e3 * dt, updates_per_s=tot.row_updates / dt, solve_ms=tot.solve_ms,
                    max_iters=int(reps["iters"].max()), mean_iters=float(reps["iters"].mean()),
                    groups=int(reps["groups"].sum()),
                    converged=int(np.sum((reps["merit"] <= 1e-6) & (reps["error"] <= eps))),
                    merit_max=float(reps["merit"].max())), R, reps

    # off-diagonal blocks per scene: one per shared rigid body (LOCDYN_Insert, ldy.c:492-568)
    kind, cm, csl = np.asarray(B["kind"]), np.asarray(Cn["master"]), np.asarray(Cn["slave"])
    rig_m, rig_s = kind[cm] == 0, (csl >= 0) & (kind[np.maximum(csl, 0)] == 0)
    deg = np.bincount(np.r_[cm[rig_m], csl[rig_s]], minlength=len(kind))
    row_nnz = np.where(rig_m, deg[cm] - 1, 0) + np.where(rig_s, deg[np.maximum(csl, 0)] - 1, 0)
    nnz_s = np.bincount(cs, weights=row_nnz, minlength=ns)
    n_s = np.bincount(cs, minlength=ns).astype(float)
    nnz_ok = int(row_nnz.sum()) == dev.sizes()[1]

    def sweep_bytes(reps, prec):
        """Algorithmic bytes of all sweeps of a batched solve: per scene, s (9 nnz + 26 n) + 4 nnz + 4 n
        per sweep, s = 4 for the FP32 sweeps of MIXED (the last 2 are FP64 refinement), 8 for FP64."""
        b32 = 4 * (9 * nnz_s + 26 * n_s) + 4 * nnz_s + 4 * n_s
        b64 = 8 * (9 * nnz_s + 26 * n_s) + 4 * nnz_s + 4 * n_s
        it = reps["iters"].astype(float)
        if prec == "mixed":
            ref = np.minimum(it, 2.0)
            return float(np.sum((it - ref) * b32 + ref * b64))
        return float(np.sum(it * b64))

    head, Rh, rh = run(pm, have_model, args.steps, args.warmup)
    log(f"[batch] learned+mixed: {head['ms_per_step']:.1f} ms/step, iters max {head['max_iters']}")
    # the FP64 baseline on its own best shape (the autotuner re-assembles the store for it)
    tune64, tms64 = dev.batch_autotune(pf, trial_iters=20)
    log(f"[batch] autotuned FP64 shape {tune64} ({tms64:.2f} ms for 20 iterations)")
    plain, Rp, rp = run(pf, False, args.steps, args.warmup)
    log(f"[batch] plain FP64: {plain['ms_per_step']:.1f} ms/step, iters max {plain['max_iters']}")
    # per-scene agreement with the FP64 run, relative to each scene's max|R|, where both converged
    errs = []
    ok = (rh["iters"] - 2 < maxiter) & (rp["iters"] < maxiter) & (rh["rows"] > 0)
    for k, (Ra, Rb_) in enumerate(zip(batch.split(Rh, cs, 3), batch.split(Rp, cs, 3))):
        sc = np.abs(Rb_).max() if len(Rb_) else 0.0
        if ok[k] and sc > 0:
            errs.append(float(np.abs(Ra - Rb_).max() / sc))
    errs = errs or [float("nan")]
    out = {
        "metric": "batched solve: contact-block updates/s, learned+mixed vs plain FP64 (config 5)",
        "value": head["updates_per_s"], "unit": "contact-block updates/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": head["ms_per_step"], "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32+f64", "data": "synthetic",
        "config": {"workload": f"{args.scenes} independent lattice scenes, contacts log-uniform [500, 50000], "
                               + ("one launch per colour over all running scenes" if tune[3] == 1
                                  else "one persistent launch"), "scenes": args.scenes, "contacts": ncon,
                   "epsilon": eps, "meritval": 1e-6, "maxiter": maxiter, "shape": list(tune),
                   "shape_fp64": list(tune64),
                   "learned_model": have_model, "learned_mixed": head, "plain_fp64": plain,
                   "speedup_vs_plain_fp64": plain["ms_per_step"] / head["ms_per_step"],
                   "scenes_converged_both": int(ok.sum()),
                   "scene_err_vs_fp64_max": max(errs), "scene_err_vs_fp64_median": float(np.median(errs))},
    }
    if nnz_ok:
        gbs = sweep_bytes(rh, "mixed") / (head["solve_ms"] * 1e-3) / 1e9
        out["roofline"] = dict(bound="hbm", achieved=round(gbs, 2), peak=HBM_PEAK_GBS, unit="GB/s",
                               frac=round(gbs / HBM_PEAK_GBS, 4), traffic=None,
                               kernel="k_usweep (algorithmic bytes of all sweeps / device time of the whole solve)")
        log(f"[batch] whole-solve sweep bytes rate {gbs:.1f} GB/s")
    print(json.dumps(out), flush=True)
    dev.close()


REF_ITERS = 2000


def iso_merit(res, ref_merit):
    """Time to a merit level: per timed step, the first iteration whose merit is at or below
    ref_merit, and the step's device time scaled to that iteration."""
    h = res["merhist_last"]
    k = int(np.argmax(h <= ref_merit)) + 1 if np.any(h <= ref_merit) else None
    res["iters_to_merit_ref"] = k
    res["ms_to_merit_ref"] = float(np.mean(res["device_ms"])) * k / res["iters"][-1] if k else None


def make_parser():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--spheres", type=int, default=200000)
    ap.add_argument("--box", type=float, default=6.0)
    ap.add_argument("--settle", type=int, default=500)
    ap.add_argument("--seed", type=int, default=2)
    ap.add_argument("--friction", type=float, default=0.5)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--lanes", type=int, default=2, help="lanes of a wavefront per contact row (1, 2, 4, 8)")
    ap.add_argument("--save-snapshot", default=None, help="write the settled state to this .npz")
    ap.add_argument("--snapshot", default=None, help="load a settled state instead of settling")
    ap.add_argument("--no-mixed", dest="mixed", action="store_false", help="skip the mixed-precision mode")
    ap.add_argument("--headline", default="learned", choices=["learned", "mixed", "fp64"])
    ap.add_argument("--modes", default="fp64,mixed,learned",
                    help="solve modes to time (comma list of fp64, mixed, learned = learned warm start + mixed)")
    ap.add_argument("--maxiter", type=int, default=5000, help="iteration bound (5000 is the workload's)")
    ap.add_argument("--row-order", type=int, default=3, choices=[0, 1, 2, 3],
                    help="rows inside a colour: 0 by decreasing length, 1 by contact index, 2 by body block, "
                         "3 by the Z-order of the master body's position")
    ap.add_argument("--schedule", type=int, default=0, choices=[0, 1, 2, 3],
                    help="sweep schedule: 0 one launch per colour (XCD-aware), 1 persistent dataflow, "
                         "2 one launch per colour (hardware order), 3 persistent with a barrier per colour")
    ap.add_argument("--wpb", type=int, default=0, choices=[0, 1, 2, 4],
                    help="wavefronts per workgroup of the per-colour sweep launches (0: 2 for FP64, 4 for mixed)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run 3 device copies of MIB MiB (scale of the traffic counters) and exit")
    ap.add_argument("--shape", default="sphere", choices=["sphere", "ellipsoid"],
                    help="packing bodies: spheres (config 2) or rigid ellipsoids (the config-3 variant)")
    ap.add_argument("--workload", default="packing", choices=["packing", "batch", "run", "consumers"],
                    help="packing: the 200k-sphere settled packing (headline); batch: independent scenes; "
                         "run: whole time steps, device loop against host loop")
    ap.add_argument("--run-steps", type=int, default=100, help="config-2 settle steps timed by --workload run")
    ap.add_argument("--run-configs", default="config1,config2_settle", help="cases of --workload run")
    ap.add_argument("--host-loop", action="store_true", help="settle with the host loop instead of the device loop")
    ap.add_argument("--nt-maxiter", type=int, default=500, help="Newton steps bound of --workload consumers (DIAG: 4x)")
    ap.add_argument("--scenes", type=int, default=256)
    ap.add_argument("--ws-new-only", type=int, default=None,
                    help="learned warm start only on new contacts (default: the shipped model's variant)")
    ap.add_argument("--ws-scale", type=float, default=None,
                    help="trust factor of the learned warm start (default: the shipped model's variant)")
    return ap


def main():
    args = make_parser().parse_args()
    if args.ws_new_only is None or args.ws_scale is None:
        from solfec import learned
        meta = learned.load_meta() if os.path.exists(learned.DEFAULT_MODEL) else {}
        if args.ws_new_only is None:
            args.ws_new_only = int(meta.get("new_only", 1))
        if args.ws_scale is None:
            args.ws_scale = float(meta.get("trust", 0.5))
    if args.workload in ("batch", "run", "consumers"):
        if args.gpus != 1:
            print("bench.py measures one GPU", file=sys.stderr)
            sys.exit(2)
        {"batch": bench_batch, "run": bench_run, "consumers": bench_consumers}[args.workload](args)
        return
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU: this build runs the contact solve on a single MI355X "
              f"(multi-GPU sharding is a design only); got --gpus {args.gpus}", file=sys.stderr)
        sys.exit(2)

    import solfec
    h = 1e-3
    b, c = settled_state(args)
    nb = len(b["kind"])
    dev = solfec.Device(0)
    dev.set_layout(args.lanes)
    dev.set_schedule(args.schedule)
    dev.set_sweep_block(args.wpb)
    dev.set_row_order(args.row_order)
    dev.bodies(b["kind"], b["rot"], b["x0"], b["jinv"], b["minv"], b["velo"], b["velo0"])
    dev.contacts(c["master"], c["slave"], c["mpnt"], c["spnt"], c["base"], c["gap"], c["friction"],
                 c["restitution"], c["cohesion"], c["R"], c["U"])
    n, nnz, ncol = 0, 0, 0

    def run_mode(name, precision, steps, warmup, learned_ws=False, ref_merit=None):
        par = solfec.make_param(1e-6, args.maxiter, meritval=1e-6, precision=precision, refine=2)
        # every mode starts from the step's own R0 (reading R back overwrites the uploaded one)
        dev.contacts(c["master"], c["slave"], c["mpnt"], c["spnt"], c["base"], c["gap"], c["friction"],
                     c["restitution"], c["cohesion"], c["R"], c["U"])

        def step():
            dev.assemble(h)
            if learned_ws:
                dev.warm_start()
            rep, rer, mer = dev.solve(par, history=True)
            dev.body_update(h, nb, download=False)
            return rep, mer

        for _ in range(warmup):
            step()
        reps, mers = [], []
        t0 = time.perf_counter()
        for _ in range(steps):
            rep, mer = step()
            reps.append(rep)
            mers.append(mer)
        dt = time.perf_counter() - t0
        R, _ = dev.get_RU()
        n_ = dev.sizes()[0]
        iters = [r.iters for r in reps]
        # merit <= 1e-6 (the workload's meritval): first iteration that reaches it, per step
        first = [int(np.argmax(m <= 1e-6)) + 1 if np.any(m <= 1e-6) else None for m in mers]
        # sweeps run (the merit's closing sweep included) and the FP64 ones among them
        sweeps = [r.row_updates // max(n_, 1) for r in reps]
        fp64 = [r.fp64_sweeps for r in reps]
        res = dict(ms_per_step=1e3 * dt / steps, iters=iters, merit=reps[-1].merit, rel_error=reps[-1].rel_error,
                   updates_per_s=sum(iters) * n_ / dt,
                   merit_at_end=[float(r.merit) for r in reps], first_iter_merit_le_1e6=first,
                   converged=[bool(r.iters < args.maxiter) for r in reps],
                   merit_at=dict((str(k), float(mers[-1][k - 1])) for k in (100, 1000, 2000, 5000)
                                 if k <= len(mers[-1])),
                   device_ms=[float(r.solve_ms) for r in reps], sweeps=sweeps, fp64_sweeps=fp64,
                   launches_per_sweep=int(reps[-1].launches_per_sweep))
        res["merhist_last"] = mers[-1]
        if ref_merit is not None:
            iso_merit(res, ref_merit)
        log(f"[bench] {name}: {res['ms_per_step']:.2f} ms/step, iters={iters}, merit={res['merit']:.3e}, "
            f"first merit<=1e-6: {first}")
        return res, R

    modes = {}
    want = [m.strip() for m in args.modes.split(",") if m.strip()]
    if not args.mixed and "mixed" in want:
        want.remove("mixed")
    R64 = Rmx = None
    if "fp64" in want:
        modes["fp64"], R64 = run_mode("plain FP64 GS", solfec.device.FP64, args.steps, args.warmup)
    # reference merit of the time-to-merit comparison: what plain FP64 reaches after REF_ITERS
    # iterations, on the falling part of its merit curve (it plateaus at ~1.5e-5 further on)
    mref = None
    if "fp64" in modes:
        f = modes["fp64"]
        h64 = f["merhist_last"]
        mref = float(h64[min(REF_ITERS, len(h64)) - 1])
        iso_merit(f, mref)
    if "mixed" in want:
        modes["mixed"], Rmx = run_mode("mixed FP32/FP64 GS", solfec.device.MIXED, args.steps, args.warmup,
                                       ref_merit=mref)
        if R64 is not None:
            modes["mixed"]["err_vs_fp64"] = float(np.abs(Rmx - R64).max() / max(np.abs(R64).max(), 1e-300))
    if "learned" in want:
        from solfec import learned
        learned.install(dev)
        dev.warm_start_mode(args.ws_new_only, args.ws_scale)
        modes["learned"], Rl = run_mode("learned warm start + mixed GS", solfec.device.MIXED, args.steps, args.warmup,
                                        learned_ws=True, ref_merit=mref)
        if R64 is not None:
            modes["learned"]["err_vs_fp64"] = float(np.abs(Rl - R64).max() / max(np.abs(R64).max(), 1e-300))
    head = args.headline if args.headline in modes else ("mixed" if "mixed" in modes else "fp64")
    n, nnz, ncol = dev.sizes()
    value = modes[head]["updates_per_s"]
    ms_per_step = modes[head]["ms_per_step"]
    iters = modes[head]["iters"]

    # roofline of the sweep kernel from the timed steps themselves: algorithmic
    # bytes of every sweep of every timed solve (FP32 sweeps at s = 4, FP64 at
    # s = 8) over the solves' device time (HIP events on the library's stream
    # around each solve, so the termination kernels and launch gaps count
    # against the sweeps); per colour and per launch, the same device time over
    # the colours and over the launches (the trailing colours of a few slices
    # each share one launch)
    is32 = head in ("mixed", "learned")
    hm = modes[head]
    b32 = 4 * (9 * nnz + 26 * n) + 4 * nnz + 4 * n
    b64 = 8 * (9 * nnz + 26 * n) + 4 * nnz + 4 * n
    nbytes = sum((sw - f) * (b32 if is32 else b64) + f * b64 for sw, f in zip(hm["sweeps"], hm["fp64_sweeps"]))
    dev_s = 1e-3 * sum(hm["device_ms"])
    colour_steps = sum(hm["sweeps"]) * ncol
    launches = sum(hm["sweeps"]) * (hm["launches_per_sweep"] or ncol)
    achieved = nbytes / dev_s / 1e9
    kname = ({1: "k_flow<%s>", 3: "k_bar<%s>"}.get(args.schedule, "k_sweep<%s>")) % ("float" if is32 else "double")
    roof = dict(bound="hbm", achieved=round(achieved, 2), peak=HBM_PEAK_GBS, unit="GB/s",
                frac=round(achieved / HBM_PEAK_GBS, 4), traffic=measured_traffic(kname, args),
                kernel=kname, avg_colour_us=round(1e6 * dev_s / max(colour_steps, 1), 3),
                avg_launch_us=round(1e6 * dev_s / max(launches, 1), 3), launches=int(launches),
                launches_per_sweep=int(hm["launches_per_sweep"] or ncol),
                bytes_per_sweep=int(b32 if is32 else b64),
                how="timed steps: algorithmic bytes of all sweeps / device time of the solves (HIP events)")
    log(f"[bench] sweeps of the timed {head} steps: {launches} launches, {1e3 * dev_s:.2f} ms device, "
        f"{achieved:.1f} GB/s")
    copy = stream_copy_rate()
    roof["copy_rate"] = dict(value=round(copy, 1), unit="GB/s", frac_of_copy=round(achieved / copy, 4),
                             what="hipMemcpyAsync device to device, 1 GiB x 10, read + write bytes")
    log(f"[bench] device copy rate {copy:.1f} GB/s")

    cpu = None
    if not args.no_cpu:
        colours = dev.locdyn()["colour"]
        cpu = cpu_baseline(b, c, colours, n, args.cpu_seconds, args.maxiter)
        log(f"[bench] cpu baseline: {cpu['value']:.3e} updates/s")

    for m in modes.values():
        m.pop("merhist_last", None)     # arrays stay out of the JSON line
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "contact-block updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64" if is32 else "f64",
        "data": "synthetic",
        "config": {
            "workload": f"{args.spheres}-{args.shape} packing settled {args.settle} steps, "
                        + {"learned": "learned warm start + mixed FP32/FP64 GPU GS", "mixed": "mixed FP32/FP64 GPU GS",
                           "fp64": "plain FP64 GPU GS"}[head],
            "spheres": args.spheres, "contacts": n, "offdiag_blocks": nnz, "colours": ncol,
            "epsilon": 1e-6, "meritval": 1e-6, "maxiter": args.maxiter, "friction": args.friction,
            "iters_per_step": iters, "merit": modes[head]["merit"], "rel_error": modes[head]["rel_error"],
            "precision": "mixed" if is32 else "fp64",
            "warm_start": ("learned (%s, trust %.2g)" % ("new contacts" if args.ws_new_only else "all contacts", args.ws_scale)
                           if head == "learned" else "previous-step R"),
            "lanes_per_row": args.lanes,
            "waves_per_workgroup": args.wpb if args.wpb else "2 for FP64 solves, 4 for mixed",
            "schedule": ["per-colour launches, XCD-aware", "dataflow", "per-colour launches",
                         "persistent, barrier per colour"][args.schedule],
            "row_order": ["decreasing row length", "contact index", "body block, row length",
                          "Z-order of the master body's position"][args.row_order],
            "modes": modes,
            "speedup_vs_plain_fp64": modes["fp64"]["ms_per_step"] / ms_per_step if "fp64" in modes else None,
            "merit_1e-6_reached": {k: any(f is not None for f in m["first_iter_merit_le_1e6"]) for k, m in modes.items()},
            "merit_ref": {"value": mref, "defined_as": f"plain FP64 merit after {REF_ITERS} iterations"},
            "time_to_merit_note": ("ms_per_step is the time to the reference termination test (bgs.c:1036: relative "
                                   "error <= epsilon and merit <= meritval) or to maxiter; first_iter_merit_le_1e6 "
                                   "lists, per timed step, the first iteration whose merit is <= 1e-6 (null: never "
                                   "within maxiter; the reference GS plateaus near 1.5e-5 on this packing); "
                                   "ms_to_merit_ref is each mode's device time to merit_ref"),
            "speedup_to_merit_ref": (modes["fp64"]["ms_to_merit_ref"] / modes[head]["ms_to_merit_ref"]
                                     if "fp64" in modes and modes[head].get("ms_to_merit_ref") else None),
        },
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    dev.close()


if __name__ == "__main__":
    main()
