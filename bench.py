"""Benchmark of the per-time-step fluid solve on one MI355X.

Workload (default): the fixed random assembly -- 128^3 periodic box, 195
spheres of 16 cells (solid fraction 0.2) from seeded RSA, Re_m = 50 -- run
as FP32-bulk (the headline) beside plain FP64.  A step is one full time
step (compute_new_timestep, nonlinear + bcset, velstep).

Prints one JSON line with Mgridpt-steps/s, the roofline of the dominant
kernel (algorithmic bytes per launch / average launch time from HIP events
recorded on the solver's stream during the timed region) and a bounded
single-core sample of the CPU oracle as the CPU baseline.
"""
import argparse
import ctypes
import json
import os
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from pureibm_ps2d_fd1d_amd import _abi, configs, solver  # noqa: E402

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, GB/s

# Algorithmic field passes per launch (reads + writes of N-word fields,
# spectral and physical fields both counted as N words).
PASSES = {
    "c2r_ubcp_pbcp": 8, "form_nl": 6, "form_nl_yz": 6, "form_nl_x": 9, "c2r_nlbc": 6, "c2r_onlbc": 6,
    "c2r_visc_omega": 14, "c2r_pgrad": 4, "r2c_force": 6, "velstep": 17, "c2r_inner": 12, "c2r_omega": 8,
    "nl_products": 9, "c2r_inner_nl": 18,
}


def shell_fraction(st):
    """Fraction s of grid nodes touched by the 2x2x2 IBM stencils (surface,
    pressure, inner and outer reversal points), exact for this geometry."""
    nx, my, mz = st.nx, st.my, st.mz
    b = st.bndpts
    nodes = set()
    dr = st.config.dr
    for m in range(st.nbody):
        c = st.xc[m][:, None]
        r = st.radbdy[m]
        for pos in (c + b * r, c + b * r - np.array([[0.5], [0], [0]]), c + b * (r - dr),
                    c + b * (r - dr) - np.array([[0.5], [0], [0]]), c + b * (r + dr)):
            base = np.floor(pos).astype(np.int64)
            for dx in (0, 1):
                for dy in (0, 1):
                    for dz in (0, 1):
                        i = (base[0] + dx - 1) % nx
                        j = (base[1] + dy - 1) % my
                        k = (base[2] + dz - 1) % mz
                        nodes.update((i * my + j) * mz + k)
    return len(nodes) / float(nx * my * mz)


def run_gpu(setup, steps, warmup, ksteps, autotune=True):
    """Timed run (whole steps replayed from HIP graphs, no per-launch timing),
    then ksteps more steps with HIP events around every launch for the
    per-kernel figures.  With autotune the step form is first picked by
    pib_autotune (the state is left as it was)."""
    g = solver.Solver(setup)
    lib = g.lib
    picks = None
    if autotune:
        g.step(2)
        p, ms = g.autotune(10)
        picks = {"fft": p[0], "velstep": p[1], "streams": p[2], "nl": g.tuning(_abi.PIB_TUNE_NL),
                 "ms_per_step": ms}
    g.step(warmup)
    lib.pib_sync(g.ctx)
    t0 = time.perf_counter()
    st = lib.pib_step(g.ctx, steps)
    t1 = time.perf_counter()
    if st != 0:
        raise RuntimeError(lib.pib_last_error(g.ctx).decode())
    lib.pib_set_timing(g.ctx, 1)
    st = lib.pib_step(g.ctx, ksteps)
    if st != 0:
        raise RuntimeError(lib.pib_last_error(g.ctx).decode())
    maxn = 64
    names = ctypes.create_string_buffer(32 * maxn)
    tot = np.zeros(maxn)
    cnt = np.zeros(maxn, dtype=np.int32)
    n = ctypes.c_int()
    lib.pib_kernel_times(g.ctx, maxn, names, _abi.dptr(tot),
                         cnt.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), ctypes.byref(n))
    lib.pib_set_timing(g.ctx, 0)
    kern = {}
    for q in range(n.value):
        nm = names.raw[32 * q:32 * q + 32].split(b"\0")[0].decode()
        kern[nm] = (float(tot[q]), int(cnt[q]))
    res = dict(wall_s=t1 - t0, kernels=kern, ksteps=ksteps, scalars=g.scalars(), forces=g.forces()["force"],
               autotuned=picks)
    g.close()
    return res


# bench kernel names -> the FP32 kernel symbols rocprofv3 reports for them
KERNEL_SYMBOLS = {
    "form_nl": "k_form_nl<float", "velstep": "k_velstep_tiled<float", "forcing_points": "k_forcing_points<float",
    "spread": "k_spread<float", "c2r_visc_omega": "k_c2r_pair<float, 4,", "c2r_ubcp_pbcp": "k_c2r_pair<float, 0,",
    "c2r_pgrad": "k_c2r<float, 6,", "c2r_nlbc": "k_c2r<float, 2,", "r2c_force": "k_r2c_force<float",
    "c2r_inner": "k_c2r<float, 7,", "c2r_omega": "k_c2r<float, 5,", "nl_products": "k_nl_products<float",
    "c2r_inner_nl": "k_c2r<float, 8,",
    "drag_points": "k_drag_points<float",
}
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "hbm_traffic_fluid_step_128cube_fp32.md")
COPY_MIB = 256


def _pmc_pass(counter, args, outdir):
    """One rocprofv3 counter pass over a child bench process; returns
    {kernel symbol: [values per dispatch]}."""
    import csv
    import subprocess
    env = dict(os.environ, TMPDIR="/tmp")
    cmd = ["timeout", "-s", "KILL", "90", "rocprofv3", "--pmc", counter, "-f", "csv", "-d", outdir, "-o", "p",
           "--", sys.executable, os.path.abspath(__file__)] + args
    print("+", " ".join(cmd), file=sys.stderr, flush=True)
    subprocess.run(cmd, cwd="/tmp", env=env, check=True, stdout=subprocess.DEVNULL)
    vals = {}
    for root, _, files in os.walk(outdir):
        for f in files:
            if f.endswith("counter_collection.csv"):
                for r in csv.DictReader(open(os.path.join(root, f))):
                    if r["Counter_Name"] == counter:
                        vals.setdefault(r["Kernel_Name"], []).append(float(r["Counter_Value"]))
    return vals


def collect_traffic():
    """HBM bytes per dispatch of the FP32 step kernels: FETCH_SIZE and
    WRITE_SIZE passes (KiB per dispatch) over a short FP32 run, each scaled by
    what the counter reads for plain device-to-device copies of known size.
    Writes TRAFFIC_FILE.  Runs the GPU only in child processes."""
    work = tempfile.mkdtemp(prefix="pib_pmc_")
    run_args = ["--only", "fp32", "--steps", "3", "--warmup", "2", "--kernel-steps", "1", "--no-cpu", "--no-batched",
                "--no-dem", "--no-adjoint", "--no-scalar", "--no-autotune"]
    copy_args = ["--calib-copy", str(COPY_MIB)]
    fetch = _pmc_pass("FETCH_SIZE", run_args, os.path.join(work, "f"))
    write = _pmc_pass("WRITE_SIZE", run_args, os.path.join(work, "w"))
    cf = _pmc_pass("FETCH_SIZE", copy_args, os.path.join(work, "cf"))
    cw = _pmc_pass("WRITE_SIZE", copy_args, os.path.join(work, "cw"))
    copy_kib = COPY_MIB * 1024.0

    def copy_scale(v):
        # the copy kernels are the only dispatches of that run above 1 MiB
        big = [x for xs in v.values() for x in xs if x > 1024.0]
        return copy_kib / (sum(big) / len(big))

    ff, wf = copy_scale(cf), copy_scale(cw)
    table = {}
    for name, sym in KERNEL_SYMBOLS.items():
        fk = [x for k, xs in fetch.items() if sym in k for x in xs]
        wk = [x for k, xs in write.items() if sym in k for x in xs]
        if fk and wk:
            f_kib, w_kib = sum(fk) / len(fk), sum(wk) / len(wk)
            table[name] = {"fetch_kib": f_kib, "write_kib": w_kib,
                           "hbm_bytes": f_kib * 1024 * ff + w_kib * 1024 * wf, "dispatches": len(fk)}
    cmd = "cd /tmp && export TMPDIR=/tmp && python3 /root/repo/bench.py --collect-traffic"
    lines = ["# HBM traffic per dispatch -- FP32 fluid step kernels, 128^3 bed, 195 spheres", "",
             "Command (one MI355X, gfx950):", "", "```", cmd, "```", "",
             "It runs, as child processes, `rocprofv3 --pmc FETCH_SIZE` and then `--pmc WRITE_SIZE` over",
             "`bench.py " + " ".join(run_args) + "`, and the same two counters over `bench.py --calib-copy "
             + str(COPY_MIB) + "` (three device-to-device copies of " + str(COPY_MIB) + " MiB) to scale them:",
             f"FETCH_SIZE factor {ff:.3f}, WRITE_SIZE factor {wf:.3f} (copy bytes / counter bytes).", "",
             "| kernel | dispatches | FETCH_SIZE KiB | WRITE_SIZE KiB | HBM bytes per dispatch (scaled) |",
             "|---|---:|---:|---:|---:|"]
    for name, r in sorted(table.items(), key=lambda kv: -kv[1]["hbm_bytes"]):
        lines.append(f"| {name} | {r['dispatches']} | {r['fetch_kib']:.0f} | {r['write_kib']:.0f} | "
                     f"{r['hbm_bytes']:.4g} |")
    lines += ["", "Machine-readable copy (read by bench.py for `roofline.traffic`):", "", "```json",
              json.dumps({"fetch_factor": ff, "write_factor": wf, "kernels": table}), "```", ""]
    os.makedirs(os.path.dirname(TRAFFIC_FILE), exist_ok=True)
    with open(TRAFFIC_FILE, "w") as f:
        f.write("\n".join(lines))
    print(json.dumps({"traffic_file": TRAFFIC_FILE, "fetch_factor": ff, "write_factor": wf,
                      "kernels": {k: v["hbm_bytes"] for k, v in table.items()}}))


def traffic_of(kernel):
    """HBM bytes per dispatch of an FP32 kernel from TRAFFIC_FILE, or None."""
    try:
        txt = open(TRAFFIC_FILE).read()
        blob = txt.split("```json", 1)[1].split("```", 1)[0]
        return json.loads(blob)["kernels"][kernel]["hbm_bytes"]
    except (OSError, IndexError, KeyError, ValueError):
        return None


def run_batched(nreal, steps, warmup, tune_steps=10):
    """Batched realizations (FP32-bulk): nreal independent contexts of the C3
    mix (96^3/128^3/160^3 beds).  Each shape's step form is picked by the
    autotuner (pib_autotune, cached per shape).  Times the contexts one after
    the other (pib_step each), launched together on their own streams
    (pib_step_async on every context, then pib_sync), and as one batch
    (pib_batch_step: one graph launch a step for all of them) -- the value."""
    work = tempfile.mkdtemp(prefix="pib_c3_")
    sts = [solver.RunSetup(configs.c3_realization(work, k), precision=_abi.PIB_FP32_BULK) for k in range(nreal)]
    gs = [solver.Solver(st) for st in sts]
    picks = {}
    for g, st in zip(gs, sts):
        g.step(2)
        p, ms = g.autotune(tune_steps)
        if ms > 0:
            picks[f"{st.nx}^3"] = {"fft": p[0], "velstep": p[1], "streams": p[2],
                                   "nl": g.tuning(_abi.PIB_TUNE_NL), "ms_per_step": ms}
    B = solver.Batch(gs)
    B.step(warmup)
    alone = 0.0
    for g in gs:
        t0 = time.perf_counter()
        if g.lib.pib_step(g.ctx, steps) != 0:
            raise RuntimeError(g.lib.pib_last_error(g.ctx).decode())
        alone += time.perf_counter() - t0
    t0 = time.perf_counter()
    for g in gs:
        if g.lib.pib_step_async(g.ctx, steps) != 0:
            raise RuntimeError(g.lib.pib_last_error(g.ctx).decode())
    for g in gs:
        if g.lib.pib_sync(g.ctx) != 0:
            raise RuntimeError(g.lib.pib_last_error(g.ctx).decode())
    streams = time.perf_counter() - t0
    B.step(2)  # both batch graphs captured before the timed region
    t0 = time.perf_counter()
    B.step(steps)
    batch = time.perf_counter() - t0
    pts = sum(st.nx * st.my * st.mz for st in sts)
    B.close()
    for g in gs:
        g.close()
    return {"realizations": nreal, "steps": steps, "grid_points": pts,
            "value": pts * steps / batch / 1e6, "unit": "Mgridpt-steps/s",
            "ms_per_batch_step": batch / steps * 1e3,
            "streams": pts * steps / streams / 1e6,
            "one_after_another": pts * steps / alone / 1e6,
            "autotuned": picks,
            "sizes": [f"{st.nx}^3/{st.nbody}" for st in sts]}


def run_dem(marches, per_march, cpu_substeps):
    """Soft-sphere DEM sub-cycle at the 256^3 bed size: 1200 spheres of 20
    cells (dx = 1/20, box 12.8, solid fraction 0.30), coeff_rest 0.8, random
    velocities in [-0.5, 0.5] and fixed random hydrodynamic terms.  Each march
    is des_time_march with dt = per_march*DTSOLID; the GPU figure is wall time
    over the timed marches (host copies of XC/VELBDY included), the CPU figure
    the oracle on the same bed for cpu_substeps sub-steps."""
    n, box, dx = 1200, 12.8, 1.0 / 20
    xc, rad, vel = configs.moving_bed(n, box, dx, (10.0, 10.0), overlap=0.0, vel=0.5, seed=2024)
    rng = np.random.default_rng(7)
    hyd = dict(pres=rng.normal(0, 0.05, (n, 3)), visc=rng.normal(0, 0.05, (n, 3)),
               mpg=np.array([0.1, 0.0, 0.0]), frame_accln=np.zeros(3))
    kw = dict(dx=dx, xlength=box, ylength=box, zlength=box, coeff_rest=0.8, rhos=2.0, rhof=1.0)
    g = solver.Dem(xc, rad, velbdy=vel, **kw)
    dts = g.state()[4].dtsolid
    g.march(per_march * dts, **hyd)
    r0 = g.state()[4].rebuilds
    t0 = time.perf_counter()
    for _ in range(marches):
        g.march(per_march * dts, **hyd)
    wall = time.perf_counter() - t0
    info = g.state()[4]
    g.close()
    nsub = marches * per_march
    out = {"workload": "1200 spheres d=20 cells in a 12.8^3 periodic box (256^3 grid), phi=0.30",
           "substeps": nsub, "us_per_substep": wall / nsub * 1e6,
           "value": n * nsub / wall / 1e6, "unit": "M particle-substeps/s",
           "contacts_last_substep": info.ncontacts, "candidate_rebuilds": info.rebuilds - r0,
           "dtsolid": dts}
    if cpu_substeps:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle
        oracle.build()
        o = solver.Dem(xc, rad, velbdy=vel, lib=oracle.load(), prefix="orc_", **kw)
        t0 = time.perf_counter()
        o.march(cpu_substeps * dts, **hyd)
        tc = time.perf_counter() - t0
        o.close()
        out["cpu_baseline"] = {"value": n * cpu_substeps / tc / 1e6, "unit": "M particle-substeps/s",
                               "cores": 1, "kind": "port",
                               "sample": f"{cpu_substeps} sub-steps of the same bed on the CPU oracle, {tc:.2f} s"}
    return out


def time_ibm_update(setup, reps=5):
    """Per-step cost of moving the bed in the fluid solver: the solid mask,
    neighbour sets, active reversal pairs and spreading records re-derived on
    the device for shifted centres (grid_nodes_insphere + update_nrpr_array,
    as bcset runs them after des_time_march)."""
    g = solver.Solver(setup)
    rng = np.random.default_rng(3)
    ts = []
    for _ in range(reps):
        xc = setup.xc + rng.uniform(-0.2, 0.2, setup.xc.shape)
        t0 = time.perf_counter()
        g.update_particles(xc)
        ts.append(time.perf_counter() - t0)
    g.close()
    return {"bed": f"{setup.nx}^3, {setup.nbody} spheres, {setup.nbnd} surface points each",
            "ms": 1e3 * float(np.median(ts))}


def time_scalar_solve(setup, reps=20):
    """One species' per-mode implicit solve (pib_scalstep_modes, FP64) on the
    workload's grid from random phif, nlphif, onlphif, ffphi.  Wall time a
    call: the calls are enqueued back to back and the stream drained at the
    end (pib_sync, which also reads the status back); the algorithmic
    traffic is 4 complex reads and 1 complex write a grid point."""
    g = solver.Solver(setup)
    g.scalar_set(0.02, 1.05, 0.97)
    rng = np.random.default_rng(5)
    shp = (setup.nx + 1, setup.my // 2 + 1, setup.mz)
    for f in ("phi", "nlphi", "onlphi", "ffphi"):
        g.set_field(f, rng.standard_normal(shp) + 1j * rng.standard_normal(shp))
    g.scalstep_modes(1)
    g.lib.pib_sync(g.ctx)
    t0 = time.perf_counter()
    for _ in range(reps):
        g.scalstep_modes(1)
    if g.lib.pib_sync(g.ctx) != 0:
        raise RuntimeError(g.lib.pib_last_error(g.ctx).decode())
    t = (time.perf_counter() - t0) / reps
    g.close()
    pts = shp[0] * shp[1] * shp[2]
    return {"grid": f"{setup.nx} x {setup.my // 2 + 1} x {setup.mz} modes", "ms": 1e3 * t,
            "algorithmic_GB_s": 80.0 * pts / t / 1e9}


def _kernel_times(g):
    lib = g.lib
    maxn = 64
    names = ctypes.create_string_buffer(32 * maxn)
    tot = np.zeros(maxn)
    cnt = np.zeros(maxn, dtype=np.int32)
    n = ctypes.c_int()
    lib.pib_kernel_times(g.ctx, maxn, names, _abi.dptr(tot),
                         cnt.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), ctypes.byref(n))
    out = {}
    for q in range(n.value):
        nm = names.raw[32 * q:32 * q + 32].split(b"\0")[0].decode()
        out[nm] = (float(tot[q]), int(cnt[q]))
    return out


def run_scalar(steps=20, warmup=3, ksteps=3):
    """Scalar transport (ISCALON = 1) on the C2 bed (128^3, 195 spheres):
    FP32-bulk fluid, scalar transforms in FP32, scalar fields and mode solve
    in FP64.  Each step is the fluid step's stage calls and then pib_scalstep
    (the whole scalstep: nlphi, the reversal-pair forcing, calc_diffn and the
    surface flux, the mode solve), enqueued back to back and drained at the
    end; the same steps without the scalar give the fluid's share.  Per-kernel
    times of the scalar step from HIP events over ksteps more steps."""
    work = tempfile.mkdtemp(prefix="pib_scal_")
    configs.write_workload(work, "C2", iscalon=1, discr_scheme="upwind")
    st = solver.RunSetup(os.path.join(work, "C2"), precision=_abi.PIB_FP32_BULK)
    g = solver.Solver(st)

    def stage(scal):
        g.compute_new_timestep(1)
        g.nonlinear(1)
        g.velstep(1)
        if scal:
            g.scalstep(1)

    for _ in range(warmup):
        stage(True)
    g.lib.pib_sync(g.ctx)
    res = {}
    for scal in (True, False):
        t0 = time.perf_counter()
        for _ in range(steps):
            stage(scal)
        if g.lib.pib_sync(g.ctx) != 0:
            raise RuntimeError(g.lib.pib_last_error(g.ctx).decode())
        res[scal] = (time.perf_counter() - t0) / steps
    g.lib.pib_set_timing(g.ctx, 1)
    for _ in range(ksteps):
        stage(True)
    g.lib.pib_sync(g.ctx)
    kern = {k: v[0] / max(v[1], 1) * 1e3 for k, v in _kernel_times(g).items()
            if k.startswith("scal_") or k == "phistep"}
    g.lib.pib_set_timing(g.ctx, 0)
    fb, fg, pf, ps = g.scalar_flux()
    sc = st.scalar
    nus = -(fg * sc["char_length"] ** 2) / (6 * sc["voldom"] * g.scalars().maxvolfrac * (sc["phisurf"] - sc["phistream"]))
    hr = g.scalar_get()[0]
    g.close()
    pts = st.nx * st.my * st.mz
    return {"workload": "C2 bed with the scalar (ISCALON = 1, Pr 0.7, phisurf 0, phistream 1), upwind x-fluxes",
            "steps": steps, "ms_per_step_with_scalar": 1e3 * res[True], "ms_per_step_fluid_only": 1e3 * res[False],
            "scalar_ms_per_step": 1e3 * (res[True] - res[False]),
            "value": pts / (res[True] - res[False]) / 1e6, "unit": "Mgridpt-scalar-steps/s",
            "kernel_us_per_launch": kern, "nusselt": nus, "heat_ratio": hr, "phi_fluid_mean": pf}


def time_moving_bed(workload, steps=5, warmup=2):
    """Freely moving bed on the default workload's 128^3/195-sphere bed,
    FP32-bulk fluid with the FP64 DEM: per step nonlinear to the drag, the DEM
    over dt, the IBM update for the moved centres, the forcing, velstep
    (Solver.step_moving).  Wall time a step, host orchestration included."""
    work = tempfile.mkdtemp(prefix="pib_mv_")
    configs.write_workload(work, workload)
    st = solver.RunSetup(os.path.join(work, workload), precision=_abi.PIB_FP32_BULK)
    st.config.move_particles = 1
    st.config.rhos, st.config.rhof = 2.0, 1.0
    g = solver.Solver(st)
    dx = st.config.dx
    L = (st.nx * dx, st.my * dx, st.mz * dx)
    vel0 = np.random.default_rng(11).uniform(-0.1, 0.1, (st.nbody, 3))
    dem = solver.Dem(st.xc, st.radbdy, velbdy=vel0, dx=dx, xlength=L[0], ylength=L[1], zlength=L[2],
                     coeff_rest=0.8, rhos=2.0, rhof=1.0, dia_phys=st.config.char_length)
    g.update_particles(st.xc, vel0)
    g.step_moving(dem, warmup)
    t0 = time.perf_counter()
    g.step_moving(dem, steps)
    wall = time.perf_counter() - t0
    info = dem.state()[4]
    g.close()
    dem.close()
    return {"bed": f"{st.nx}^3, {st.nbody} spheres, FP32-bulk fluid, FP64 DEM", "steps": steps,
            "ms_per_step": wall / steps * 1e3, "dem_substeps_per_step": info.factor}


def run_adjoint(nsteps=64, every=8):
    """Adjoint workload C4 (128^3, 50 spheres of 16 cells, Re_m = 20, SET_MPG
    at the Stokes balance, fixed dt), FP32-bulk: J = total streamwise drag
    after nsteps steps from the uniform stream.  Times the forward steps
    alone (pib_step, graph replay), the forward with a checkpoint every
    `every` steps, and the backward sweep (segment re-runs + adjoint steps)
    that returns dJ/dxc and dJ/dmpg; adjoint_over_forward is the backward
    sweep over the plain forward run."""
    from pureibm_ps2d_fd1d_amd import adjoint
    work = tempfile.mkdtemp(prefix="pib_c4_")
    configs.write_workload(work, "C4")
    st = solver.RunSetup(os.path.join(work, "C4"), precision=_abi.PIB_FP32_BULK)
    P = adjoint.DragProblem(st, nsteps, every)
    P.objective(None, None, keep=False)
    mpg = adjoint.stokes_mpg(P.s)
    P.objective(None, mpg, keep=False)  # warm: graphs, kernel attributes
    s = P._fresh(None, mpg)
    s.step(nsteps)
    s = P._fresh(None, mpg)
    s.lib.pib_sync(s.ctx)
    t0 = time.perf_counter()
    s.step(nsteps)
    t_fwd = time.perf_counter() - t0
    J = float(np.sum(P.wforce * s.forces()["force"]))
    s = P._fresh(None, mpg)
    t0 = time.perf_counter()
    adjoint.forward_checkpointed(s, nsteps, every)
    t_ck = time.perf_counter() - t0
    t0 = time.perf_counter()
    dxc, dmpg = P.gradient()
    t_bwd = time.perf_counter() - t0
    P.close()
    return {"workload": f"C4: {st.nx}^3, {st.nbody} spheres d=16 cells, Re_m=20, SET_MPG, fixed dt, FP32-bulk",
            "steps": nsteps, "checkpoint_every": every, "J": J,
            "forward_ms_per_step": t_fwd / nsteps * 1e3,
            "forward_checkpointed_ms_per_step": t_ck / nsteps * 1e3,
            "backward_ms_per_step": t_bwd / nsteps * 1e3,
            "adjoint_over_forward": t_bwd / t_fwd,
            "grad_norm_xc": float(np.linalg.norm(dxc)), "dJ_dmpg": [float(x) for x in dmpg]}


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(setup, nsteps, threads):
    """The oracle on the same workload, on ``threads`` host threads (its plane
    transforms and mode solves run under OpenMP; sums stay serial): one
    untimed first step, then nsteps timed steps."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    oracle.build()
    lib = oracle.load()
    lib.orc_set_threads(threads)
    o = solver.Solver(setup, lib=lib, prefix="orc_")
    o.step(1)
    t0 = time.perf_counter()
    o.step(nsteps)
    dt = time.perf_counter() - t0
    o.close()
    return dt


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--workload", default="C2")
    ap.add_argument("--cpu-steps", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--only", choices=["both", "fp32", "fp64"], default="both")
    ap.add_argument("--kernel-steps", type=int, default=20, help="steps timed per launch for the roofline")
    ap.add_argument("--calib-copy", type=int, default=0, metavar="MIB",
                    help="only run 3 device-to-device copies of MIB MiB (scale for FETCH_SIZE/WRITE_SIZE)")
    ap.add_argument("--no-batched", action="store_true", help="skip the batched-realization figure")
    ap.add_argument("--batched-steps", type=int, default=100)
    ap.add_argument("--no-dem", action="store_true", help="skip the DEM sub-cycle figure")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the adjoint (C4) figure")
    ap.add_argument("--no-scalar", action="store_true", help="skip the scalar-step (ISCALON) figure")
    ap.add_argument("--no-autotune", action="store_true", help="time the default step form, not the autotuned one")
    ap.add_argument("--dem-marches", type=int, default=10)
    ap.add_argument("--collect-traffic", action="store_true",
                    help="measure HBM bytes per dispatch with rocprofv3 counter passes (child processes)")
    a = ap.parse_args()
    if a.collect_traffic:
        collect_traffic()
        return
    if a.calib_copy:
        import torch
        n = a.calib_copy * (1 << 20) // 4
        x = torch.ones(n, dtype=torch.float32, device="cuda:0")
        y = torch.empty_like(x)
        for _ in range(3):
            y.copy_(x)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes": n * 4, "copies": 3}))
        return
    if a.gpus != 1:
        print(json.dumps({"error": "bench.py measures one GPU; sharding the x-slabs across GPUs is "
                                   "designed but not launched", "n_gpus": a.gpus}))
        sys.exit(2)

    work = tempfile.mkdtemp(prefix="pib_bench_")
    configs.write_workload(work, a.workload)
    run = os.path.join(work, a.workload)
    st64 = solver.RunSetup(run, precision=_abi.PIB_FP64)
    st32 = solver.RunSetup(run, precision=_abi.PIB_FP32_BULK)
    N = st64.nx * st64.my * st64.mz
    s = shell_fraction(st64)

    r32 = run_gpu(st32, a.steps, a.warmup, a.kernel_steps, not a.no_autotune) if a.only in ("both", "fp32") else None
    r64 = run_gpu(st64, a.steps, a.warmup, a.kernel_steps, not a.no_autotune) if a.only in ("both", "fp64") else None
    head = r32 if r32 is not None else r64
    w = 4 if r32 is not None else 8
    ms_step = head["wall_s"] / a.steps * 1e3
    value = N * a.steps / head["wall_s"] / 1e6

    # dominant kernel and its roofline
    kern = head["kernels"]
    dom = max(kern, key=lambda k: kern[k][0])
    tot_ms, launches = kern[dom]
    avg_s = tot_ms / launches / 1e3
    passes = PASSES.get(dom)
    achieved = (passes * w * N / avg_s / 1e9) if passes else None
    roof = {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": (achieved / HBM_PEAK_GBS) if achieved else None,
            "traffic": traffic_of(dom) if w == 4 else None,
            "algorithmic_bytes_per_launch": passes * w * N if passes else None,
            "avg_launch_us": avg_s * 1e6, "launches": launches}
    step_bytes = w * N * (70 + 22 * s)
    step_roof = {"bytes_per_step": step_bytes, "achieved_GBs": step_bytes / (ms_step / 1e3) / 1e9,
                 "frac": step_bytes / (ms_step / 1e3) / 1e9 / HBM_PEAK_GBS, "shell_fraction": s}

    out = {
        "metric": "Mgridpt-steps/s (FP32-bulk fixed-bed step, 128^3, 195 spheres)",
        "value": value, "unit": "Mgridpt-steps/s", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
        "ms_per_step": ms_step, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32" if r32 is not None else "f64", "data": "synthetic (seeded RSA bed, flow started from rest)",
        "config": {"workload": "C2: 128^3 periodic box, 195 spheres d=16 cells (phi=0.1994), Re_m=50, "
                               "fixed bed, AB2/CN step, central x-flux",
                   "grid_points": N, "nbody": st64.nbody, "precision": "fp32-bulk" if r32 else "fp64"},
        "roofline": roof,
        "step_roofline": step_roof,
        "kernel_ms_per_step": {k: v[0] / head["ksteps"] for k, v in sorted(kern.items(), key=lambda kv: -kv[1][0])},
        "autotuned": head["autotuned"],
    }
    if r64 is not None and r32 is not None:
        v64 = N * a.steps / r64["wall_s"] / 1e6
        d64, d32 = r64["scalars"].norm_drag, r32["scalars"].norm_drag
        f64, f32 = r64["forces"], r32["forces"]
        out["fp64"] = {"value": v64, "ms_per_step": r64["wall_s"] / a.steps * 1e3,
                       "speedup_fp32_bulk": value / v64,
                       "mean_drag_rel_err_fp32_vs_fp64": abs(d32 - d64) / abs(d64),
                       "max_particle_drag_rel_err": float(np.max(np.abs(f32[:, 0] - f64[:, 0])) /
                                                          np.max(np.abs(f64[:, 0]))),
                       "norm_drag_fp64": d64, "norm_drag_fp32": d32, "autotuned": r64["autotuned"]}
    if not a.no_batched and a.only in ("both", "fp32"):
        try:
            out["batched_c3"] = run_batched(12, a.batched_steps, 5)
        except Exception as e:  # reported, never fatal for the headline figure
            out["batched_c3"] = {"error": str(e)}
    if not a.no_adjoint and a.only in ("both", "fp32"):
        try:
            out["adjoint_c4"] = run_adjoint()
        except Exception as e:  # reported, never fatal for the headline figure
            out["adjoint_c4"] = {"error": str(e)}
    if not a.no_dem:
        try:
            out["dem"] = run_dem(a.dem_marches, 500, 0 if a.no_cpu else 2000)
            out["dem"]["ibm_update"] = time_ibm_update(st32)
            out["dem"]["moving_bed"] = time_moving_bed(a.workload)
        except Exception as e:  # reported, never fatal for the headline figure
            out["dem"] = {"error": str(e)}
        try:
            out["scalar_solve"] = time_scalar_solve(st32)
        except Exception as e:  # reported, never fatal for the headline figure
            out["scalar_solve"] = {"error": str(e)}
    if not a.no_scalar and a.only in ("both", "fp32"):
        try:
            out["scalar_step_c2"] = run_scalar()
        except Exception as e:  # reported, never fatal for the headline figure
            out["scalar_step_c2"] = {"error": str(e)}
    if not a.no_cpu:
        ncpu = max(1, a.cpu_steps)
        nthr = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or (os.cpu_count() or 1)
        tall = cpu_baseline(st64, 2 * ncpu, nthr)
        t1 = cpu_baseline(st64, ncpu, 1)
        out["cpu_baseline"] = {"value": N * 2 * ncpu / tall / 1e6, "unit": "Mgridpt-steps/s", "cores": nthr,
                               "kind": "port",
                               "sample": f"{2 * ncpu} FP64 step(s) of the same 128^3 bed on the CPU oracle "
                                         f"with {nthr} OpenMP threads, {tall:.1f} s",
                               "nproc": os.cpu_count(), "cpu_model": _cpu_model(),
                               "serial": {"value": N * ncpu / t1 / 1e6, "cores": 1,
                                          "sample": f"{ncpu} FP64 step(s), one thread, {t1:.1f} s"}}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
