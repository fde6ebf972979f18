"""Input decks for the fixed-bed workloads, written in the reference's formats.

* ``<RUN>_floparam.in`` -- Fortran namelists FLOANDGRID / PART_PROPT / FLAGS
  (+ FILES, PSD, GCG) as read by read_nmls (initialize.F90:274-359).
* ``<RUN>_sphr_center.inp`` -- one header line, then ``x y z r`` per sphere in
  diameters, read by generate_configuration 'default'
  (initialize_flo.F90:1485-1605): xc = x*DBYDX.

Random assemblies come from this project's own seeded random sequential
addition (RSA): centres uniform in [0, LYBYD)^3, accepted when the periodic
minimum-image distance to every accepted centre is at least ``min_gap``
diameters.  The reference builds random beds differently (lattice + DEM
shaking, gener_config_random); RSA is only a source of reproducible inputs.
"""
import math
import os

import numpy as np


def _too_close(pts, c, lybyd, gap2):
    """Rows of candidates ``c`` (k, 3) within the gap of any of ``pts`` (n, 3),
    by the periodic minimum image."""
    if len(pts) == 0 or len(c) == 0:
        return np.zeros(len(c), bool)
    d = pts[None, :, :] - c[:, None, :]
    d -= lybyd * np.round(d / lybyd)
    return np.min(np.einsum("kij,kij->ki", d, d), axis=1) < gap2


def rsa_centers(nbody, lybyd, seed, min_gap=1.0, max_tries=50_000_000, batch=4096):
    """Seeded RSA of ``nbody`` unit-diameter spheres in a periodic cube.

    Candidates are drawn in batches but accepted strictly in draw order: a
    candidate is tested against every centre accepted before it, including
    those of its own batch, so the result is the one-at-a-time RSA of the
    same random stream.  A periodic k-d tree of the centres accepted before
    the batch settles most candidates; those near the gap, and the centres
    accepted within the batch, are tested exactly."""
    from scipy.spatial import cKDTree

    rng = np.random.Generator(np.random.PCG64(seed))
    gap2 = min_gap * min_gap
    pts = np.zeros((0, 3))
    tries = 0
    while len(pts) < nbody:
        if tries >= max_tries:
            raise RuntimeError(f"RSA jammed after {len(pts)} of {nbody} spheres")
        cand = rng.uniform(0.0, lybyd, size=(batch, 3))
        tries += batch
        if len(pts):
            tree = cKDTree(pts, boxsize=lybyd)
            dist, idx = tree.query(cand, k=1, distance_upper_bound=min_gap * 1.001)
            ok = dist > min_gap * 1.000001
            near = np.nonzero((~ok) & (dist >= min_gap * 0.999999))[0]
            if len(near):
                ok[near] = ~_too_close(pts, cand[near], lybyd, gap2)
        else:
            ok = np.ones(batch, bool)
        new = []
        for r in np.nonzero(ok)[0]:
            c = cand[r]
            if new and _too_close(np.array(new), c[None, :], lybyd, gap2)[0]:
                continue
            new.append(c)
            if len(pts) + len(new) == nbody:
                break
        if new:
            pts = np.vstack([pts, np.array(new)])
    return pts


def _fmt(v):
    if isinstance(v, bool):
        return "T" if v else "F"
    if isinstance(v, str):
        return "'" + v + "'"
    if isinstance(v, (list, tuple)):
        return ", ".join(_fmt(x) for x in v)
    if isinstance(v, float):
        return repr(v).replace("e", "d") if "e" in repr(v) else repr(v) + "d0"
    return str(v)


def write_namelists(path, groups):
    with open(path, "w") as f:
        for gname, entries in groups.items():
            f.write(f" &{gname}\n")
            for k, v in entries.items():
                f.write(f" {k} = {_fmt(v)},\n")
            f.write(" /\n")


def write_case(directory, run_name, *, my, lybyd, dbydx, centers=None, input_type=None,
               vol_frac1=None, re=20.0, vis=0.012, cfl=0.1, cfl_max=0.5, adaptive=True,
               discr_scheme="center", limiter="none", rk3=False, ramp_frac_steps=0,
               set_mpg=False, mpg=(0.0, 0.0, 0.0), revernorm=1, dr=1.0, f2=1.6,
               include_frmeanfluid=False, dorpr=True, dobnd=False, imove=0, rhos=1.0, coeff_rest=1.0,
               iscalon=0, phisurf=0.0, phistream=1.0, pr=0.7, nerr_steps=100, tol_ferror=None):
    """Write ``<directory>/<run_name>_floparam.in`` (+ sphr_center file).
    imove=1 makes a freely moving bed (soft-sphere collisions, particle
    density rhos, restitution coeff_rest).  iscalon=1 adds the scalar
    (SCAL_PROPT: one species, PHISURF, PHISTREAM, PR_OR_SC_NU = pr)."""
    os.makedirs(directory, exist_ok=True)
    base = os.path.join(directory, run_name)
    if input_type is None:
        input_type = "default" if centers is not None else "simple"
    part = {
        "NBODY": int(len(centers)) if centers is not None else 1,
        "DR": float(dr), "F1": 1.6, "F2": float(f2), "LYBYD": float(lybyd),
        "DBYDX": float(dbydx), "DIA_PHYS": 1.0, "RHOS": float(rhos),
        "INPUT_TYPE": input_type, "DISCR_SCHEME": discr_scheme, "LIMITER": limiter,
        "COLLISION_TYPE": "softsphere" if imove else "none",
    }
    if vol_frac1 is not None:
        part["VOL_FRAC1"] = float(vol_frac1)
    groups = {
        "FLOANDGRID": {
            "MY": int(my), "TEND": 1.0, "NERR_STEPS": int(nerr_steps), "RAMP_FRAC_STEPS": int(ramp_frac_steps),
            "VIS": float(vis), "RE": float(re), "RHOF": 1.0, "CFL": float(cfl),
            "CFL_MAX": float(cfl_max), "MPG": [float(x) for x in mpg],
            "FLO_ANG": [0.0, 90.0, 90.0],
        },
        "PART_PROPT": part,
        "FLAGS": {
            "IMOVE": int(imove), "ISCALON": int(iscalon), "IRESTART": 0, "ADAPTIVE": bool(adaptive),
            "RK3": bool(rk3), "DORPR": bool(dorpr), "DOBND": bool(dobnd),
            "INCLUDE_FRMEANFLUID": bool(include_frmeanfluid), "SET_MPG": bool(set_mpg),
            "REVERNORM": int(revernorm),
        },
        "FILES": {"SAVEITNS": 10, "WRITE_OUTPUT": False},
        "PSD": {"PSD_TYPE": "discrete", "NPHASES": 1},
        "GCG": {"COEFF_REST": float(coeff_rest)},
    }
    if tol_ferror is not None:
        groups["FLOANDGRID"]["TOL_FERROR"] = float(tol_ferror)
    if iscalon:
        groups["FLAGS"]["ISCAL_RESTART"] = 0
        groups["SCAL_PROPT"] = {"NSPMX": 1, "PHISTREAM": float(phistream), "PHISURF": float(phisurf),
                                "PR_OR_SC_NU": float(pr), "SOURCEPRESENT": False, "SETPHIMEAN": False}
    write_namelists(base + "_floparam.in", groups)
    if centers is not None:
        with open(base + "_sphr_center.inp", "w") as f:
            f.write(" x y z r (diameters)\n")
            for c in np.asarray(centers):
                f.write(f" {c[0]:.15f} {c[1]:.15f} {c[2]:.15f} 0.5\n")
    return base


def simple_cubic(directory, run_name, *, my, dbydx, **kw):
    """One sphere of DBYDX cells centred in a my^3 box ('simple')."""
    lybyd = my / dbydx
    vol_frac1 = math.pi / (6.0 * lybyd ** 3)
    return write_case(directory, run_name, my=my, lybyd=lybyd, dbydx=dbydx,
                      input_type="simple", vol_frac1=vol_frac1, **kw)


def random_bed(directory, run_name, *, my, dbydx, nbody, seed, **kw):
    """``nbody`` spheres of DBYDX cells placed by seeded RSA ('default')."""
    lybyd = my / dbydx
    centers = rsa_centers(nbody, lybyd, seed)
    return write_case(directory, run_name, my=my, lybyd=lybyd, dbydx=dbydx,
                      centers=centers, **kw)


C3_PHI_CAP = 0.35  # random sequential addition jams near 0.38


C3_CELL_RE_MAX = 4.5  # Re_m / d: the resolution the scheme needs (measured)


def c3_realization(directory, k):
    """Realization k (0..11) of the batched workload: grids 96^3/128^3/160^3 in
    turn, sphere diameter d ~ U{12..20} cells, solid fraction ~ U[0.1, 0.4]
    capped at 0.35 (RSA jams near 0.38), Re_m ~ U[10, min(100, 4.5 d)], all
    drawn from seed 1000 + k.  The x-fluxes use the MUSCL/Rusanov 'upwind'
    scheme, as the reference's own input decks do (sample_inputs/floparam.in,
    MIS1_floparam.in: DISCR_SCHEME = 'upwind').  The cap on Re_m is the grid
    resolution the method needs: with the pseudo-spectral y-z derivatives of
    the products (no dealiasing), beds whose cell Reynolds number Re_m/d
    exceeds 5 diverge within a few hundred steps in FP64 and FP32
    alike, with any time step, with RK3 and with a ramped start (Re_m/d =
    5.1 and 5.7 diverged; 4.6 and below ran 320 steps); d = 20 reaches
    Re_m = 90."""
    rng = np.random.Generator(np.random.PCG64(1000 + k))
    my = (96, 128, 160)[k % 3]
    d = int(rng.integers(12, 21))
    phi = min(float(rng.uniform(0.1, 0.4)), C3_PHI_CAP)
    re = 10.0 + (min(100.0, C3_CELL_RE_MAX * d) - 10.0) * float(rng.uniform(0.0, 1.0))
    lybyd = my / d
    nbody = max(1, int(round(phi * lybyd ** 3 / (math.pi / 6.0))))
    return random_bed(directory, f"C3_{k}", my=my, dbydx=d, nbody=nbody, seed=k + 1, re=re,
                      discr_scheme="upwind")


# Workloads of the benchmark table (box, spheres, Re); C2 is the headline.
WORKLOADS = {
    "C1": dict(kind="simple", my=64, dbydx=16, re=20.0, adaptive=False, ramp_frac_steps=0),
    "C2": dict(kind="random", my=128, dbydx=16, nbody=195, seed=20261202, re=50.0),
    # the adjoint workload: SET_MPG with a fixed dt, so that J(xc, mpg) is smooth
    "C4": dict(kind="random", my=128, dbydx=16, nbody=50, seed=20261204, re=20.0, set_mpg=True,
               adaptive=False),
}


def write_workload(directory, name, **over):
    w = dict(WORKLOADS[name])
    w.update(over)
    kind = w.pop("kind")
    if kind == "simple":
        return simple_cubic(directory, name, **w)
    return random_bed(directory, name, **w)


def moving_bed(n, box, dx, rad_range, overlap=0.01, vel=0.5, seed=0):
    """Freely moving bed for the soft-sphere DEM: n spheres placed by random
    sequential addition in a periodic cube of side ``box`` (physical units),
    pairs allowed to overlap by ``overlap`` of r_a + r_b so that contacts exist
    from the start, velocities uniform in [-vel, vel].  Returns xc (n,3) and
    radbdy (n,) in grid units of spacing dx, and velbdy (n,3)."""
    rng = np.random.default_rng(seed)
    rad = rng.uniform(rad_range[0], rad_range[1], n)
    pos = np.zeros((n, 3))
    k = tries = 0
    while k < n:
        tries += 1
        if tries > 200 * n:
            raise RuntimeError("bed too dense for random sequential addition")
        p = rng.uniform(0.0, box, 3)
        if k:
            d = pos[:k] - p
            d -= box * np.round(d / box)
            if np.any(np.sqrt((d * d).sum(1)) < (1.0 - overlap) * (rad[:k] + rad[k]) * dx):
                continue
        pos[k] = p
        k += 1
    return np.asfortranarray(pos / dx + 1.0), rad, np.asfortranarray(rng.uniform(-vel, vel, (n, 3)))
