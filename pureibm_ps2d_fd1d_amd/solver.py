"""Host-side mirror of the reference's per-step operator interface.

The reference's time loop (dtibm_main_scal_flo.F90:426-527) calls, for each
Runge-Kutta stage rks,

    compute_new_timestep(rks)   dependent_functions.F90:1054
    nonlinear(rks)              nl.F90:63  (form_nl, transforms, bcset)
    velstep(sflag, rks)         steptheflow.F90:52

``Solver`` exposes the same three calls over the C-ABI of
include/pureibm_gpu.h, with the state owned by the device.  ``RunSetup`` is
the reference's initialisation (read_nmls + generate_configuration + initflo)
done by the Fortran host library from a ``<RUN>_floparam.in`` namelist file.
"""
import ctypes
import os

import numpy as np

from . import _abi
from ._abi import PibConfig, PibScalars, dptr, PIB_OK


class PibError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"status {code}: {msg}")
        self.code = code


class RunSetup:
    """Namelist input + configuration for one run (Fortran host library)."""

    def __init__(self, run_name, precision=_abi.PIB_FP64):
        self.run_name = run_name
        self._load()
        host = _abi.load_host()
        nb, npnt, nx, my, mz = (ctypes.c_int() for _ in range(5))
        host.pib_host_counts(ctypes.byref(nb), ctypes.byref(npnt), ctypes.byref(nx),
                             ctypes.byref(my), ctypes.byref(mz))
        self.nbody, self.nbnd = nb.value, npnt.value
        self.nx, self.my, self.mz = nx.value, my.value, mz.value
        self.config = PibConfig()
        host.pib_host_get_config(ctypes.byref(self.config))
        self.config.precision = precision
        self.xc = np.zeros((self.nbody, 3), order="F")
        self.radbdy = np.zeros(self.nbody)
        host.pib_host_get_particles(dptr(self.xc), dptr(self.radbdy))
        self.bndpts = np.zeros((3, self.nbnd), order="F")
        self.cd = np.zeros((3, self.nbnd), order="F")
        host.pib_host_get_points(dptr(self.bndpts), dptr(self.cd))
        sc = np.zeros(8)
        host.pib_host_get_scalar(dptr(sc))
        # scal_propt of an ISCALON = 1 deck (scalar_init.F90:58-204)
        self.scalar = dict(iscalon=int(sc[0]), gamma=sc[1], phisurf=sc[2], phistream=sc[3], pr=sc[4],
                           iscal_restart=int(sc[5]), voldom=sc[6], char_length=sc[7])

    def _load(self):
        """(Re)read the namelists and configure; the Fortran host holds one run."""
        host = _abi.load_host()
        nml = (self.run_name + "_floparam.in").encode()
        st = host.pib_host_read(nml, len(nml))
        if st != PIB_OK:
            raise PibError(st, _abi.host_error())
        rn = self.run_name.encode()
        st = host.pib_host_configure(rn, len(rn))
        if st != PIB_OK:
            raise PibError(st, _abi.host_error())

    def finish(self, count_solid):
        """Mask-dependent initial scalars (dt, cf, umean, ufmean_des ...)."""
        self._load()
        s = PibScalars()
        st = _abi.load_host().pib_host_finish(int(count_solid), ctypes.byref(s))
        if st != PIB_OK:
            raise PibError(st, _abi.host_error())
        return s


class Batch:
    """Independent realizations (Solver objects on the HIP library) stepped
    in lockstep by pib_batch_step: one graph launch a step for all of them."""

    def __init__(self, solvers):
        self.lib = _abi.load_gpu()
        self.solvers = list(solvers)
        arr = (ctypes.c_void_p * len(self.solvers))(*[g.ctx.value for g in self.solvers])
        self.b = ctypes.c_void_p()
        st = self.lib.pib_batch_create(arr, len(self.solvers), ctypes.byref(self.b))
        if st != PIB_OK:
            raise PibError(st, "pib_batch_create")

    def _check(self, st):
        if st != PIB_OK:
            raise PibError(st, (self.lib.pib_batch_last_error(self.b) or b"").decode(errors="replace"))

    def step_async(self, nsteps=1):
        self._check(self.lib.pib_batch_step(self.b, nsteps))

    def sync(self):
        self._check(self.lib.pib_batch_sync(self.b))

    def step(self, nsteps=1):
        self.step_async(nsteps)
        self.sync()

    def close(self):
        if self.b:
            self.lib.pib_batch_destroy(self.b)
            self.b = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Solver:
    """One realization on one device context.

    ``lib``/``prefix`` select the library implementing the C-ABI; the default
    is the HIP library (``pib_``).  A missing HIP library raises.
    """

    def __init__(self, setup, lib=None, prefix="pib_"):
        self.lib = lib if lib is not None else _abi.load_gpu()
        self.p = prefix
        self.setup = setup
        self.cfg = setup.config
        self.ctx = ctypes.c_void_p()
        self._call("create", ctypes.byref(self.cfg), ctypes.byref(self.ctx), ctx=False)
        self._call("set_points", self.ctx, setup.nbnd, dptr(setup.bndpts), dptr(setup.cd))
        cs = ctypes.c_longlong()
        self.nrpr_active = np.zeros(setup.nbody, dtype=np.int32)
        self._call("set_particles", self.ctx, setup.nbody, dptr(setup.xc), dptr(setup.radbdy),
                   None, None, ctypes.byref(cs),
                   self.nrpr_active.ctypes.data_as(ctypes.POINTER(ctypes.c_int)))
        self.count_solid = cs.value
        s0 = setup.finish(self.count_solid)
        self._call("set_scalars", self.ctx, ctypes.byref(s0))
        c = self.cfg
        self.nx, self.my, self.mz = c.nx, c.my, c.mz
        self.my2 = c.my // 2 + 1
        self.itrmax = 3 if c.rk3 else 1
        sc = getattr(setup, "scalar", None)
        self.scalar_on = bool(sc and sc["iscalon"] == 1)
        if self.scalar_on:
            # initscal (scalar_init.F90:73, 101, 117-171): gamma = vis/Pr, the
            # ratios 1, phisurf on every body, phi = 1 on planes 1..nx+1
            self.scalar_set(sc["gamma"], 1.0, 1.0)
            self.scalar_surface(np.full(setup.nbody, sc["phisurf"]))
            phi = np.zeros((self.nx + 1, self.my2, self.mz), dtype=np.complex128)
            phi[:, 0, 0] = 1.0
            self.set_field("phi", phi)

    def _fn(self, name):
        return getattr(self.lib, self.p + name)

    def _call(self, name, *args, ctx=True):
        st = self._fn(name)(*args)
        if st != PIB_OK:
            msg = self._fn("last_error")(self.ctx if ctx else None)
            raise PibError(st, (msg or b"").decode(errors="replace"))
        return st

    # ---- the reference's per-step operators ----
    def compute_new_timestep(self, rks=1):
        self._call("compute_new_timestep", self.ctx, rks)

    def nonlinear(self, rks=1):
        self._call("nonlinear", self.ctx, rks)

    def velstep(self, rks=1):
        self._call("velstep", self.ctx, rks)

    def step(self, nsteps=1):
        self._call("step", self.ctx, nsteps)

    def scalar_set(self, gamma, heat_ratio=1.0, heat_ratio_new=None):
        """Diffusivity GAMMA and the x-wrap ratios of the one scalar species
        (scalar_data.F90; heat_ratio_new is the value the solve uses at the
        x wrap, heat_ratio the one of the previous solve).  Allocates the
        scalar fields, zero, on first use."""
        hrn = heat_ratio if heat_ratio_new is None else heat_ratio_new
        self._call("scalar_set", self.ctx, gamma, heat_ratio, hrn)

    def scalar_get(self):
        """-> (heat_ratio, heat_ratio_new, heat_ratio_old, gamma)."""
        out = (ctypes.c_double * 4)()
        self._call("scalar_get", self.ctx, out)
        return tuple(out)

    def scalstep_modes(self, rks=1):
        """The per-mode implicit solve of scalstep (stepthescalar.F90:67-91)."""
        self._call("scalstep_modes", self.ctx, rks)

    def scalar_surface(self, phisurf):
        """Surface value of the scalar on each body (phisurfall, one species;
        scalar_init.F90:101 sets PHISURF on every body)."""
        a = np.ascontiguousarray(np.broadcast_to(np.asarray(phisurf, dtype=np.float64), (self.setup.nbody,)))
        self._call("scalar_surface", self.ctx, dptr(a))

    def scalstep(self, rks=1):
        """The whole scalstep of stage rks (stepthescalar.F90:51-95): the
        nonlinear term, the immersed-boundary forcing and the mode solve.
        Call after the stage's velstep."""
        self._call("scalstep", self.ctx, rks)

    def scalar_flux(self):
        """-> (flux_body (nbody,), flux_global, phi_fluid_mean, phi_solid_mean)
        of the last scalstep (bc_scalar.F90:524-551, 185; nl_allphi.F90:380-419)."""
        fb = np.zeros(max(self.setup.nbody, 1))
        out = np.zeros(4)
        self._call("scalar_flux", self.ctx, dptr(fb), dptr(out))
        return fb[: self.setup.nbody], out[0], out[1], out[2]

    def nonlinear_begin(self, rks=1):
        """nonlinear(rks) up to where bcset moves particles (form_nl, the
        transforms, calc_visc, compute_mpg_at_n)."""
        self._call("nonlinear_begin", self.ctx, rks)

    def nonlinear_finish(self, rks=1):
        """The rest of bcset: calc_pgrad and the reversal forcing."""
        self._call("nonlinear_finish", self.ctx, rks)

    def step_moving(self, dem, nsteps=1):
        """Time steps of a freely moving bed: per stage rks, compute_new_timestep,
        nonlinear up to the drag, des_time_march over dt with the step's
        PRES/VISC/mpg/frame_accln (``dem``, a ``Dem``), grid_nodes_insphere and
        update_nrpr_array for the moved centres, the forcing, velstep
        (dtibm_main_scal_flo.F90:426-527 with boundary_condition.F90:539-554),
        then the scalar step when the deck has ISCALON = 1 (:489-515).
        The context must have been created with move_particles set."""
        if not self.cfg.move_particles:
            raise ValueError("the run set-up must set move_particles")
        for _ in range(nsteps):
            for rks in range(1, self.itrmax + 1):
                self.compute_new_timestep(rks)
                self.nonlinear_begin(rks)
                f, s = self.forces(), self.scalars()
                xc, vel, angv = dem.march(s.dt, pres=f["pres"], visc=f["visc"], mpg=np.array(s.mpg),
                                          frame_accln=np.array(s.frame_accln))
                self.update_particles(xc, vel, angv)
                self.nonlinear_finish(rks)
                self.velstep(rks)
                if self.scalar_on:
                    self.scalstep(rks)

    def update_particles(self, xc, velbdy=None, angv=None):
        """Moved particles: what bcset does after des_time_march when particles
        move (boundary_condition.F90:539-554) -- the solid mask of
        grid_nodes_insphere (dependent_functions.F90:783-963) and the active
        reversal pairs of update_nrpr_array (:966-990) for the new centres, with
        the new surface velocities.  xc, velbdy, angv: (nbody, 3).  Returns
        (count_solid, nrpr_active)."""
        st = self.setup
        xc = np.asfortranarray(xc, dtype=np.float64)
        vel = None if velbdy is None else np.asfortranarray(velbdy, dtype=np.float64)
        av = None if angv is None else np.asfortranarray(angv, dtype=np.float64)
        cs = ctypes.c_longlong()
        nr = np.zeros(st.nbody, dtype=np.int32)
        self._call("set_particles", self.ctx, st.nbody, dptr(xc), dptr(st.radbdy),
                   dptr(vel) if vel is not None else None, dptr(av) if av is not None else None,
                   ctypes.byref(cs), nr.ctypes.data_as(ctypes.POINTER(ctypes.c_int)))
        self.count_solid = cs.value
        self.nrpr_active = nr
        return cs.value, nr

    # ---- adjoint of the fixed-bed step and the checkpoint stack ----
    def checkpoint_push(self):
        """Copy the forward state (u, p, nl and its physical copy, scalars)
        onto the context's checkpoint stack."""
        self._call("checkpoint_push", self.ctx)

    def checkpoint_pop(self):
        """Restore the forward state from the top of the stack and drop it."""
        self._call("checkpoint_pop", self.ctx)

    def checkpoint_steps(self, nsteps):
        """``nsteps`` steps, pushing the state at the start of each."""
        self._call("checkpoint_steps", self.ctx, nsteps)

    def adjoint_begin(self):
        """Allocate and zero the adjoint state and the accumulated gradients."""
        self._call("adjoint_begin", self.ctx)

    def adjoint_step(self, wforce=None):
        """Reverse the step that starts from the current forward state; wforce
        (nbody, 3) weighs that step's per-particle force in the objective."""
        w = None if wforce is None else np.asfortranarray(wforce, dtype=np.float64)
        if w is not None and w.shape != (self.setup.nbody, 3):
            raise ValueError("wforce must be (nbody, 3)")
        self._call("adjoint_step", self.ctx, dptr(w) if w is not None else None)

    def adjoint_record(self, nsteps):
        """``nsteps`` steps, keeping each step's forward fields for its
        adjoint step (pib_adjoint_record)."""
        self._call("adjoint_record", self.ctx, nsteps)

    def adjoint_get(self):
        """-> (dJ/dxc (nbody, 3), dJ/dmpg (3,), adjoint of umean (3,))."""
        dxc = np.zeros((self.setup.nbody, 3), order="F")
        dmpg, dum = np.zeros(3), np.zeros(3)
        self._call("adjoint_get", self.ctx, dptr(dxc), dptr(dmpg), dptr(dum))
        return dxc, dmpg, dum

    def adjoint_set_umean(self, dumean):
        a = np.ascontiguousarray(dumean, dtype=np.float64)
        self._call("adjoint_set_umean", self.ctx, dptr(a))

    # ---- restart files (restart_funcs.F90) ----
    def save_restart(self, xc=None, velbdy=None, ferror_array=None):
        """save_part_restart (restart_funcs.F90:69-128): writes the reference's
        restart files <RUN>_u_<c>.rst, _nl_<c>.rst, _p_<c>.rst,
        _sphere_config.rst and <RUN>_RESTART for this context's state, the
        count c alternating 1, 2 (and <RUN>_scal_<c>.rst for an ISCALON
        deck).  xc/velbdy default to the set-up's centres
        at rest (pass the current ones for a moving bed); ferror_array is the
        NERR_STEPS drag-error history (by default the context's own, from
        pib_get_ferror_history, when its length is NERR_STEPS; ones when it is
        not).  Returns c."""
        st = self.setup
        st._load()
        host = _abi.load_host()
        nerr = max(1, host.pib_host_get_nsteps_hint())
        u, nl, p = (np.asfortranarray(self.field(k)) for k in ("u", "nl", "p"))
        s = self.scalars()
        xc = np.asfortranarray(st.xc if xc is None else xc, dtype=np.float64)
        vel = np.asfortranarray(np.zeros((st.nbody, 3)) if velbdy is None else velbdy, dtype=np.float64)
        if ferror_array is not None:
            ferr = np.ascontiguousarray(ferror_array, dtype=np.float64)
        elif nerr <= _abi.PIB_NERR_MAX and self._ferror_len() == nerr:
            ferr = self.ferror_history(nerr)[0]
        else:
            ferr = np.ones(nerr)
        if ferr.size < nerr:
            raise ValueError(f"ferror_array needs NERR_STEPS = {nerr} entries")
        cnt = ctypes.c_int()
        rn = st.run_name.encode()
        rc = host.pib_host_save_restart(rn, len(rn), ctypes.byref(s), st.nbody, dptr(xc), dptr(st.radbdy),
                                        dptr(vel), self.nx, self.my2, self.mz, dptr(u), dptr(nl), dptr(p),
                                        dptr(ferr), ctypes.byref(cnt))
        if rc != PIB_OK:
            raise PibError(rc, _abi.host_error())
        if st.scalar["iscalon"] == 1:
            self._save_scal_restart(cnt.value, nerr)
        return cnt.value

    def load_restart(self):
        """read_restart (restart_funcs.F90:411-616): reads the files of the
        count named in <RUN>_RESTART into this context (u, nl, p and the
        restart scalars; the centres too when they differ from the set-up's;
        the scalar's file too for an ISCALON deck with ISCAL_RESTART = 1).
        Returns (count, ferror_array, xc, velbdy)."""
        st = self.setup
        st._load()
        host = _abi.load_host()
        nerr = max(1, host.pib_host_get_nsteps_hint())
        nx, my2, mz = self.nx, self.my2, self.mz
        u = np.zeros((nx + 1, my2, mz, 3), dtype=np.complex128, order="F")
        nl = np.zeros_like(u)
        p = np.zeros((nx, my2, mz), dtype=np.complex128, order="F")
        xc = np.zeros((st.nbody, 3), order="F")
        rad = np.zeros(st.nbody)
        vel = np.zeros((st.nbody, 3), order="F")
        ferr = np.zeros(nerr)
        s = self.scalars()
        cnt = ctypes.c_int()
        rn = st.run_name.encode()
        rc = host.pib_host_read_restart(rn, len(rn), ctypes.byref(s), st.nbody, dptr(xc), dptr(rad), dptr(vel),
                                        nx, my2, mz, dptr(u), dptr(nl), dptr(p), dptr(ferr), ctypes.byref(cnt))
        if rc != PIB_OK:
            raise PibError(rc, _abi.host_error())
        self.set_field("u", u)
        self.set_field("nl", nl)
        self.set_field("p", p)
        self.set_scalars(s)
        if not np.array_equal(xc, st.xc) or vel.any():
            self.update_particles(xc, vel)
        if st.scalar["iscalon"] == 1 and st.scalar["iscal_restart"] == 1:
            self._load_scal_restart(cnt.value, nerr)
        if nerr <= _abi.PIB_NERR_MAX:  # the history the next steps extend
            self.set_ferror_history(nerr, ferr)
        return cnt.value, ferr, xc, vel

    def _save_scal_restart(self, count, nerr, nu_error_array=None, nu=(1.0, 0.0)):
        """<RUN>_scal_<count>.rst (restart_funcs.F90:315-333): phif, nlphif,
        phisurfall and the heat ratios; the Nusselt history as given (ones
        when not kept)."""
        st = self.setup
        host = _abi.load_host()
        phif, nlphif = (np.asfortranarray(self.field(k)) for k in ("phi", "nlphi"))
        hr = self.scalar_get()
        ps = np.full(max(st.nbody, 1), st.scalar["phisurf"])
        nue = np.ones(nerr) if nu_error_array is None else np.ascontiguousarray(nu_error_array, dtype=np.float64)
        nu3 = np.array([nu[0], nu[1], 0.0])
        hr2 = np.array([hr[0], hr[2]])
        rn = st.run_name.encode()
        rc = host.pib_host_save_scal_restart(rn, len(rn), count, self.nx, self.my2, self.mz, st.nbody,
                                             dptr(phif), dptr(nlphif), dptr(ps), dptr(nu3), dptr(hr2), dptr(nue))
        if rc != PIB_OK:
            raise PibError(rc, _abi.host_error())

    def _load_scal_restart(self, count, nerr):
        st = self.setup
        host = _abi.load_host()
        nx, my2, mz = self.nx, self.my2, self.mz
        phif = np.zeros((nx + 1, my2, mz), dtype=np.complex128, order="F")
        nlphif = np.zeros_like(phif)
        ps = np.zeros(max(st.nbody, 1))
        nu3, hr2, nue = np.zeros(3), np.zeros(2), np.zeros(nerr)
        rn = st.run_name.encode()
        rc = host.pib_host_read_scal_restart(rn, len(rn), count, nx, my2, mz, st.nbody, dptr(phif), dptr(nlphif),
                                             dptr(ps), dptr(nu3), dptr(hr2), dptr(nue))
        if rc != PIB_OK:
            raise PibError(rc, _abi.host_error())
        self.set_field("phi", phif)
        self.set_field("nlphi", nlphif)
        self.scalar_surface(ps[: st.nbody])
        self.scalar_ratios(hr2[0], hr2[0], hr2[1])
        return nue, nu3

    def scalar_ratios(self, heat_ratio, heat_ratio_new, heat_ratio_old):
        r = np.array([heat_ratio, heat_ratio_new, heat_ratio_old], dtype=np.float64)
        self._call("scalar_ratios", self.ctx, dptr(r))

    # ---- state access in the reference layout ----
    def _shape(self, name):
        nx, my, mz, my2 = self.nx, self.my, self.mz, self.my2
        return {
            "u": (nx + 1, my2, mz, 3), "nl": (nx + 1, my2, mz, 3), "onl": (nx + 1, my2, mz, 3),
            "ff": (nx + 1, my2, mz, 3), "p": (nx, my2, mz),
            "ubcp": (nx + 1, my, mz, 3), "nlbcp": (nx + 1, my, mz, 3),
            "onlbcp": (nx + 1, my, mz, 3), "diffn": (nx + 1, my, mz, 3),
            "fr": (nx + 1, my, mz, 3), "omega": (nx + 1, my, mz, 3),
            "pbcp": (nx, my, mz), "ppr": (nx, my, mz, 3), "mask": (nx, my, mz),
            "phi": (nx + 1, my2, mz), "nlphi": (nx + 1, my2, mz), "onlphi": (nx + 1, my2, mz),
            "ffphi": (nx + 1, my2, mz),
            "au": (nx + 1, my2, mz, 3), "ap": (nx, my2, mz), "anl": (nx + 1, my2, mz, 3),
        }[name]

    def field(self, name):
        shp = self._shape(name)
        if name in _abi.COMPLEX_FIELDS:
            a = np.zeros(shp, dtype=np.complex128, order="F")
        else:
            a = np.zeros(shp, dtype=np.float64, order="F")
        self._call("get_field", self.ctx, _abi.FIELDS[name], a.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
        return a

    def set_field(self, name, arr):
        shp = self._shape(name)
        dt = np.complex128 if name in _abi.COMPLEX_FIELDS else np.float64
        a = np.asfortranarray(np.asarray(arr, dtype=dt).reshape(shp, order="F"))
        self._call("set_field", self.ctx, _abi.FIELDS[name], a.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))

    def scalars(self):
        s = PibScalars()
        self._call("get_scalars", self.ctx, ctypes.byref(s))
        return s

    def set_scalars(self, s):
        self._call("set_scalars", self.ctx, ctypes.byref(s))

    def autotune(self, nsteps=10, force=False):
        """pib_autotune: time the step forms of this bed and keep the
        fastest (cached per shape for the process).  -> (picks, ms): picks
        = (PIB_TUNE_FFT, PIB_TUNE_VELSTEP, PIB_TUNE_STREAMS), ms the GPU ms
        per step of the pick (0 when it came from the cache)."""
        picks = (ctypes.c_int * 3)()
        ms = ctypes.c_double()
        self._call("autotune", self.ctx, nsteps, 1 if force else 0, picks, ctypes.byref(ms))
        return tuple(picks), ms.value

    def ferror_history(self, n=None):
        """bcset's drag-error history (boundary_condition.F90:828-833):
        (ferror_array newest first, ferror_hist).  n is NERR_STEPS, by
        default the length last set (PIB_NERR_DEFAULT from creation)."""
        if n is None:
            n = getattr(self, "_nerr", _abi.PIB_NERR_DEFAULT)
        a = np.empty(n, dtype=np.float64)
        h = ctypes.c_double()
        self._call("get_ferror_history", self.ctx, n, a.ctypes.data_as(_abi._PD), ctypes.byref(h))
        return a, h.value

    def _ferror_len(self):
        """Length of the context's drag-error history (the n that
        pib_get_ferror_history accepts), found by asking for NULL outputs."""
        for n in (getattr(self, "_nerr", None), _abi.PIB_NERR_DEFAULT):
            if n and self._fn("get_ferror_history")(self.ctx, n, None, None) == PIB_OK:
                return n
        return None

    def set_ferror_history(self, n, ferror_array=None):
        """A history of n entries, newest first (None: ones, as at start-up)."""
        a = None if ferror_array is None else np.ascontiguousarray(ferror_array, dtype=np.float64)
        if a is not None and a.size != n:
            raise ValueError("ferror_array needs n entries")
        self._call("set_ferror_history", self.ctx, n, None if a is None else a.ctypes.data_as(_abi._PD))
        self._nerr = n

    def tuning(self, key):
        """pib_get_tuning: the current value of a PIB_TUNE_* key."""
        v = ctypes.c_int()
        self._call("get_tuning", self.ctx, key, ctypes.byref(v))
        return v.value

    def forces(self):
        nb = self.setup.nbody
        out = [np.zeros((nb, 3), order="F") for _ in range(4)]
        self._call("get_forces", self.ctx, *[dptr(a) for a in out])
        return dict(zip(("force", "pres", "visc", "torq"), out))

    def close(self):
        if self.ctx:
            self._fn("destroy")(self.ctx)
            self.ctx = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Dem:
    """Soft-sphere DEM of a freely moving bed, one context per realization.

    Mirrors des_time_march (collision_mod.F90:49-344) over the C-ABI of
    include/pureibm_dem.h: the constructor is des_time_march(.TRUE.) (coarse
    grid, DES_INIT_ARRAYS, CFASSIGN, first neighbour search) and ``march`` is
    des_time_march(.FALSE.) followed by IBMUPDATE, returning the new XC,
    VELBDY and ANGV.  Per-particle arrays are (nbody, 3); xc and radbdy are in
    grid units as the fluid solver holds them.  ``lib``/``prefix`` select the
    implementation (HIP library by default; a missing library raises).
    """

    def __init__(self, xc, radbdy, *, dx, xlength, ylength, zlength, velbdy=None, dy=None, dz=None,
                 foffset=0.0, rhos=1.0, rhof=1.0, dia_phys=1.0, coeff_rest=0.9, kn=0.0,
                 only_dem=False, lib=None, prefix="pib_"):
        self.lib = lib if lib is not None else _abi.load_gpu()
        self.p = prefix
        xc = np.asfortranarray(xc, dtype=np.float64)
        rad = np.ascontiguousarray(radbdy, dtype=np.float64)
        self.nbody = n = xc.shape[0]
        if xc.shape != (n, 3) or rad.shape != (n,):
            raise ValueError("xc must be (nbody, 3) and radbdy (nbody,)")
        c = _abi.PibDemConfig()
        c.nbody = n
        c.only_dem = 1 if only_dem else 0
        c.dx = dx
        c.dy = dx if dy is None else dy
        c.dz = dx if dz is None else dz
        c.xlength, c.ylength, c.zlength = xlength, ylength, zlength
        c.foffset, c.rhos, c.rhof = foffset, rhos, rhof
        c.dia_phys, c.coeff_rest, c.kn = dia_phys, coeff_rest, kn
        self.cfg = c
        self.only_dem = bool(only_dem)
        vel = None if velbdy is None else np.asfortranarray(velbdy, dtype=np.float64)
        if vel is not None and vel.shape != (n, 3):
            raise ValueError("velbdy must be (nbody, 3)")
        self.ctx = ctypes.c_void_p()
        st = self._fn("create")(ctypes.byref(c), dptr(xc), dptr(rad), dptr(vel) if vel is not None else None,
                                ctypes.byref(self.ctx))
        if st != PIB_OK:
            msg = self._fn("last_error")(self.ctx) if self.ctx else b"create failed"
            self.close()  # a failed create still hands back a context to destroy
            raise PibError(st, (msg or b"").decode(errors="replace"))

    def _fn(self, name):
        return getattr(self.lib, self.p + "dem_" + name)

    def _raise(self, st):
        msg = self._fn("last_error")(self.ctx) if self.ctx else b"create failed"
        raise PibError(st, (msg or b"").decode(errors="replace"))

    def _call(self, name, *args):
        st = self._fn(name)(*args)
        if st != PIB_OK:
            self._raise(st)

    def march(self, dt, pres=None, visc=None, mpg=None, frame_accln=None):
        """des_time_march(.FALSE.) + IBMUPDATE -> (xc, velbdy, angv), each (nbody, 3)."""
        n = self.nbody

        def arr(a, shape):
            return None if a is None else np.asfortranarray(np.broadcast_to(a, shape), dtype=np.float64)
        pres, visc = arr(pres, (n, 3)), arr(visc, (n, 3))
        mpg, fa = arr(mpg, (3,)), arr(frame_accln, (3,))
        xc = np.zeros((n, 3), order="F")
        vel = np.zeros((n, 3), order="F")
        angv = np.zeros((n, 3), order="F")
        self._call("march", self.ctx, float(dt), *(dptr(a) if a is not None else None for a in (pres, visc, mpg, fa)),
                   dptr(xc), dptr(vel), dptr(angv))
        return xc, vel, angv

    def state(self):
        """(pos, vel, omega, contact_force, info) in physical units."""
        n = self.nbody
        out = [np.zeros((n, 3), order="F") for _ in range(4)]
        info = _abi.PibDemInfo()
        self._call("get_state", self.ctx, *(dptr(a) for a in out), ctypes.byref(info))
        return (*out, info)

    def set_state(self, pos, vel):
        pos = np.asfortranarray(pos, dtype=np.float64)
        vel = np.asfortranarray(vel, dtype=np.float64)
        self._call("set_state", self.ctx, dptr(pos), dptr(vel))

    def close(self):
        if self.ctx:
            self._fn("destroy")(self.ctx)
            self.ctx = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
