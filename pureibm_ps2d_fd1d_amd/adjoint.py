"""Gradients of the fixed-bed drag through the time steps, over the C-ABI.

The reference has no adjoint.  This module drives the hand-written adjoint
step of include/pureibm_gpu.h (``pib_adjoint_step``) with the checkpoint
stack (``pib_checkpoint_push``/``_pop``): the forward run keeps the state at
the start of every ``every``-th step; the backward sweep restores each
segment, re-runs it forward keeping the forward fields of each of its steps
(``pib_adjoint_record``), and reverses the steps last to first.  At most
``ceil(nsteps/every)`` states and ``every`` records are held at once.

``DragObjective`` wraps the pair as a ``torch.autograd.Function`` of the
particle centres and the mean pressure gradient: a thin test harness over the
same C-ABI (torch is used for nothing else).
"""
import math

import numpy as np

from . import solver as _solver


def stokes_mpg(s):
    """Mean pressure gradient (x) that balances the Stokes drag
    3 pi vis dia U of every sphere of Solver ``s`` at the desired mean flow
    U = ufmean_des: nbody * 3 pi vis dia U / (voldom (1 - maxvolfrac))."""
    sc, c = s.scalars(), s.cfg
    u = sc.ufmean_des[0]
    g = s.setup.nbody * 3.0 * math.pi * c.vis * c.dia * u / (c.voldom * (1.0 - sc.maxvolfrac))
    return np.array([g, 0.0, 0.0])


def forward_checkpointed(s, nsteps, every):
    """Run ``nsteps`` steps of Solver ``s``, pushing the state at the start of
    steps 0, every, 2*every, ..."""
    for n0 in range(0, nsteps, every):
        s.checkpoint_push()
        s.step(min(every, nsteps - n0))


def backward(s, nsteps, every, wforce_of_step):
    """Reverse ``nsteps`` steps whose segment checkpoints ``forward_checkpointed``
    pushed.  ``wforce_of_step(n)`` gives the (nbody, 3) weight of step n's
    per-particle force in the objective, or None.  Each segment is restored
    and re-run: a step without a weight keeps its forward fields on the
    record stack (pib_adjoint_record), one with a weight keeps the state at
    its start (its adjoint recomputes the fields, the vorticity included);
    then its steps are reversed last to first.  Returns (dJ/dxc, dJ/dmpg,
    adjoint of the initial umean)."""
    s.adjoint_begin()
    nseg = math.ceil(nsteps / every)
    for seg in reversed(range(nseg)):
        n0, n1 = seg * every, min((seg + 1) * every, nsteps)
        s.checkpoint_pop()  # the state at the start of step n0
        last_w = wforce_of_step(n1 - 1) is not None
        stop = n1 - 1 if last_w else n1
        n = n0
        while n < stop:
            if wforce_of_step(n) is not None:
                s.checkpoint_push()
                s.step(1)
                n += 1
                continue
            k = 1
            while n + k < stop and wforce_of_step(n + k) is None:
                k += 1
            s.adjoint_record(k)
            n += k
        if last_w:  # the state at the start of step n1-1 is current
            s.adjoint_step(wforce_of_step(n1 - 1))
        for n in reversed(range(n0, stop)):
            w = wforce_of_step(n)
            if w is not None:
                s.checkpoint_pop()
            s.adjoint_step(w)
    return s.adjoint_get()


class DragProblem:
    """J(xc, mpg) = sum_m wforce(m, :) . force(m, :) after ``nsteps`` steps of
    the bed in ``setup`` (a RunSetup with SET_MPG), with the centres ``xc``
    (nbody, 3, grid units) and the mean pressure gradient ``mpg`` (3).  The
    flow starts as the uniform stream umean = ufmean_des (from rest, a
    SET_MPG run has no slip at its first step, and the drag normalisation of
    bcset reports a blow-up).  The default weight is (1, 0, 0) per body: the
    total streamwise drag.  ``lib``/``prefix`` select the implementation of
    the C-ABI (the HIP library by default)."""

    def __init__(self, setup, nsteps, every=8, wforce=None, lib=None, prefix="pib_"):
        self.setup, self.nsteps, self.every = setup, int(nsteps), int(every)
        self.lib, self.prefix = lib, prefix
        w = np.zeros((setup.nbody, 3), order="F")
        w[:, 0] = 1.0
        self.wforce = w if wforce is None else np.asfortranarray(wforce, dtype=np.float64)
        self.s = None

    def _fresh(self, xc, mpg):
        if self.s is not None:
            self.s.close()
        kw = {} if self.lib is None else dict(lib=self.lib, prefix=self.prefix)
        s = _solver.Solver(self.setup, **kw)
        if xc is not None:
            s.update_particles(np.asarray(xc, dtype=np.float64))
        sc = s.scalars()
        for n in range(3):
            sc.umean[n] = sc.ufmean_des[n]
            if mpg is not None:
                sc.mpg[n] = float(mpg[n])
        s.set_scalars(sc)
        self.s = s
        return s

    def objective(self, xc=None, mpg=None, keep=True):
        """Run the steps from rest; with ``keep`` the segment checkpoints stay
        on the context for ``gradient``."""
        s = self._fresh(xc, mpg)
        if keep:
            forward_checkpointed(s, self.nsteps, self.every)
        else:
            s.step(self.nsteps)
        f = s.forces()["force"]
        return float(np.sum(self.wforce * f))

    def gradient(self):
        """(dJ/dxc, dJ/dmpg) of the last ``objective(keep=True)``."""
        last = self.nsteps - 1
        dxc, dmpg, _ = backward(self.s, self.nsteps, self.every,
                                lambda n: self.wforce if n == last else None)
        return dxc, dmpg

    def close(self):
        if self.s is not None:
            self.s.close()
            self.s = None


def drag_objective():
    """The torch.autograd.Function (built on first use, so that importing
    this module does not need torch): ``apply(xc, mpg, problem)`` with float64
    tensors xc (nbody, 3) and mpg (3,) returns J as a 0-d tensor; backward
    gives dJ/dxc and dJ/dmpg."""
    import torch

    class DragObjective(torch.autograd.Function):
        @staticmethod
        def forward(ctx, xc, mpg, problem):
            J = problem.objective(xc.detach().cpu().numpy(), mpg.detach().cpu().numpy(), keep=True)
            ctx.problem = problem
            return torch.tensor(J, dtype=torch.float64)

        @staticmethod
        def backward(ctx, gJ):
            dxc, dmpg = ctx.problem.gradient()
            g = gJ.detach().cpu().double()
            return (g * torch.from_numpy(np.ascontiguousarray(dxc)),
                    g * torch.from_numpy(np.ascontiguousarray(dmpg)), None)

    return DragObjective
