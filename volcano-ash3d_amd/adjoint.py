"""Source inversion with the discrete adjoint of the transport step.

The eruption source is a table s[p, h] (kg/km3/h per unit grain fraction)
of emission rates by time period p and height layer h: during the steps of
period p the vent column receives, at the top of each step (after the step
glue, as src/Ash3d.F90:428-431 adds SourceNodeFlux),

    concen(iv, jv, h+1, n, ts0) += dt * s[p, h] * frac[n].

The observations are column loads L_t (ConcentrationCalculator's CloudLoad
without its display threshold, src/Output_Vars.f90:1061-1066) after every
obs_every-th step.  SourceInversion.forward(s) runs the device model and
returns the loads; SourceInversion.gradient(w) returns dJ/ds for
J = sum_t <w_t, L_t>, by one reverse sweep of the discrete adjoint.

Checkpointing.  Every adjoint step needs the input state of its forward
step.  The steps form segments of K; the forward run stores the input of
the first step of every segment but the last, and the input of every step
of the last segment.  The reverse sweep walks the last segment backwards
directly, and replays every other segment from its checkpoint (keeping
each step input) before walking it backwards.  With keep_snapshots (the
default) a kept step also keeps the inputs of its second and third
advection sweeps (ash3d_tr_step_keep), which the adjoint step would
otherwise recompute.  Device memory: n_seg + K (3K with snapshots) + 1
states (the +1 holds the initial state); extra work: one forward step per
step outside the last segment.  A run whose step inputs all fit the device-memory
budget is a single segment and never replays; otherwise K is the longest
segment length whose states fit.

SourceToLoads wraps the pair as a torch.autograd.Function, so an inversion
can use torch optimisers on s.
"""
import math

import numpy as np

from . import transport as _tr
from .state import TransportState


class SourceInversion:
    def __init__(self, st: TransportState, ivent: int, jvent: int, n_heights: int, n_periods: int,
                 steps_per_period: int, frac, dz, obs_every: int = 1, use_diffusion: bool = True,
                 segment: int = None, itime0: int = 1, device: int = 0, mem_budget: float = 64e9,
                 keep_snapshots: bool = True):
        if not 1 <= n_heights <= st.nz + 1:
            raise ValueError(f"n_heights must be in 1..nz+1 (got {n_heights})")
        self.st = st
        self.ivent, self.jvent = int(ivent), int(jvent)
        self.H, self.P, self.spp = int(n_heights), int(n_periods), int(steps_per_period)
        self.frac = np.asarray(frac, dtype=np.float64)
        assert self.frac.shape == (st.ns,)
        self.dz = np.ascontiguousarray(dz, dtype=np.float64)
        assert self.dz.shape == (st.nz,)
        self.use_diffusion = bool(use_diffusion)
        self.itime0 = int(itime0)
        self.n_steps = self.P * self.spp
        self.obs_steps = [t for t in range(self.n_steps) if (t + 1) % obs_every == 0]
        self._obs_index = {t: i for i, t in enumerate(self.obs_steps)}
        # states kept per step of the segment being reversed: its input, and
        # with keep_snapshots the inputs of its second and third advection
        # sweeps (the adjoint step then need not recompute them)
        self.per_step = 3 if keep_snapshots else 1
        self.keep = bool(keep_snapshots)
        if segment:
            self.K = int(segment)
        else:
            state_bytes = 8.0 * st.ns * (st.nx + 4) * (st.ny + 4) * (st.nz + 4)
            slots = max(1 + 2 * self.per_step, int(mem_budget // state_bytes) - 1)
            if self.per_step * self.n_steps + 1 <= slots:
                self.K = self.n_steps  # every step input stored: no replay
            else:  # the longest segments whose states fit
                self.K = max([K for K in range(1, self.n_steps + 1)
                              if -(-self.n_steps // K) + self.per_step * K <= slots]
                             or [max(1, int(math.ceil(math.sqrt(self.n_steps))))])
        self.n_seg = (self.n_steps + self.K - 1) // self.K
        self.T = _tr.Transport(st, device=device)
        # slot 0: initial state; 1..n_seg: segment inputs; then K step inputs
        # (and 2K snapshots)
        self.T.checkpoints(1 + self.n_seg + self.per_step * self.K)
        self.T.step_glue()  # ts0 <- the initial ts1, as the first step's glue would
        self.T.checkpoint_save(0)
        self.T.adj_init(self.P)
        self._s = None  # source table of the last forward run

    def close(self):
        self.T.close()

    # -- forward -------------------------------------------------------------
    def flux(self, s, p):
        """SourceNodeFlux(1:nz+1, 1:ns) of period p."""
        f = np.zeros((self.st.nz + 1, self.st.ns), order="F")
        f[:self.H, :] = np.outer(s[p, :self.H], self.frac)
        return f

    def _inner(self, t):
        return 1 + self.n_seg + t % self.K

    def _snaps(self, t):
        """The two snapshot slots of step t of the segment being reversed."""
        base = 1 + self.n_seg + self.K + 2 * (t % self.K)
        return base, base + 1

    def _step(self, t, keep):
        if keep and self.keep:
            self.T.transport_step_keep(self.itime0 + t, self.use_diffusion, *self._snaps(t))
        else:
            self.T.transport_step(self.itime0 + t, self.use_diffusion)

    def _begin_step(self, s, t):
        """Glue and source insertion of step t (the forward step's input)."""
        self.T.step_glue()
        self.T.source_column_add(self.ivent, self.jvent, self.flux(s, t // self.spp))

    def forward(self, s) -> np.ndarray:
        """Column loads after the observed steps, shape (n_obs, nx, ny)."""
        s = np.asarray(s, dtype=np.float64)
        assert s.shape == (self.P, self.H)
        self._s = s.copy()
        self.T.checkpoint_load(0)
        loads = np.zeros((len(self.obs_steps), self.st.nx, self.st.ny))
        last = self.n_seg - 1
        for t in range(self.n_steps):
            self._begin_step(s, t)
            if t // self.K == last:
                self.T.checkpoint_save(self._inner(t))
            elif t % self.K == 0:
                self.T.checkpoint_save(1 + t // self.K)
            self._step(t, t // self.K == last)
            if t in self._obs_index:
                loads[self._obs_index[t]] = self.T.column_load(self.dz)
        return loads

    # -- reverse ---------------------------------------------------------------
    def gradient(self, w) -> np.ndarray:
        """dJ/ds for J = sum_t <w_t, L_t> at the last forward run's s."""
        if self._s is None:
            raise RuntimeError("gradient() needs a forward() run first")
        w = np.asarray(w, dtype=np.float64)
        assert w.shape == (len(self.obs_steps), self.st.nx, self.st.ny)
        T, s = self.T, self._s
        T.adj_init(self.P)
        for seg in reversed(range(self.n_seg)):
            t0, t1 = seg * self.K, min((seg + 1) * self.K, self.n_steps)
            kept = set()  # steps whose sweep inputs are in their snapshot slots
            if seg < self.n_seg - 1:
                # replay the segment from its checkpoint, keeping every step
                # input (and the sweep inputs of all but its last step)
                T.checkpoint_load(1 + seg)
                for t in range(t0, t1):
                    if t > t0:
                        self._begin_step(s, t)
                    T.checkpoint_save(self._inner(t))
                    if t < t1 - 1:
                        self._step(t, True)
                        kept.add(t)
            elif self.keep:
                kept = set(range(t0, t1))  # the forward run kept them
            for t in reversed(range(t0, t1)):
                if t in self._obs_index:
                    T.adj_column_load(self.dz, w[self._obs_index[t]])
                T.checkpoint_load(self._inner(t))
                if self.keep and t in kept:
                    T.adj_step_kept(self.itime0 + t, self.use_diffusion, *self._snaps(t))
                else:
                    T.adj_step(self.itime0 + t, self.use_diffusion)
                T.adj_source(self.ivent, self.jvent, t // self.spp)
        G = T.adj_source_read(self.P)  # (P, ns, nz+1): sum over steps of dt * lambda(vent)
        grad = np.zeros((self.P, self.H))
        for n in range(self.st.ns):  # bins in order (the Fortran entry sums the same way)
            grad = grad + self.frac[n] * G[:, n, :self.H]
        return grad


def source_to_loads(inv: SourceInversion):
    """torch.autograd.Function over the inversion: loads = f(s)."""
    import torch

    class SourceToLoads(torch.autograd.Function):
        @staticmethod
        def forward(ctx, s):
            loads = inv.forward(s.detach().cpu().double().numpy())
            return torch.from_numpy(loads).to(device=s.device, dtype=s.dtype)

        @staticmethod
        def backward(ctx, g):
            grad = inv.gradient(g.detach().cpu().double().numpy())
            return torch.from_numpy(grad).to(device=g.device, dtype=g.dtype)

    return SourceToLoads.apply
