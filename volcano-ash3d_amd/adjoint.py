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
segment length whose states fit, and when no segment length fits the
budget the reverse sweep uses binomial checkpointing instead.

Binomial checkpointing (schedule="revolve"; Griewank, Optimization Methods
and Software 1, 1992; Griewank & Walther, ACM TOMS 26, 2000): with S free
slots and r repetitions, beta(S, r) = C(S + r, S) steps can be reversed,
each forward step recomputed at most r times.  The reverse sweep of n
steps from a checkpointed state at t0 advances m = max(1, n - beta(S-1, r))
steps, checkpoints that state in a free slot, reverses the last n - m steps
with S - 1 slots, then the first m with S slots and r - 1 repetitions.
Device memory: S + 2 states; work: about r forward runs.  The adjoint
steps recompute their sweep inputs (ash3d_tr_adj_step).

SourceToLoads wraps the pair as a torch.autograd.Function, so an inversion
can use torch optimisers on s.
"""
import math

import numpy as np

from . import transport as _tr
from .state import TransportState


def beta(slots: int, reps: int) -> int:
    """Steps reversible with `slots` checkpoints and `reps` repetitions."""
    return math.comb(slots + reps, slots)


def revolve_actions(n: int, slots: int):
    """The binomial checkpointing schedule for reversing steps 0..n-1 from
    the input of step 0 (kept in slot 0), with `slots` further slots
    (1..slots).  Yields ("advance", t_from, t_to, slot) -- load slot (the
    input of step t_from), run steps t_from..t_to-1; ("save", slot, t);
    ("reverse", t, slot) -- load slot (the input of step t) and apply the
    adjoint of step t.  Steps are reversed in the order n-1, ..., 0."""
    if n <= 0:
        return
    if slots < 0:
        raise ValueError("slots must be >= 0")

    def rec(t0, n, s, r, slot):
        if n == 1:
            yield ("reverse", t0, slot)
            return
        if s == 0:
            # no free slot: recompute from t0 for every step
            for t in reversed(range(t0, t0 + n)):
                if t > t0:
                    yield ("advance", t0, t, slot)
                    yield ("reverse", t, None)
                else:
                    yield ("reverse", t0, slot)
            return
        m = max(1, n - beta(s - 1, r))
        new = slot_of[s]
        yield ("advance", t0, t0 + m, slot)
        yield ("save", new, t0 + m)
        yield from rec(t0 + m, n - m, s - 1, r, new)
        yield from rec(t0, m, s, r - 1, slot)

    r = 0
    while slots > 0 and beta(slots, r) < n:
        r += 1
    # the slot taken at depth with s free slots left: slots - s + 1
    slot_of = {s: slots - s + 1 for s in range(1, slots + 1)}
    yield from rec(0, n, slots, r, 0)


def plan_checkpoints(n_steps: int, per_step: int, state_bytes: float, mem_budget: float, segment: int = None,
                     schedule: str = "segments", revolve_slots: int = None) -> dict:
    """Checkpoint plan of a reverse sweep over n_steps steps: a dict with
    schedule ("segments" or "revolve"), K and n_seg (segments), S (free
    revolve slots) and n_slots (device states to allocate besides the
    context's own).  Segments: every step input kept when they fit the
    budget, else the longest segment length whose states fit, else
    binomial checkpointing.  Raises ValueError when the budget holds fewer
    than the two states binomial checkpointing needs."""
    if schedule not in ("segments", "revolve"):
        raise ValueError(f"schedule must be 'segments' or 'revolve' (got {schedule!r})")
    budget_slots = int(mem_budget // state_bytes) - 1  # states beyond the initial one
    plan = {"schedule": schedule}
    if schedule == "segments":
        if segment:
            K = int(segment)
        elif per_step * n_steps + 1 <= budget_slots:
            K = n_steps  # every step input stored: no replay
        else:  # the longest segments whose states fit
            fits = [k for k in range(1, n_steps + 1) if -(-n_steps // k) + per_step * k <= budget_slots]
            K = max(fits) if fits else None
        if K is not None:
            n_seg = (n_steps + K - 1) // K
            # slot 0: initial state; 1..n_seg: segment inputs; then K step
            # inputs (and 2K snapshots)
            plan.update(K=K, n_seg=n_seg, S=None, n_slots=1 + n_seg + per_step * K)
            return plan
        plan["schedule"] = "revolve"  # no segment length fits
    S = budget_slots - 1 if revolve_slots is None else int(revolve_slots)
    if S < 0:
        raise ValueError(f"the device-memory budget of {mem_budget:.3g} B holds {budget_slots + 1} states of "
                         f"{state_bytes:.3g} B; binomial checkpointing needs at least 2")
    # slot 0: initial state; 1: input of step 0; 2..S+1: free slots
    plan.update(K=n_steps, n_seg=1, S=S, n_slots=2 + S)
    return plan


class SourceInversion:
    def __init__(self, st: TransportState, ivent: int, jvent: int, n_heights: int, n_periods: int,
                 steps_per_period: int, frac, dz, obs_every: int = 1, use_diffusion: bool = True,
                 segment: int = None, itime0: int = 1, device: int = 0, mem_budget: float = 64e9,
                 keep_snapshots: bool = True, schedule: str = "segments", revolve_slots: int = None):
        """schedule "segments" (K from `segment` or the memory budget) or
        "revolve" (binomial checkpointing with revolve_slots free slots, by
        default as many as the budget holds)."""
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
        self.state_bytes = 8.0 * st.ns * (st.nx + 4) * (st.ny + 4) * (st.nz + 4)
        plan = plan_checkpoints(self.n_steps, self.per_step, self.state_bytes, mem_budget, segment, schedule,
                                revolve_slots)
        self.schedule, self.K, self.n_seg, self.S = plan["schedule"], plan["K"], plan["n_seg"], plan["S"]
        self.n_slots = plan["n_slots"]
        self.T = _tr.Transport(st, device=device)
        self.T.checkpoints(self.n_slots)
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
            if self.schedule == "revolve":
                if t == 0:
                    self.T.checkpoint_save(1)
                self.T.transport_step(self.itime0 + t, self.use_diffusion)
                if t in self._obs_index:
                    loads[self._obs_index[t]] = self.T.column_load(self.dz)
                continue
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
        if self.schedule == "revolve":
            self._reverse_revolve(w)
            return self._source_gradient()
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
        return self._source_gradient()

    def _reverse_revolve(self, w):
        """The reverse sweep under binomial checkpointing (slot 1 holds the
        input of step 0, slots 2..S+1 are free).  A slot saved by the
        previous action still equals the state, so it is not loaded back
        (half of the loads: every save is followed by an advance or a
        reverse from that slot)."""
        T, s = self.T, self._s
        self.forward_steps = 0
        held = None  # the slot that the state equals, while it does
        for act in revolve_actions(self.n_steps, self.S):
            if act[0] == "advance":
                _, t_from, t_to, slot = act
                if held != slot:
                    T.checkpoint_load(1 + slot)
                for t in range(t_from, t_to):
                    T.transport_step(self.itime0 + t, self.use_diffusion)
                    self._begin_step(s, t + 1)
                self.forward_steps += t_to - t_from
                held = None
            elif act[0] == "save":
                T.checkpoint_save(1 + act[1])
                held = act[1]
            else:
                _, t, slot = act
                if slot is not None and held != slot:
                    T.checkpoint_load(1 + slot)
                held = None  # the adjoint step rewrites the state's ghosts (Set_BC(1))
                if t in self._obs_index:
                    T.adj_column_load(self.dz, w[self._obs_index[t]])
                T.adj_step(self.itime0 + t, self.use_diffusion)
                T.adj_source(self.ivent, self.jvent, t // self.spp)

    def _source_gradient(self):
        T = self.T
        G = T.adj_source_read(self.P)  # (P, ns, nz+1): sum over steps of dt * lambda(vent)
        grad = np.zeros((self.P, self.H))
        for n in range(self.st.ns):  # bins in order (the Fortran entry sums the same way)
            grad = grad + self.frac[n] * G[:, n, :self.H]
        return grad


def misfit(inv: SourceInversion, s, obs, weight):
    """The weighted squared mismatch J(s) = sum W (L(s) - obs)^2 of the
    column loads and its gradient dJ/ds = dL/ds^T (2 W (L(s) - obs)), by
    one forward run and one reverse sweep.  Returns (J, gradient, loads)."""
    loads = inv.forward(s)
    r = loads - obs
    J = float(np.sum(weight * r * r))
    return J, inv.gradient(2.0 * weight * r), loads


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
