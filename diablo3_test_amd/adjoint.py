"""Gradients through the channel substep: checkpointed adjoint runs on the
HIP adjoint kernels, and a torch.autograd.Function on top of them.

The loss is the time average of the wall shear stress on both walls
(statistics.f90:143-146, see stats.py) over N steps of the driver loop
(for/diablo.f90:98-108):

    L = (1/N) sum_{n=1..N} (tau_lower(S_n) + tau_upper(S_n)) / 2,

a linear function of the mean profile Re cu1(0, 0, :) of each step's state.
Its gradient with respect to the initial state S_0 comes from the adjoint
substeps run backwards.  The forward run stores the state every
`checkpoint_every` steps; the backward sweep restores each checkpoint,
recomputes that segment storing every substep's input, and runs the adjoint
substeps over it.
"""
import numpy as np
import torch

from .binding import ChannelSolver, FIELD_IDS, PER_SCALAR
from .grid import metrics

STATE = ("cu1", "cu2", "cu3", "cp", "cth", "cf1", "cf2", "cf3", "cfth")


def wall_shear_weights(p, gy):
    """w with (tau_lower + tau_upper)/2 = w . Re cu1(0,0,:)."""
    gyf, _, _ = metrics(gy)
    ny = p["ny"]
    nu = 1.0 / p["re"]
    w = np.zeros(ny + 2)
    h1 = gyf[2] - gyf[1]
    h2 = gyf[ny] - gyf[ny - 1]
    w[2] += 0.5 * nu / h1
    w[1] -= 0.5 * nu / h1
    w[ny] -= 0.5 * nu / h2
    w[ny - 1] += 0.5 * nu / h2
    return w


class WallShearRun:
    """N steps from a given initial state on one solver handle (FP64, DNS or LES)."""

    def __init__(self, solver, nsteps, checkpoint_every=10, time0=0.0, time_step0=0):
        if not isinstance(solver, ChannelSolver):
            raise TypeError("solver must be a ChannelSolver")
        self.s = solver
        self.n = int(nsteps)
        self.k = max(1, int(checkpoint_every))
        self.t0 = float(time0)
        self.ts0 = int(time_step0)
        p = solver.p
        self.dt = p.delta_t
        pd = {"ny": p.ny, "re": p.re}
        self.w = wall_shear_weights(pd, solver.gy)

    def _step(self, n):
        """Step n (0-based) of the driver loop: time_step = ts0 + n + 1, time fixed over the substeps."""
        t = self.t0 + n * self.dt
        ts = self.ts0 + n + 1
        for rk in (1, 2, 3):
            self.s.rk_chan_1(rk, t, ts)

    def _nseg(self):
        return (self.n + self.k - 1) // self.k

    def _snapshot_step(self, n, n0):
        """Step n, saving each substep's input in the snapshot slots of the
        segment that starts at step n0 (what the adjoint substep reads: cu1..3
        and cth, or the whole state for the LES adjoint, which reruns the
        substep)."""
        t = self.t0 + n * self.dt
        ts = self.ts0 + n + 1
        light = not self.s.p.use_les
        base = self._nseg()  # substep snapshots follow the segment checkpoints
        for rk in (1, 2, 3):
            self.s.checkpoint(base + 3 * (n - n0) + rk - 1, save=True, light=light)
            self.s.rk_chan_1(rk, t, ts)

    def loss(self, state=None, save_checkpoints=False):
        """Run N steps from `state` (None: the handle's current state), return L.
        With save_checkpoints the run also keeps the segment checkpoints and the
        substep snapshots of the last segment, which the adjoint sweep then
        needs no recomputation for."""
        if state is not None:
            self.s.set_state(state)
        last0 = (self._nseg() - 1) * self.k
        total = 0.0
        for n in range(self.n):
            if save_checkpoints and n % self.k == 0:
                self.s.checkpoint(n // self.k, save=True)
            if save_checkpoints and n >= last0:
                self._snapshot_step(n, last0)
            else:
                self._step(n)
            total += float(self.w @ self.s.mean_profile("cu1"))
        return total / self.n

    def gradient(self, state=None, fetch=True):
        """(L, dL/dS_0), dL/dS_0 a dict of reference-layout complex arrays.
        state None starts from the handle's current state; fetch False leaves
        the gradient in the handle's adjoint state (and returns None for it).
        Afterwards the forward state is a mix of recomputed fields; checkpoint
        slot 0 holds the initial state (restore it with checkpoint(0, False))."""
        L = self.loss(state, save_checkpoints=True)
        nseg = self._nseg()
        base = nseg  # substep snapshots follow the segment checkpoints
        s = self.s
        s.adj_zero()
        wN = self.w / self.n
        for seg in reversed(range(nseg)):
            n0, n1 = seg * self.k, min(self.n, (seg + 1) * self.k)
            if seg < nseg - 1:  # the forward run kept the last segment's snapshots
                s.checkpoint(seg, save=False)
                for n in range(n0, n1):
                    self._snapshot_step(n, n0)
            for n in reversed(range(n0, n1)):
                t = self.t0 + n * self.dt
                ts = self.ts0 + n + 1
                s.adj_add_mean("cu1", wN)  # the loss term of step n's output
                for rk in (3, 2, 1):
                    s.checkpoint(base + 3 * (n - n0) + rk - 1, save=False)
                    s.adj_rk_chan_1(rk, t, ts)
        return L, (s.adj_get_state(STATE) if fetch else None)


def _load(solver, fields, rest):
    """Put (cu1, cu2, cu3, cth) and the held fields `rest` into the handle:
    device tensors through their data pointers on torch's current stream,
    host tensors and arrays through host copies."""
    for k, v in list(rest.items()) + list(fields.items()):
        if isinstance(v, torch.Tensor) and v.is_cuda:
            solver.set_field_device(k, v.detach().contiguous())
        elif isinstance(v, torch.Tensor):
            solver.set_field(k, v.detach().numpy())
        else:
            solver.set_field(k, v)


class WallShearLoss(torch.autograd.Function):
    """L(cu1, cu2, cu3, cth) for a WallShearRun; the other state fields (cp,
    the RK carry) are held fixed at `rest`.  Tensors are complex128 in the
    reference layout; backward runs the checkpointed HIP adjoint.  Device
    tensors go in and the gradients come out through the handle's device
    entry points on torch's current stream, with no host copy of a field;
    host tensors take the host copies."""

    @staticmethod
    def forward(ctx, cu1, cu2, cu3, cth, run, rest):
        fields = {"cu1": cu1.detach(), "cu2": cu2.detach(), "cu3": cu3.detach(), "cth": cth.detach()}
        ctx.run, ctx.rest, ctx.fields = run, rest, fields
        ctx.device = cu1.device
        _load(run.s, fields, rest)
        return torch.tensor(run.loss(None), dtype=torch.float64, device=cu1.device)

    @staticmethod
    def backward(ctx, g):
        run = ctx.run
        _load(run.s, ctx.fields, ctx.rest)
        run.gradient(None, fetch=False)
        out = []
        for k in ("cu1", "cu2", "cu3", "cth"):
            if ctx.device.type == "cuda":
                grad = run.s.adj_get_field_device(k, device=ctx.device)
            else:
                grad = torch.from_numpy(run.s.adj_get_field(k))
            out.append(g * grad)
        return out[0], out[1], out[2], out[3], None, None
