"""FWI gradients over the GPU adjoint (awcpml_checkpoint_run / awcpml_adjoint).

The reference lists the adjoint method and checkpointing as work in
progress (README.md:297) and its forward solve is the function to be
differentiated (src/shared/acoustic_2D_xPU.jl:133-444).  Here:

  misfit_gradient(plan, obs)   J = 1/2 sum (traces - obs)^2 and dJ/dfact
  velocity_gradient(gf, vel, dt)  dJ/dvel = dJ/dfact * 2 dt^2 vel
  solve_traces(vel, problem)   traces as a differentiable torch function
                               (torch.autograd.Function: the counterpart of
                               a ChainRulesCore rrule of the solve)

dt, the CPML coefficients and the source wavelets are held fixed when the
velocity changes (they are set once from the starting model, as an FWI
loop does); fact = dt^2 vel^2 is the only model-dependent input.
"""
import numpy as np

from .solver import Plan


def misfit_gradient(plan, obs, every=0):
    """(J, dJ/dfact) for observed traces ``obs`` (one (nt, nrecs) array per
    shot, or one array for a single-shot plan)."""
    if isinstance(obs, np.ndarray) and obs.ndim == 2:
        obs = [obs]
    plan.checkpoint_run(every)
    res = []
    J = 0.0
    for s in range(plan.nshots):
        r = plan.traces(s) - np.asarray(obs[s], dtype=np.float64)
        J += 0.5 * float(np.sum(r * r))
        res.append(r)
    return J, plan.adjoint(res)


def velocity_gradient(fact_grad, vel, dt):
    """Chain rule through fact = dt^2 vel^2."""
    return fact_grad * (2.0 * dt * dt) * np.asarray(vel, dtype=np.float64)


class Problem:
    """A plan plus the fixed quantities of the forward map vel -> traces."""

    def __init__(self, S, nshots=1, src_positions=None, dtype="f64", every=0, wavelets=None):
        self.S = S
        self.dt = S.dt
        self.every = every
        # number of checkpointed forward runs so far: a backward call checks
        # that the plan still holds the checkpoints of its own forward
        self.forward_count = 0
        self.plan = Plan(S.n, S.halo, dtype=dtype, nshots=nshots)
        self.plan.set_model(S.fact, S.inv_d, S.inv_d2)
        for a in range(S.ndim):
            self.plan.set_cpml(a, S.coef[a])
        for s in range(nshots):
            pos = S.possrcs if src_positions is None else src_positions[s]
            wav = S.dt2srctf if wavelets is None else wavelets[s]
            self.plan.set_sources(s, pos, wav)
            self.plan.set_receivers(s, S.posrecs)

    def set_velocity(self, vel):
        v = np.asarray(vel, dtype=np.float64)
        fact = np.asfortranarray((self.dt * self.dt) * (v * v))
        self.plan.set_model(fact, self.S.inv_d, self.S.inv_d2)

    def forward(self, vel):
        """Checkpointed forward run for ``vel``; returns the traces
        (nshots, nt, nrecs) and the run's number, which :meth:`pullback`
        needs."""
        self.set_velocity(vel)
        P = self.plan
        P.checkpoint_run(self.every)
        self.forward_count += 1
        return np.stack([P.traces(s) for s in range(P.nshots)]), self.forward_count

    def pullback(self, run, trace_bar):
        """dJ/dfact for the cotangents ``trace_bar`` of forward run ``run``.
        Raises if another forward ran on this problem since: the plan's
        checkpoints and model then belong to that run, and the adjoint would
        mix the two models."""
        if run != self.forward_count:
            raise RuntimeError(
                f"stale backward: forward run {run} was followed by run {self.forward_count} on "
                "the same Problem; the checkpoints of run %d are gone (call backward before the "
                "next forward, or use one Problem per pending backward)" % run)
        return self.plan.adjoint(trace_bar)

    def close(self):
        self.plan.close()


def _function():
    import torch

    class SolveTraces(torch.autograd.Function):
        """vel (float64 CPU tensor, grid shape) -> traces (nshots, nt, nrecs).

        forward: one checkpointed solve on the GPU; backward: the adjoint
        solve for the incoming trace cotangents."""

        @staticmethod
        def forward(ctx, vel, problem):
            tr, run = problem.forward(vel.detach().cpu().numpy())
            ctx.problem = problem
            ctx.run = run
            ctx.save_for_backward(vel)
            return torch.from_numpy(tr)

        @staticmethod
        def backward(ctx, tbar):
            (vel,) = ctx.saved_tensors
            problem = ctx.problem
            tb = tbar.detach().cpu().numpy().astype(np.float64)
            gf = problem.pullback(ctx.run, [tb[s] for s in range(tb.shape[0])])
            gv = velocity_gradient(gf, vel.detach().cpu().numpy(), problem.dt)
            return torch.from_numpy(gv).to(vel.dtype), None

    return SolveTraces


_SolveTraces = None


def solve_traces(vel, problem):
    """Differentiable traces of ``problem`` for the velocity tensor ``vel``.

    The checkpoints of the forward call are reused by the backward call, so
    each forward must be followed by its backward before the next forward
    on the same problem; a backward after a later forward raises
    RuntimeError instead of returning a gradient of the wrong model."""
    global _SolveTraces
    if _SolveTraces is None:
        _SolveTraces = _function()
    return _SolveTraces.apply(vel, problem)
