"""Reverse-mode derivatives of integrate_step! on the device.

The reference differentiates its Julia step through ChainRulesCore: an
rrule(integrate_step!, ...) returns the primal result and a pullback that
maps the cotangent of the output to the cotangents of the inputs.  This
module is that rule over libfsi_hip.so's discrete adjoint (fsi_adjoint_* in
include/fsi_api.h):

  rrule_integrate_step(state, t, dt, ctx, work)            -> (adj, pullback)
  rrule_integrate_steps(state, t0, dt, n, ctx, work, snapshots)
                                                           -> (adj, pullback)
  pullback(bar: FSIState) -> d_shear_modulus   (bar becomes the input's cotangent)

`centroid_objective` forms J = |x_c(phi) - target|^2 of the phase centroid
(scripts/benchmarks/run_sugiyama_lid_driven_cavity.jl:496-509) with its
cotangent, and `disk_gradient` chains it to dJ/d(shear modulus, disk radius,
disk centre) of the lid-driven cavity (the initial tanh disk,
InterfaceInitialization.jl:46-95).

Coverage: walled domains, ACDI with a fixed gamma, FP64
(FSI_E_UNSUPPORTED -> the library's status error otherwise).  Derivatives
are those of the discrete step along the branches the forward took (upwind
sides, the psi clamp and the safety reset frozen), with exact linear solves:
use tight solver settings (velocity_method "direct").
"""
from __future__ import annotations

import ctypes
import time
from dataclasses import dataclass

import numpy as np

from . import _abi
from .imex import FSIState, IMEXContext, IMEXWorkspace, build_imex_context, allocate_imex_workspace


def zero_cotangent(ctx: IMEXContext) -> FSIState:
    """An FSIState of zeros shaped for ctx (a cotangent of the state)."""
    m = ctx.metrics
    nx, ny = m.Nx, m.Ny
    z = lambda *s: np.zeros(s, order="F")  # noqa: E731
    return FSIState(z(nx + 1, ny), z(nx, ny + 1), z(nx, ny), z(nx, ny), z(nx, ny), z(nx, ny), z(nx, ny))


@dataclass
class AdjointStats:
    steps: int
    advances: int
    recordings: int
    adjoint_steps: int
    snapshots_used: int
    repetitions: int
    device_bytes: float
    replanned: int = 0
    last_shortfall: int = 0
    solve_iterations: tuple = (0, 0, 0)


class StepAdjoint:
    """fsi_adjoint: the tape, the checkpoints and the transposed solves of
    runs on (ctx, work)."""

    def __init__(self, ctx: IMEXContext, work: IMEXWorkspace, snapshots: int = 45):
        self.ctx, self.work = ctx, work
        self.lib = ctx.lib
        h = ctypes.c_void_p()
        st = self.lib.fsi_adjoint_create(ctx.handle, work.handle, int(snapshots), ctypes.byref(h))
        self.handle = h
        _abi.check(self.lib, ctx.handle, st, "fsi_adjoint_create")

    def close(self):
        if self.handle and self.ctx.handle:
            self.lib.fsi_adjoint_destroy(self.handle)
        self.handle = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, st, what):
        _abi.check(self.lib, self.ctx.handle, st, what)

    # -- one step --------------------------------------------------------------
    def record_step(self, state: FSIState, t: float, dt: float):
        d = state._device(self.ctx)
        st = self.lib.fsi_adjoint_record_step(self.handle, d, float(t), float(dt))
        state._device_modified()
        self._check(st, "integrate_step! (taped)")

    def step(self, bar: FSIState) -> float:
        g = ctypes.c_double(0.0)
        d = bar._device(self.ctx)
        st = self.lib.fsi_adjoint_step(self.handle, d, ctypes.byref(g))
        bar._device_modified()
        self._check(st, "adjoint of integrate_step!")
        return g.value

    # -- a run with checkpoints --------------------------------------------------
    def forward(self, state: FSIState, t0: float, dt: float, n_steps: int):
        d = state._device(self.ctx)
        st = self.lib.fsi_adjoint_forward(self.handle, d, float(t0), float(dt), int(n_steps))
        state._device_modified()
        self._check(st, "integrate_step! run (checkpointed)")

    def backward(self, bar: FSIState) -> float:
        g = ctypes.c_double(0.0)
        d = bar._device(self.ctx)
        st = self.lib.fsi_adjoint_backward(self.handle, d, ctypes.byref(g))
        bar._device_modified()
        self._check(st, "adjoint of the run")
        return g.value

    @property
    def stats(self) -> AdjointStats:
        s = _abi.AdjointStatsC()
        self._check(self.lib.fsi_adjoint_get_stats(self.handle, ctypes.byref(s)), "fsi_adjoint_get_stats")
        return AdjointStats(s.steps, s.advances, s.recordings, s.adjoint_steps, s.snapshots_used, s.repetitions,
                            s.device_bytes, s.replanned, s.last_shortfall, tuple(s.solve_iterations))


def rrule_integrate_step(state: FSIState, t: float, dt: float, ctx: IMEXContext, work: IMEXWorkspace):
    """rrule(integrate_step!, state, t, dt, ctx, work): steps `state` in place;
    the pullback maps a cotangent of the new state (in place) to that of the
    old one and returns dJ/d(shear modulus)."""
    adj = StepAdjoint(ctx, work, snapshots=1)
    adj.record_step(state, t, dt)
    return adj, adj.step


def rrule_integrate_steps(state: FSIState, t0: float, dt: float, n_steps: int, ctx: IMEXContext,
                          work: IMEXWorkspace, snapshots: int = 45):
    """The rule of n_steps integrate_step! calls with binomial checkpointing
    (`snapshots` states held)."""
    adj = StepAdjoint(ctx, work, snapshots=snapshots)
    adj.forward(state, t0, dt, n_steps)
    return adj, adj.backward


def centroid_objective(state: FSIState, ctx: IMEXContext, target) -> tuple:
    """(J, bar): J = |x_c(phi) - target|^2 and its cotangent with respect to the state."""
    bar = zero_cotangent(ctx)
    J = ctypes.c_double(0.0)
    st = ctx.lib.fsi_adjoint_centroid_objective(ctx.handle, state._device(ctx), float(target[0]), float(target[1]),
                                                bar._device(ctx), ctypes.byref(J))
    bar._device_modified()
    _abi.check(ctx.lib, ctx.handle, st, "centroid objective")
    return J.value, bar


def plan(n_steps: int, snapshots: int, with_actions: bool = False) -> dict:
    """The checkpoint schedule of a run (host only)."""
    lib = _abi.load_library()
    adv, held, rep, nact = ctypes.c_int64(), ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int64()
    st = lib.fsi_adjoint_plan(int(n_steps), int(snapshots), ctypes.byref(adv), ctypes.byref(held), ctypes.byref(rep),
                              None, 0, ctypes.byref(nact))
    if st != _abi.FSI_OK:
        raise ValueError("fsi_adjoint_plan: invalid arguments")
    out = dict(advances=adv.value, most_held=held.value, repetitions=rep.value, n_actions=nact.value)
    if with_actions:
        acts = np.zeros((nact.value, 4), dtype=np.int32)
        lib.fsi_adjoint_plan(int(n_steps), int(snapshots), None, None, None,
                             acts.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), nact.value, None)
        out["actions"] = acts
    return out


def disk_gradient(case, n_steps: int, target, snapshots: int = 45, t0: float = 0.0) -> dict:
    """dJ/d(shear modulus, radius, centre x, centre y) of J = |x_c(phi(T)) -
    target|^2 for a lid-driven cavity case (setups.lid_driven_cavity with a
    tanh disk): the run with checkpoints, the objective's cotangent, the
    pullback, and the initial disk's derivatives."""
    from .setups import circle_parameter_derivatives

    ctx = build_imex_context(case.metrics, case.bcs, case.models, case.numerics)
    work = allocate_imex_workspace(ctx)
    state = case.state
    p = case.params
    t_start = time.perf_counter()
    adj, pullback = rrule_integrate_steps(state, t0, case.dt, n_steps, ctx, work, snapshots)
    J, bar = centroid_objective(state, ctx, target)
    t_fwd = time.perf_counter()
    dG = pullback(bar)
    t_end = time.perf_counter()
    dr, dcx, dcy = circle_parameter_derivatives(case.grid, p["center"], p["radius"], p["init_epsilon"])
    bphi = bar.phi
    grad = np.array([dG, float(np.sum(bphi * dr)), float(np.sum(bphi * dcx)), float(np.sum(bphi * dcy))])
    out = dict(J=J, grad=grad, stats=adj.stats, forward_s=t_fwd - t_start, backward_s=t_end - t_fwd,
               state=state, bar=bar)
    adj.close()
    return out


# ---- kernel-level entry points (parity of the transposed tendencies) ------------
_BAR_KEYS = ("E_u", "E_v", "I_u", "I_v", "E_phi", "E_sigma_xx", "E_sigma_yy", "E_sigma_xy")
_GRAD_KEYS = ("U", "V", "phi", "sigma_xx", "sigma_yy", "sigma_xy")


def _shapes(ctx):
    nx, ny = ctx.metrics.Nx, ctx.metrics.Ny
    face = {"E_u": (nx + 1, ny), "E_v": (nx, ny + 1), "I_u": (nx + 1, ny), "I_v": (nx, ny + 1),
            "U": (nx + 1, ny), "V": (nx, ny + 1)}
    return lambda k: face.get(k, (nx, ny))


def primal_tendencies(state: FSIState, ctx: IMEXContext) -> dict:
    """The stage tendencies from the point functions the adjoint
    differentiates (E_u, E_v on x-/y-faces, I_u, I_v, E_phi, E_sigma_*)."""
    shape = _shapes(ctx)
    out = {k: np.zeros(shape(k), order="F") for k in _BAR_KEYS}
    arr = (ctypes.POINTER(ctypes.c_double) * 8)(*(_abi.as_ptr(out[k]) for k in _BAR_KEYS))
    st = ctx.lib.fsi_adjoint_primal_tendencies(ctx.handle, state._device(ctx), arr)
    _abi.check(ctx.lib, ctx.handle, st, "fsi_adjoint_primal_tendencies")
    return out


def tendencies_transpose(state: FSIState, ctx: IMEXContext, bar: dict) -> tuple:
    """(d tendencies / d state)^T bar at the state (sigma with the safety
    reset, as the first stage reads it), and the shear-modulus part."""
    shape = _shapes(ctx)
    ins = [None if bar.get(k) is None else np.asfortranarray(bar[k], dtype=np.float64) for k in _BAR_KEYS]
    for k, a in zip(_BAR_KEYS, ins):
        if a is not None and a.shape != shape(k):
            raise ValueError(f"{k}: shape {a.shape}, expected {shape(k)}")
    out = {k: np.zeros(shape(k), order="F") for k in _GRAD_KEYS}
    pin = (ctypes.POINTER(ctypes.c_double) * 8)(*(_abi.as_ptr(a) for a in ins))
    pout = (ctypes.POINTER(ctypes.c_double) * 6)(*(_abi.as_ptr(out[k]) for k in _GRAD_KEYS))
    g = ctypes.c_double(0.0)
    st = ctx.lib.fsi_adjoint_tendencies(ctx.handle, state._device(ctx), pin, pout, ctypes.byref(g))
    _abi.check(ctx.lib, ctx.handle, st, "fsi_adjoint_tendencies")
    return out, g.value
