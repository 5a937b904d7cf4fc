"""Hierarchical (block time-bin) KDK of collisionless particles on the device.

Restates the synchronized hierarchical DMO KDK of the reference
(StepOrchestrator::executeHierarchicalBlockWithDispatcher,
time_integration.cpp:1356-1600, and docs/time_integration.md "P6 synchronized
hierarchical DMO KDK") over the TreePM coordinator:

- state: comoving positions x and u = p / a (p = a^2 dx/dt) at a common epoch;
- a block opens with a synchronized Tree+PM solve; each particle's interval
  follows the gravity criterion dt_i = eta sqrt(eps / (|A_i| / a^3)), the
  block spans at most 2^M min_i dt_i, the quantum is span / 2^M and particle
  i sits in the largest bin b <= M with quantum 2^b <= dt_i;
- per fine tick: opening half-kicks of the bins whose interval starts, a drift
  of every particle (x += a0 u0 int dt/a^2, u *= a0/a1), a short-range-only
  solve on the bins whose interval ends and their closing half-kicks;
- PM half-kicks of every particle at the two block surfaces; the last tick's
  solve is the synchronized Tree+PM solve of the block end;
- kick increments are A int(dt/a) / a_endpoint over half intervals of the
  block's scale-factor grid (half ticks at equal cosmic time), with the
  64-sample midpoint integrals of time_integration.cpp:242-266 and
  bisection for a(t) as advanceScaleFactorByCosmicTime (:3469-3516).

Forces come from chui_treepm_solve_adaptive_device (short-range lanes per
source row, PM lanes per active slot, short_range_only on fine ticks); kicks
and drifts from chui_treepm_kick_device / chui_treepm_drift_device.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np
import torch

from . import _lib
from .treepm import ForceAccumulator, TreePmCoordinator


# ---- FLRW timeline (code units: H in inverse code time) -----------------------
@dataclass
class Cosmology:
    hubble0: float
    omega_m: float
    omega_lambda: float

    def hubble(self, a):
        ok = 1.0 - self.omega_m - self.omega_lambda
        return self.hubble0 * np.sqrt(self.omega_m / a**3 + ok / a**2 + self.omega_lambda)

    def _midpoint(self, f, b, e, samples: int = 64):
        """Midpoint rule in a over [b, e] (elementwise over array endpoints);
        the samples are summed pairwise."""
        b = np.asarray(b, dtype=np.float64)
        e = np.asarray(e, dtype=np.float64)
        da = (e - b) / samples
        a = b[..., None] + (np.arange(samples) + 0.5) * da[..., None]
        r = f(a).sum(axis=-1) * da
        return float(r) if r.ndim == 0 else r

    def cosmic_time(self, b, e, samples: int = 64):
        """int dt = int da / (a H), midpoint rule in a."""
        return self._midpoint(lambda a: 1.0 / (a * self.hubble(a)), b, e, samples)

    def drift_like(self, b, e):
        """int dt / a = int da / (a^2 H) (computeComovingDriftFactor)."""
        return self._midpoint(lambda a: 1.0 / (a * a * self.hubble(a)), b, e)

    def position_drift(self, b, e):
        """int dt / a^2 = int da / (a^3 H)."""
        return self._midpoint(lambda a: 1.0 / (a * a * a * self.hubble(a)), b, e)

    def advance_many(self, a: float, dts, samples: int = 64) -> np.ndarray:
        """Scale factors cosmic times dts after a, each bracketed by doubling
        from 1.001 a and bisected 96 times (advanceScaleFactorByCosmicTime),
        all at once. Bisection stops early once no bracket changes any more:
        when mid rounds to lo or hi the bracket is a fixed point, so the
        result equals the 96-step one."""
        dts = np.asarray(dts, dtype=np.float64)
        base = np.full(dts.shape, float(a))
        hi = base * 1.001
        for _ in range(256):
            short = self.cosmic_time(base, hi, samples) < dts
            if not short.any():
                break
            hi = np.where(short, hi * 2.0, hi)
        lo = base.copy()
        for _ in range(96):
            mid = 0.5 * (lo + hi)
            below = self.cosmic_time(base, mid, samples) < dts
            nlo = np.where(below, mid, lo)
            nhi = np.where(below, hi, mid)
            if np.array_equal(nlo, lo) and np.array_equal(nhi, hi):
                break
            lo, hi = nlo, nhi
        return np.where(dts == 0.0, base, 0.5 * (lo + hi))

    def advance(self, a: float, dt: float, samples: int = 64) -> float:
        """Scale factor a cosmic time dt after a (advance_many of one)."""
        return float(self.advance_many(a, [dt], samples)[0])


@dataclass
class BlockPlan:
    """One block: bins, quantum, the half-tick scale-factor grid and the kick
    and drift factors of every tick."""
    max_bin: int
    quantum: float
    scale: list
    bins: object  # int64 bin per particle: numpy array, or a device tensor when planned from one
    opening: list = field(default_factory=list)   # per tick: {bin: factor}
    closing: list = field(default_factory=list)   # per tick: {bin: factor}
    drift: list = field(default_factory=list)     # per tick: a_k int dt/a^2
    drag: list = field(default_factory=list)      # per tick: a_k / a_{k+1}
    pm_open: float = 0.0
    pm_close: float = 0.0


def plan_block(cosmo: Cosmology, a_begin: float, accel_norm, softening: float, max_bin: int,
               eta: float = 0.2, max_span: float | None = None, bins=None) -> BlockPlan:
    """Bins and factors of one block (time_integration.cpp:1356-1600;
    docs/time_integration.md:392-470). accel_norm = |A| per particle at the
    block start, a numpy array or a device tensor (the bins are then computed
    on the device and plan.bins is a device int64 tensor); bins, when given,
    replaces the criterion's assignment."""
    if not 0 <= max_bin <= 12:
        raise ValueError("hierarchical_max_rung must be in [0, 12]")
    a = a_begin
    periods = 1 << max_bin
    if torch.is_tensor(accel_norm):
        dt = eta * torch.sqrt(softening / (accel_norm.double() / a**3))
        dt_min = float(dt.min()) if dt.numel() else math.inf
    else:
        with np.errstate(divide="ignore"):
            dt = eta * np.sqrt(softening / (np.asarray(accel_norm, dtype=np.float64) / a**3))
        dt_min = float(np.min(dt)) if dt.size else math.inf
    span = periods * dt_min
    if max_span is not None:
        span = min(span, max_span)
    if not (math.isfinite(span) and span > 0.0):
        raise ValueError("hierarchical block span is not finite positive (no force or no explicit span)")
    quantum = span / periods
    if bins is None:
        if torch.is_tensor(dt):
            b = torch.floor(torch.log2(dt / quantum))
            bins = torch.clip(torch.nan_to_num(b, nan=0.0, posinf=max_bin, neginf=0.0), 0, max_bin).to(torch.int64)
        else:
            with np.errstate(divide="ignore"):
                b = np.floor(np.log2(dt / quantum))
            bins = np.clip(np.nan_to_num(b, nan=0.0, posinf=max_bin, neginf=0.0), 0, max_bin).astype(np.int64)
    halves = cosmo.advance_many(a_begin, 0.5 * quantum * np.arange(1, 2 * periods + 1))
    scale = [a_begin] + [float(v) for v in halves]
    plan = BlockPlan(max_bin, quantum, scale, bins)

    def kick(begin: int, end: int, a_now: float) -> float:
        return cosmo.drift_like(scale[begin], scale[end]) / a_now

    for off in range(periods):
        a0, a1 = scale[2 * off], scale[2 * off + 2]
        plan.opening.append({bb: kick(2 * off, 2 * off + (1 << bb), a0)
                             for bb in range(max_bin + 1) if off % (1 << bb) == 0})
        plan.drift.append(a0 * cosmo.position_drift(a0, a1))
        plan.drag.append(a0 / a1)
        plan.closing.append({bb: kick(2 * (off + 1) - (1 << bb), 2 * (off + 1), a1)
                             for bb in range(max_bin + 1) if (off + 1) % (1 << bb) == 0})
    plan.pm_open = kick(0, periods, scale[0])
    plan.pm_close = kick(periods, 2 * periods, scale[-1])
    return plan


class HierarchicalKdk:
    """Blocks of the hierarchical KDK for one periodic box (FP64 state on the
    device; forces in options.precision)."""

    def __init__(self, mesh, x, u, mass, options: _lib.Options, cosmo: Cosmology, max_bin: int,
                 eta: float = 0.2, device: str = "cuda:0"):
        self.coord = TreePmCoordinator(mesh)
        self.x = [t.contiguous() for t in x]
        self.u = [t.contiguous() for t in u]
        self.m = mass.contiguous()
        # a private copy: the fine ticks switch short_range_only on it, and the
        # caller's options must stay as given even when a solve raises
        self.options = _lib.Options.from_buffer_copy(options)
        self.cosmo = cosmo
        self.max_bin = max_bin
        self.eta = eta
        self.device = device
        self.n = int(self.m.shape[0])
        self.dtype = torch.float32 if options.precision == _lib.PRECISION_MIXED else torch.float64
        self.box = tuple(options.box_size)
        self.solves = 0
        self.short_range_solves = 0
        # torch's gathers and the library's solves, kicks and drifts share one
        # stream (default priority), so a block needs no host synchronisation
        # between them
        self.stream = torch.cuda.Stream(device=device)
        self.coord.set_stream(self.stream.cuda_stream)

    def close(self):
        self.coord.close()

    # -- device pieces (enqueued on self.stream) ----------------------------------
    def _solve(self, active, refresh: bool):
        """Short-range lanes per source row; PM lanes per active slot when refreshing."""
        na = self.n if active is None else int(active.shape[0])
        acc = tuple(torch.empty(na, dtype=self.dtype, device=self.device) for _ in range(3))
        sr = tuple(torch.zeros(self.n, dtype=self.dtype, device=self.device) for _ in range(3))
        pm = tuple(torch.empty(na, dtype=self.dtype, device=self.device) for _ in range(3)) if refresh else None
        self.options.short_range_only = 0 if refresh else 1
        self.coord.solve_active_set_with_pm_cadence(*self.x, self.m, ForceAccumulator(acc, active, pm, sr),
                                                    self.options, refresh)
        if refresh:
            self.solves += 1
        else:
            self.short_range_solves += 1
        return acc, pm, sr

    def _kick(self, accel, rows, factor: float, drag: float = 1.0):
        """u[rows] = drag u[rows] + factor accel (accel per listed row)."""
        self.coord.kick(*self.u, *accel, factor, drag, rows)

    def block(self, a_begin: float, max_span: float | None = None, bins=None) -> BlockPlan:
        """One block from a_begin; the returned plan's bins are a device tensor."""
        self.stream.wait_stream(torch.cuda.current_stream(self.device))
        with torch.cuda.stream(self.stream):
            plan = self._block(a_begin, max_span, bins)
        self.stream.synchronize()
        return plan

    def _block(self, a_begin, max_span, bins) -> BlockPlan:
        # synchronized Tree+PM solve at the block start (the mandatory bootstrap)
        acc, pm, sr = self._solve(None, True)
        norm = torch.sqrt(acc[0].double() ** 2 + acc[1].double() ** 2 + acc[2].double() ** 2)
        if bins is not None and not torch.is_tensor(bins):
            bins = torch.from_numpy(np.asarray(bins, dtype=np.int64)).to(self.device)
        plan = plan_block(self.cosmo, a_begin, norm, self.options.softening_epsilon, self.max_bin, self.eta, max_span,
                          bins)
        bins = plan.bins
        # rows of each bin, once per block (the bins are fixed within it)
        rows_of = [torch.nonzero(bins == b).reshape(-1).to(torch.int32) for b in range(plan.max_bin + 1)]
        zeros = tuple(torch.zeros_like(t) for t in pm)
        tree = sr
        periods = 1 << plan.max_bin
        all_rows = None
        for off in range(periods):
            # opening half-kicks (tree force of the interval's start) and the PM opening kick
            for b, f in plan.opening[off].items():
                rows = rows_of[b]
                if rows.numel():
                    self._kick(tuple(t.index_select(0, rows.long()) for t in tree), rows, f)
            if off == 0:
                self._kick(pm, all_rows, plan.pm_open)
            # drift of every particle with u at the tick start, then the Hubble drag of u
            self.coord.drift(*self.x, *self.u, plan.drift[off], self.box)
            self._kick(zeros, all_rows, 0.0, plan.drag[off])
            # closing forces and half-kicks of the bins whose interval ends
            closing = list(plan.closing[off].keys())
            last = off + 1 == periods
            if last:
                acc, pm, tree = self._solve(None, True)
            else:
                rows = torch.cat([rows_of[b] for b in closing])
                if rows.numel():
                    _, _, sr = self._solve(rows, False)
                    tree = tuple(torch.where(self._member(bins, closing), s, t) for s, t in zip(sr, tree))
            for b, f in plan.closing[off].items():
                rows = rows_of[b]
                if rows.numel():
                    self._kick(tuple(t.index_select(0, rows.long()) for t in tree), rows, f)
            if last:
                self._kick(pm, all_rows, plan.pm_close)
        return plan

    @staticmethod
    def _member(bins_dev, wanted):
        mask = torch.zeros_like(bins_dev, dtype=torch.bool)
        for b in wanted:
            mask |= bins_dev == b
        return mask


__all__ = ["BlockPlan", "Cosmology", "HierarchicalKdk", "plan_block"]
