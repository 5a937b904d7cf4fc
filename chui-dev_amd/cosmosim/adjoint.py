"""Differentiable TreePM + KDK forward and its adjoint behind torch.autograd.

The reference has no autograd or adjoint code; this module is the
differentiation layer the MI355X build adds on top of the force evaluation
(tree_pm_coupling.hpp:563-584) and the FP64 KDK pieces
(time_integration.cpp:242-266, 448-450, 3530-3556).

Forward model (all FP64 on the device, caller order of particles):
  x_0 = wrap(q + psi),  v_0 = a_0 H(a_0) f(a_0) psi          (Zel'dovich start)
  step n (a_n -> a_{n+1}):  v_h     = d1_n v_n + K1_n g(x_n)
                            x_{n+1} = wrap(x_n + D_n v_h)
                            v_{n+1} = d2_n v_h + K2_n g(x_{n+1})
with g the total TreePM acceleration (chui_treepm_solve_device, every source
active). The adjoint runs the steps backwards:
  lam_g   = K2_n lam_v + c_{n+1}       (c_{n+1} = K1_{n+1} lam_vh of step n+1)
  lam_x  += J(x_{n+1})^T lam_g         (chui_treepm_force_vjp_device)
  lam_vh  = d2_n lam_v + D_n lam_x
  lam_v   = d1_n lam_vh,  c_n = K1_n lam_vh
and finally lam_x += J(x_0)^T c_0, d loss / d psi = lam_x + a_0 H f lam_v. The
wrap has unit derivative. Positions are checkpointed every `checkpoint_every`
steps; the backward pass recomputes the positions inside each segment.

The forward is the strict solve with quadrupole moments (multipole_order 2,
the benchmarked forward). The force VJP takes the walk's acceptance decisions:
pairs contribute their pairwise Jacobian, and each accepted node its
expansion's Jacobian through the target position and the node's centre of
mass and quadrupole (tree_pm_coupling.cpp:743-827). On the near-uniform
lattice workloads no node is accepted (2h/r < 0.08 needs cells far smaller
than the particle spacing), so the short-range force there is the exact pair
sum within r_cut (tree_pm_acceptance.hpp:62-66).
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np
import torch

from . import _lib, ics
from .treepm import ForceAccumulator, TreePmCoordinator, kdk_factors


@dataclass
class KdkProblem:
    """One periodic box, its lattice and its KDK schedule."""

    n_side: int
    box: float
    mesh: int
    scale_factors: list
    softening: float
    loss_mesh: int
    checkpoint_every: int = 2
    omega_m: float = ics.OMEGA_M
    device: str = "cuda:0"
    options: object = None
    coord: TreePmCoordinator | None = None
    lattice: torch.Tensor | None = None
    mass: torch.Tensor | None = None
    factors: list = field(default_factory=list)
    vel_factor: float = 0.0
    force_evaluations: int = 0
    vjp_evaluations: int = 0
    # When set, every force evaluation also returns the walk's diagnostics and
    # appends (direct pair count, pair-set checksum, accepted nodes) to
    # pair_sets: two trajectories whose lists agree summed the same pairs
    # inside r_cut and accepted as many nodes.
    record_pair_sets: bool = False
    pair_sets: list = field(default_factory=list)
    # Set while a graph is recorded: the library then enqueues on torch's
    # stream (the caller set it with coord.set_stream), so no call waits.
    graph_mode: bool = False

    def __post_init__(self):
        n = self.n_side**3
        spacing = self.box / self.n_side
        idx = torch.arange(self.n_side, dtype=torch.float64, device=self.device)
        gx, gy, gz = torch.meshgrid(idx, idx, idx, indexing="ij")
        self.lattice = torch.stack([(g.reshape(-1) + 0.5) * spacing for g in (gx, gy, gz)]).contiguous()
        rho_crit = 3.0 * ics.H0_CODE**2 / (8.0 * np.pi * ics.G_CODE)
        self.mass = torch.full((n,), rho_crit * self.omega_m * self.box**3 / n, dtype=torch.float64,
                               device=self.device)
        self.options = _lib.default_options(
            box_size=self.box, gravitational_constant=ics.G_CODE, opening_theta=0.5, asmth_cells=1.25,
            rcut_cells=4.5 * 1.25, softening_epsilon=self.softening, multipole_order=2,
            precision=_lib.PRECISION_FP64)
        self.coord = TreePmCoordinator(self.mesh)
        self.factors = [kdk_factors(a0, a1, ics.H0_CODE, self.omega_m, 1.0 - self.omega_m)
                        for a0, a1 in zip(self.scale_factors[:-1], self.scale_factors[1:])]
        a0 = self.scale_factors[0]
        self.vel_factor = a0 * ics.hubble(a0, self.omega_m) * ics.growth_rate(a0, self.omega_m)

    @property
    def steps(self) -> int:
        return len(self.scale_factors) - 1

    def close(self):
        if self.coord is not None:
            self.coord.close()
            self.coord = None

    # -- device pieces ----------------------------------------------------------
    # The library runs on its own stream: each call waits for torch's queued
    # work on its inputs and is waited for before torch reads its outputs.
    def _call(self, fn, *args):
        if self.graph_mode:
            fn(*args)
            return
        torch.cuda.current_stream(self.device).synchronize()
        fn(*args)
        self.coord.synchronize()

    def wrap(self, x: torch.Tensor) -> torch.Tensor:
        return torch.remainder(x, self.box)

    def forces(self, x: torch.Tensor) -> torch.Tensor:
        acc = torch.empty_like(x)
        if self.record_pair_sets:
            torch.cuda.current_stream(self.device).synchronize()
            res = self.coord.solve_active_set(x[0], x[1], x[2], self.mass, ForceAccumulator((acc[0], acc[1], acc[2])),
                                              self.options, want_diagnostics=True)
            d = res.diagnostics
            self.pair_sets.append((d["direct_pair_evaluations"], d["pair_set_checksum"],
                                   d["accepted_internal_multipoles"]))
        else:
            self._call(self.coord.solve_active_set, x[0], x[1], x[2], self.mass,
                       ForceAccumulator((acc[0], acc[1], acc[2])), self.options)
        self.force_evaluations += 1
        return acc

    def force_vjp(self, x: torch.Tensor, lam: torch.Tensor) -> torch.Tensor:
        out = torch.empty_like(x)
        lam = lam.contiguous()
        self._call(self.coord.force_vjp, x[0], x[1], x[2], self.mass, (lam[0], lam[1], lam[2]),
                   (out[0], out[1], out[2]), self.options)
        self.vjp_evaluations += 1
        return out

    def cic_density(self, x: torch.Tensor) -> torch.Tensor:
        m = self.loss_mesh
        delta = torch.empty(m, m, m, dtype=torch.float64, device=x.device)
        self._call(self.coord.cic_density, x[0], x[1], x[2], self.box, m, delta)
        return delta

    def cic_density_vjp(self, x: torch.Tensor, grad_delta: torch.Tensor) -> torch.Tensor:
        out = torch.empty_like(x)
        self._call(self.coord.cic_density_vjp, x[0], x[1], x[2], self.box, self.loss_mesh,
                   grad_delta.contiguous(), (out[0], out[1], out[2]))
        return out

    def kdk_step(self, n: int, x, v, g, positions_only: bool = False):
        """One KDK step; with positions_only the drifted positions alone (no
        force evaluation, v and g returned as None)."""
        f = self.factors[n]
        vh = f["first_hubble_drag"] * v + f["first_kick_factor"] * g
        x = self.wrap(x + f["drift_factor"] * vh)
        if positions_only:
            return x, None, None
        g = self.forces(x)
        v = f["second_hubble_drag"] * vh + f["second_kick_factor"] * g
        return x, v, g


def trajectory_forward(psi: torch.Tensor, prob: KdkProblem):
    """The KDK forward from the displacements psi (3, N) FP64: final positions
    and the checkpoints {step: (x, v, g)} every prob.checkpoint_every steps."""
    x = prob.wrap(prob.lattice + psi)
    v = prob.vel_factor * psi
    g = prob.forces(x)
    checkpoints = {0: (x, v, g)}
    for n in range(prob.steps):
        x, v, g = prob.kdk_step(n, x, v, g)
        if (n + 1) % prob.checkpoint_every == 0 and n + 1 < prob.steps:
            checkpoints[n + 1] = (x, v, g)
    return x, checkpoints


def trajectory_backward(prob: KdkProblem, checkpoints: dict, final: torch.Tensor,
                        lam_final: torch.Tensor) -> torch.Tensor:
    """The adjoint sweep: d loss / d psi from lam_final = d loss / d x_final."""
    S = prob.steps
    lam_x = lam_final.contiguous().clone()
    lam_v = torch.zeros_like(lam_x)
    carry = None
    starts = sorted(checkpoints)
    for seg_start in reversed(starts):
        seg_end = min(seg_start + prob.checkpoint_every, S)
        # positions x_seg_start .. x_seg_end (recomputed inside the segment);
        # the reverse sweep needs positions only, so the last recomputed
        # step drifts without evaluating the force at its new positions
        x, v, g = checkpoints[seg_start]
        xs = [x]
        for n in range(seg_start, seg_end - 1):
            x, v, g = prob.kdk_step(n, x, v, g, positions_only=n == seg_end - 2)
            xs.append(x)
        xs.append(final if seg_end == S else checkpoints[seg_end][0])
        for n in range(seg_end - 1, seg_start - 1, -1):
            f = prob.factors[n]
            lam_g = f["second_kick_factor"] * lam_v
            if carry is not None:
                lam_g = lam_g + carry
            lam_x = lam_x + prob.force_vjp(xs[n + 1 - seg_start], lam_g)
            lam_vh = f["second_hubble_drag"] * lam_v + f["drift_factor"] * lam_x
            lam_v = f["first_hubble_drag"] * lam_vh
            carry = f["first_kick_factor"] * lam_vh
    lam_x = lam_x + prob.force_vjp(checkpoints[0][0], carry)
    return lam_x + prob.vel_factor * lam_v


class KdkTrajectory(torch.autograd.Function):
    """psi (3, N) FP64 initial displacements -> final positions (3, N)."""

    @staticmethod
    def forward(ctx, psi: torch.Tensor, prob: KdkProblem):
        x, checkpoints = trajectory_forward(psi, prob)
        ctx.prob = prob
        ctx.checkpoints = checkpoints
        ctx.final = x
        return x.clone()

    @staticmethod
    def backward(ctx, lam_final: torch.Tensor):
        return trajectory_backward(ctx.prob, ctx.checkpoints, ctx.final, lam_final), None


class CicDensity(torch.autograd.Function):
    """positions (3, N) -> CIC density contrast on a loss_mesh^3 grid."""

    @staticmethod
    def forward(ctx, x: torch.Tensor, prob: KdkProblem):
        ctx.prob = prob
        ctx.save_for_backward(x)
        return prob.cic_density(x)

    @staticmethod
    def backward(ctx, grad_delta: torch.Tensor):
        (x,) = ctx.saved_tensors
        return ctx.prob.cic_density_vjp(x, grad_delta), None


def density_loss(psi: torch.Tensor, target: torch.Tensor, prob: KdkProblem) -> torch.Tensor:
    """sum over cells of (delta_final - target)^2; psi may be FP32 (promoted)."""
    x = KdkTrajectory.apply(psi.to(torch.float64), prob)
    delta = CicDensity.apply(x, prob)
    return ((delta - target) ** 2).sum()


def gradient_sweeps(psi: torch.Tensor, target: torch.Tensor, prob: KdkProblem):
    """The forward and adjoint sweeps of density_loss without its final sum:
    (delta_final - target, d loss / d psi). Element-wise torch work and library
    calls only, so the sweeps can be recorded into a graph."""
    with torch.no_grad():
        x, checkpoints = trajectory_forward(psi.to(torch.float64), prob)
        diff = prob.cic_density(x) - target
        lam_final = prob.cic_density_vjp(x, 2.0 * diff)
        grad = trajectory_backward(prob, checkpoints, x, lam_final)
    return diff, grad


def loss_and_gradient(psi: torch.Tensor, target: torch.Tensor, prob: KdkProblem):
    """density_loss and d loss / d psi by the same forward and adjoint sweeps
    as the autograd Functions, run on the calling thread and its current
    stream without the autograd engine (whose device thread would serialise
    the backward passes of boxes differentiated from several host threads)."""
    diff, grad = gradient_sweeps(psi, target, prob)
    with torch.no_grad():
        loss = (diff * diff).sum()
    return loss, grad


def central_differences(prob: KdkProblem, psi: torch.Tensor, target: torch.Tensor, grad: torch.Tensor,
                        directions: int, step: float, seed: int = ics.BASE_SEED) -> list:
    """Central differences of density_loss along `directions` random unit
    directions u of the (3, N) displacement field, (L(psi + step u) -
    L(psi - step u)) / (2 step), against grad . u. A direction is flagged
    crosses_cutoff when the two trajectories did not sum the same pairs inside
    r_cut in every force evaluation (pair counts and pair-set checksums of the
    FP64 walk): the loss jumps there, and the difference quotient measures the
    jump, not the derivative."""
    rng = np.random.default_rng(seed)
    rows = []
    saved = prob.record_pair_sets
    prob.record_pair_sets = True
    try:
        for _ in range(directions):
            d = rng.standard_normal(tuple(psi.shape))
            d /= np.linalg.norm(d)
            u = torch.tensor(d, dtype=torch.float64, device=psi.device)
            sets = []
            values = []
            for sign in (1.0, -1.0):
                prob.pair_sets = []
                with torch.no_grad():
                    values.append(float(density_loss(psi + sign * step * u, target, prob)))
                sets.append(list(prob.pair_sets))
            fd = (values[0] - values[1]) / (2.0 * step)
            an = float((grad * u).sum())
            rows.append({"finite_difference": fd, "analytic": an,
                         "relative_error": abs(fd - an) / max(abs(an), 1e-300),
                         "crosses_cutoff": sets[0] != sets[1]})
    finally:
        prob.record_pair_sets = saved
        prob.pair_sets = []
    return rows


def adjoint_problem(n_side: int = 64, steps: int = 10, a_begin: float = 0.05, a_end: float = 0.5,
                    checkpoint_every: int = 2, rms_spacing: float = 0.1, device: str = "cuda:0",
                    seed: int = ics.BASE_SEED + 3):
    """The differentiable workload: n_side^3 particles in a box of 50 (n_side/64)
    Mpc/h, mesh 2 n_side, KDK steps uniform in a, a seeded displacement
    (P(k) ~ k^-2.5, RMS rms_spacing mean spacings at a_begin: 0.1 grows to about
    one spacing by a = 0.5) and a target density field from a second seeded
    displacement of ten times that RMS. Returns (problem, psi (3, N) FP32 on the
    device, target (loss_mesh^3) FP64)."""
    box = 50.0 * n_side / 64
    rng = np.random.default_rng(seed)
    disp = ics.zeldovich_displacement(n_side, box, -2.5, rms_spacing, rng)
    tdisp = ics.zeldovich_displacement(n_side, box, -2.5, 10.0 * rms_spacing, np.random.default_rng(seed + 1000))
    prob = KdkProblem(n_side=n_side, box=box, mesh=2 * n_side,
                      scale_factors=list(np.linspace(a_begin, a_end, steps + 1)),
                      softening=box / n_side / 30.0, loss_mesh=n_side, checkpoint_every=checkpoint_every,
                      device=device)
    psi = torch.tensor(disp, dtype=torch.float32, device=device)
    # target: the CIC density contrast of the lattice displaced by the second field
    tx = prob.wrap(prob.lattice + torch.tensor(tdisp, dtype=torch.float64, device=device))
    return prob, psi, prob.cic_density(tx)
