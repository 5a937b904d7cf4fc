"""Ensembles of independent periodic boxes: workload law, batched KDK forward,
per-box gradients and the autotuner of leaf size and launch shapes.

The reference evolves one box per run; ensembles have no counterpart there.
Each box follows the single-box path (TreePmCoordinator::solveActiveSet,
tree_pm_coupling.hpp:563-584, and the KDK pieces of time_integration.cpp:
448-450, 3530-3556). The batched forward issues every box's force evaluation
through chui_treepm_ensemble_solve_device (host worker threads, one stream per
box); kicks and drifts run on each box's stream between the batched solves.
Gradients come from the differentiable forward of cosmosim.adjoint (FP64, one
box at a time).
"""
from __future__ import annotations

import ctypes as C
import time
from dataclasses import dataclass

import numpy as np
import torch

from . import _lib, adjoint, ics
from .treepm import Graph, TreePmCoordinator, kdk_factors


@dataclass
class BoxSpec:
    n_side: int
    box: float
    mesh: int
    spectral_index: float
    a_begin: float
    a_end: float
    seed: int

    @property
    def n(self) -> int:
        return self.n_side**3


def ensemble_specs(counts=((32, 16), (64, 16), (96, 12), (128, 4)), seed: int = ics.BASE_SEED + 4) -> list:
    """The ensemble law: for each (n_side, count), count boxes of n_side^3
    particles in 50 n_side / 64 Mpc/h (the density of the 64^3 / 50 Mpc/h
    adjoint workload), mesh 2 n_side, spectral index uniform in [-2.8, -2.0],
    a_begin uniform in [0.05, 0.1] and a_end = a_begin + U[0.25, 0.45]; every
    box has its own seed."""
    rng = np.random.default_rng(seed)
    specs = []
    for n_side, count in counts:
        for _ in range(count):
            a0 = float(rng.uniform(0.05, 0.1))
            specs.append(BoxSpec(n_side=n_side, box=50.0 * n_side / 64, mesh=2 * n_side,
                                 spectral_index=float(rng.uniform(-2.8, -2.0)), a_begin=a0,
                                 a_end=a0 + float(rng.uniform(0.25, 0.45)), seed=int(rng.integers(1 << 31))))
    return specs


def box_initial_conditions(spec: BoxSpec, rms_spacing: float = 0.1):
    """Lattice + Zel'dovich displacement psi (3, N) of the box's own field."""
    rng = np.random.default_rng(spec.seed)
    psi = ics.zeldovich_displacement(spec.n_side, spec.box, spec.spectral_index, rms_spacing, rng)
    spacing = spec.box / spec.n_side
    idx = (np.arange(spec.n_side, dtype=np.float64) + 0.5) * spacing
    gx, gy, gz = np.meshgrid(idx, idx, idx, indexing="ij")
    lattice = np.stack([gx.reshape(-1), gy.reshape(-1), gz.reshape(-1)])
    return lattice, psi


class Ensemble:
    """One context per box (chui_treepm_ensemble_create)."""

    def __init__(self, specs, device: int = 0, workers: int = 8):
        self._lib = _lib.load()
        self.specs = list(specs)
        meshes = (C.c_int32 * (3 * len(self.specs)))(*[s.mesh for s in self.specs for _ in range(3)])
        handle = C.c_void_p()
        _lib.check(self._lib.chui_treepm_ensemble_create(len(self.specs), meshes, int(device), int(workers),
                                                         C.byref(handle)))
        self._h = handle
        self.views = []
        for b, s in enumerate(self.specs):
            ctx = C.c_void_p()
            _lib.check(self._lib.chui_treepm_ensemble_box(self._h, b, C.byref(ctx)))
            self.views.append(TreePmCoordinator.borrowed(ctx, (s.mesh,) * 3))

    def close(self):
        if getattr(self, "_h", None):
            for v in self.views:
                v.close()
            _lib.check(self._lib.chui_treepm_ensemble_destroy(self._h))
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def solve(self, positions, masses, accels, options, refresh: bool = True):
        """positions[b] = (x, y, z), accels[b] = (ax, ay, az): device tensors;
        options[b]: cosmosim._lib.Options (kept alive by the caller)."""
        boxes = (_lib.Box * len(self.specs))()
        for b in range(len(self.specs)):
            x, y, z = positions[b]
            ax, ay, az = accels[b]
            boxes[b] = _lib.Box(x.data_ptr(), y.data_ptr(), z.data_ptr(), masses[b].data_ptr(), int(x.shape[0]),
                                ax.data_ptr(), ay.data_ptr(), az.data_ptr(), C.pointer(options[b]))
        _lib.check(self._lib.chui_treepm_ensemble_solve_device(self._h, boxes, len(self.specs), 1 if refresh else 0))

    def set_capture_safe(self, enable: bool = True):
        """Capture-safe solves on every box (chui_treepm_ensemble_set_capture_safe)."""
        _lib.check(self._lib.chui_treepm_ensemble_set_capture_safe(self._h, 1 if enable else 0))

    def graph_begin(self):
        """Record every box's stream into one HIP graph (box 0's stream records,
        the others join it as branches)."""
        _lib.check(self._lib.chui_treepm_ensemble_graph_begin(self._h))

    def graph_end(self) -> Graph:
        out = C.c_void_p()
        _lib.check(self._lib.chui_treepm_ensemble_graph_end(self._h, C.byref(out)))
        return Graph(out)


def box_options(spec: BoxSpec, precision: int, leaf: int = 16):
    spacing = spec.box / spec.n_side
    return _lib.default_options(box_size=spec.box, gravitational_constant=ics.G_CODE, opening_theta=0.5,
                                asmth_cells=1.25, rcut_cells=4.5 * 1.25, softening_epsilon=spacing / 30.0,
                                max_leaf_size=leaf, precision=precision)


def _median(v):
    return float(np.median(np.asarray(v, dtype=np.float64)))


def autotune(spec: BoxSpec, rounds: int = 5, repeats: int = 3, device: str = "cuda:0",
             leaves=(8, 12, 16, 24, 32), fft_lines=(0, 8, 4), walk_waves=(4, 8, 2)):
    """Launch parameters of one box of this size (its own initial conditions,
    mixed-precision force evaluations): leaf size, lines per block of the PM
    line passes and waves per block of the timed walk, tuned one after the
    other with the others at their current choice. Each candidate is timed in
    `rounds` rounds of `repeats` solves, the candidates interleaved within a
    round so that a drift of the clock hits all of them alike. A candidate
    replaces the current choice only when its median beats the current
    choice's median by more than the noise: twice the largest relative spread
    (median absolute deviation over median) seen among the candidates, and at
    least 2%. The first entry of each list is the default. Returns
    {"leaf", "fft_lines", "walk_waves", "steps": [per knob: medians, noise,
    choice]}."""
    lattice, psi = box_initial_conditions(spec)
    pos = np.mod(lattice + psi, spec.box)
    pos[pos >= spec.box] = 0.0
    dev = [torch.from_numpy(np.ascontiguousarray(p)).to(device) for p in pos]
    rho_crit = 3.0 * ics.H0_CODE**2 / (8.0 * np.pi * ics.G_CODE)
    m = torch.full((spec.n,), rho_crit * ics.OMEGA_M * spec.box**3 / spec.n, dtype=torch.float64, device=device)
    acc = tuple(torch.empty(spec.n, dtype=torch.float32, device=device) for _ in range(3))
    torch.cuda.synchronize(device)  # the library runs on its own stream
    from .treepm import ForceAccumulator

    choice = {"leaf": 16 if 16 in leaves else leaves[0], "fft_lines": fft_lines[0], "walk_waves": walk_waves[0]}
    steps = []
    with TreePmCoordinator(spec.mesh) as coord:
        def sample(cfg):
            coord.set_launch_shape(cfg["fft_lines"], cfg["walk_waves"])
            o = box_options(spec, _lib.PRECISION_MIXED, cfg["leaf"])
            coord.solve_active_set(*dev, m, ForceAccumulator(acc), o)
            coord.synchronize()
            t = time.perf_counter()
            for _ in range(repeats):
                coord.solve_active_set(*dev, m, ForceAccumulator(acc), o)
            coord.synchronize()
            return 1e3 * (time.perf_counter() - t) / repeats

        for knob, cands in (("leaf", leaves), ("fft_lines", fft_lines), ("walk_waves", walk_waves)):
            cands = list(dict.fromkeys([choice[knob]] + list(cands)))  # the current choice first
            times = {c: [] for c in cands}
            for _ in range(rounds):
                for c in cands:
                    times[c].append(sample({**choice, knob: c}))
            med = {c: _median(v) for c, v in times.items()}
            spread = max(_median(np.abs(np.asarray(v) - med[c])) / med[c] for c, v in times.items())
            noise = max(0.02, 2.0 * spread)
            best = min(med, key=med.get)
            current = choice[knob]
            if med[best] < med[current] * (1.0 - noise):
                choice[knob] = best
            steps.append({"knob": knob, "ms_median": {str(c): round(v, 4) for c, v in med.items()},
                          "noise": round(noise, 4), "choice": choice[knob]})
        coord.set_launch_shape(0, 0)
    return {**choice, "steps": steps}


class EnsembleForward:
    """Batched KDK forward of every box (steps uniform in a from a_begin to a_end)."""

    def __init__(self, ens: Ensemble, steps: int, precision: int, leaf_by_n_side=None, device: str = "cuda:0",
                 shape_by_n_side=None):
        """leaf_by_n_side: {n_side: leaf size}; shape_by_n_side: {n_side:
        (fft lines per block, walk waves per block)} (autotune's choices)."""
        self.ens = ens
        self.steps = steps
        self.precision = precision
        self.device = device
        leaf_by_n_side = leaf_by_n_side or {}
        self.options = [box_options(s, precision, leaf_by_n_side.get(s.n_side, 16)) for s in ens.specs]
        for b, s in enumerate(ens.specs):
            if shape_by_n_side and s.n_side in shape_by_n_side:
                ens.views[b].set_launch_shape(*shape_by_n_side[s.n_side])
        self.factors = []
        for s in ens.specs:
            a = np.linspace(s.a_begin, s.a_end, steps + 1)
            self.factors.append([kdk_factors(a0, a1, ics.H0_CODE, ics.OMEGA_M, 1.0 - ics.OMEGA_M)
                                 for a0, a1 in zip(a[:-1], a[1:])])
        adt = torch.float32 if precision == _lib.PRECISION_MIXED else torch.float64
        self.x, self.v, self.m, self.g = [], [], [], []
        rho_crit = 3.0 * ics.H0_CODE**2 / (8.0 * np.pi * ics.G_CODE)
        for s in ens.specs:
            lattice, psi = box_initial_conditions(s)
            pos = np.mod(lattice + psi, s.box)
            pos[pos >= s.box] = 0.0
            vf = s.a_begin * ics.hubble(s.a_begin) * ics.growth_rate(s.a_begin)
            self.x.append(tuple(torch.from_numpy(np.ascontiguousarray(p)).to(device) for p in pos))
            self.v.append(tuple(torch.from_numpy(np.ascontiguousarray(vf * p)).to(device) for p in psi))
            self.m.append(torch.full((s.n,), rho_crit * ics.OMEGA_M * s.box**3 / s.n, dtype=torch.float64,
                                     device=device))
            self.g.append(tuple(torch.empty(s.n, dtype=adt, device=device) for _ in range(3)))
        torch.cuda.synchronize(device)

    @property
    def particles(self) -> int:
        return sum(s.n for s in self.ens.specs)

    def issue(self):
        """Initial forces and every KDK step, issued on the boxes' streams
        without waiting (recordable into a graph)."""
        ens = self.ens
        ens.solve(self.x, self.m, self.g, self.options)
        for k in range(self.steps):
            for b, s in enumerate(ens.specs):
                f = self.factors[b][k]
                view = ens.views[b]
                view.kick(*self.v[b], *self.g[b], f["first_kick_factor"], f["first_hubble_drag"])
                view.drift(*self.x[b], *self.v[b], f["drift_factor"], (s.box,) * 3)
            ens.solve(self.x, self.m, self.g, self.options)
            for b in range(len(ens.specs)):
                f = self.factors[b][k]
                ens.views[b].kick(*self.v[b], *self.g[b], f["second_kick_factor"], f["second_hubble_drag"])

    def run(self):
        """Initial forces and every KDK step; returns when all boxes are done."""
        self.issue()
        self.synchronize()

    def synchronize(self):
        for v in self.ens.views:
            v.synchronize()

    def state(self):
        """Copies of the positions and velocities of every box."""
        return [tuple(t.clone() for t in self.x[b]) + tuple(t.clone() for t in self.v[b])
                for b in range(len(self.ens.specs))]

    def set_state(self, state):
        """Positions and velocities back to a state() copy (in place, so a
        recorded graph sees them)."""
        for b, st in enumerate(state):
            for dst, src in zip(self.x[b] + self.v[b], st):
                dst.copy_(src)
        torch.cuda.synchronize(self.device)

    def capture(self) -> Graph:
        """The whole forward (initial forces and every KDK step of every box)
        as one HIP graph. One eager forward first sizes every box's context
        for capture-safe solves (the state is restored afterwards); each
        launch of the graph then advances the boxes from their current state
        by the same steps, and equals run() bit for bit."""
        start = self.state()
        self.ens.set_capture_safe(True)
        self.run()
        self.set_state(start)
        self.ens.graph_begin()
        try:
            self.issue()
        finally:
            graph = self.ens.graph_end()
        return graph


def box_problem(spec: BoxSpec, steps: int, device: str = "cuda:0", checkpoint_every: int = 2):
    """The differentiable problem of one box, its initial displacements psi
    (3, N) and its loss target (the CIC contrast of a second seeded
    displacement of ten times the RMS)."""
    lattice, psi = box_initial_conditions(spec)
    del lattice
    prob = adjoint.KdkProblem(n_side=spec.n_side, box=spec.box, mesh=spec.mesh,
                              scale_factors=list(np.linspace(spec.a_begin, spec.a_end, steps + 1)),
                              softening=spec.box / spec.n_side / 30.0, loss_mesh=spec.n_side,
                              checkpoint_every=checkpoint_every, device=device)
    tdisp = ics.zeldovich_displacement(spec.n_side, spec.box, spec.spectral_index, 1.0,
                                       np.random.default_rng(spec.seed + 1))
    target = prob.cic_density(prob.wrap(prob.lattice + torch.tensor(tdisp, dtype=torch.float64, device=device)))
    return prob, torch.tensor(psi, dtype=torch.float64, device=device), target


class GradientGraph:
    """d loss / d psi of several boxes recorded as one HIP graph with
    torch.cuda.graph: each box's forward and adjoint sweeps (adjoint.
    gradient_sweeps: the FP64 solves and VJPs of its own capture-safe context,
    the CIC density and its transpose, the element-wise torch work of the KDK
    steps) form a branch on a stream of its own, forked from and joined into
    the recording stream. Two eager gradients first prepare every context
    (the first solve sizes it, the first VJP learns its accepted-list
    capacity). replay() copies new displacements into the graph's inputs,
    launches it once, checks every context's fault word and returns the
    losses (summed outside the graph) and the gradients."""

    def __init__(self, specs, steps: int, device: str = "cuda:0", checkpoint_every: int = 2):
        self.device = device
        made = [box_problem(s, steps, device, checkpoint_every) for s in specs]
        self.problems = [m[0] for m in made]
        self.psi = [m[1] for m in made]
        self.targets = [m[2] for m in made]
        for prob in self.problems:
            prob.coord.set_capture_safe(True)
        for _ in range(2):
            for prob, psi, target in zip(self.problems, self.psi, self.targets):
                adjoint.loss_and_gradient(psi, target, prob)
        torch.cuda.synchronize(device)
        self.main = torch.cuda.Stream(device=device)
        self.streams = [torch.cuda.Stream(device=device) for _ in specs]
        for prob, st in zip(self.problems, self.streams):
            prob.coord.set_stream(st.cuda_stream)
            prob.graph_mode = True
        self.graph = torch.cuda.CUDAGraph()
        self.diff, self.grad = [None] * len(specs), [None] * len(specs)
        try:
            with torch.cuda.graph(self.graph, stream=self.main):
                for b, (prob, st) in enumerate(zip(self.problems, self.streams)):
                    st.wait_stream(self.main)
                    with torch.cuda.stream(st):
                        self.diff[b], self.grad[b] = adjoint.gradient_sweeps(self.psi[b], self.targets[b], prob)
                for st in self.streams:
                    self.main.wait_stream(st)
        finally:
            for prob in self.problems:
                prob.graph_mode = False

    def replay(self, psi=None):
        """psi: optional list of (3, N) displacements copied into the inputs.
        Returns (losses, gradients); the gradients are the graph's outputs,
        overwritten by the next replay."""
        if psi is not None:
            for dst, src in zip(self.psi, psi):
                dst.copy_(src)
        torch.cuda.synchronize(self.device)
        self.graph.replay()
        torch.cuda.synchronize(self.device)
        for prob in self.problems:
            prob.coord.fault_check()
        losses = [float((d * d).sum()) for d in self.diff]
        return losses, self.grad

    def close(self):
        self.graph = None
        for prob in self.problems:
            prob.close()
        self.problems = []


def box_gradient(spec: BoxSpec, steps: int, device: str = "cuda:0", checkpoint_every: int = 2,
                 explicit: bool = False):
    """d loss / d psi of one box: loss = sum of squared differences between the
    final CIC density contrast (n_side^3 grid) and the CIC contrast of a second
    seeded displacement of ten times the RMS. FP64 forward and adjoint, through
    torch.autograd or (explicit) the same sweeps on the calling thread.
    Returns (loss, gradient (3, N) on the device, force evaluations, VJPs)."""
    lattice, psi = box_initial_conditions(spec)
    del lattice
    prob = adjoint.KdkProblem(n_side=spec.n_side, box=spec.box, mesh=spec.mesh,
                              scale_factors=list(np.linspace(spec.a_begin, spec.a_end, steps + 1)),
                              softening=spec.box / spec.n_side / 30.0, loss_mesh=spec.n_side,
                              checkpoint_every=checkpoint_every, device=device)
    try:
        tdisp = ics.zeldovich_displacement(spec.n_side, spec.box, spec.spectral_index, 1.0,
                                           np.random.default_rng(spec.seed + 1))
        target = prob.cic_density(prob.wrap(prob.lattice + torch.tensor(tdisp, dtype=torch.float64, device=device)))
        if explicit:
            p = torch.tensor(psi, dtype=torch.float64, device=device)
            loss, grad = adjoint.loss_and_gradient(p, target, prob)
        else:
            p = torch.tensor(psi, dtype=torch.float64, device=device).requires_grad_(True)
            loss = adjoint.density_loss(p, target, prob)
            (grad,) = torch.autograd.grad(loss, p)
        torch.cuda.current_stream(device).synchronize()
        return float(loss.detach()), grad.detach(), prob.force_evaluations, prob.vjp_evaluations
    finally:
        prob.close()


def ensemble_gradients(specs, steps: int, workers: int = 8, device: str = "cuda:0", checkpoint_every: int = 2):
    """box_gradient of every box, the boxes handed to `workers` host threads
    largest first; each thread differentiates its boxes on a torch stream of
    its own (default priority) with the explicit adjoint sweeps, so the boxes'
    force evaluations, VJPs and torch work overlap on the device. Returns the
    box_gradient results in the order of specs."""
    from concurrent.futures import ThreadPoolExecutor

    order = sorted(range(len(specs)), key=lambda b: specs[b].n, reverse=True)
    results = [None] * len(specs)

    def work(b):
        stream = torch.cuda.Stream(device=device)
        with torch.cuda.stream(stream):
            results[b] = box_gradient(specs[b], steps, device=device, checkpoint_every=checkpoint_every,
                                      explicit=True)

    with ThreadPoolExecutor(max_workers=max(1, int(workers))) as pool:
        for f in [pool.submit(work, b) for b in order]:
            f.result()
    torch.cuda.synchronize(device)
    return results


__all__ = ["BoxSpec", "Ensemble", "EnsembleForward", "autotune", "box_gradient", "box_initial_conditions",
           "box_options", "box_problem", "ensemble_gradients", "ensemble_specs", "GradientGraph"]
