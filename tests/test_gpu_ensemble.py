"""GPU checks of the ensemble entry points (no reference counterpart).

A batched solve over boxes of different sizes and meshes (including a 48^3
mesh, not a power of two) gives each box the accelerations of a separate
single-box solve (FP64: 1e-12 relative L2, the PM deposit's atomics differ by
round-off; mixed: 1e-5), a batched KDK forward equals the same steps run box by
box, errors name the box, and a box gradient matches central differences.
"""
from __future__ import annotations

import numpy as np
import pytest

from helpers import rel_l2

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
if not torch.cuda.is_available():  # pragma: no cover - GPU-only module
    pytest.skip("no HIP device", allow_module_level=True)

from cosmosim import ForceAccumulator, TreePmCoordinator  # noqa: E402
from cosmosim import _lib as L  # noqa: E402
from cosmosim import ensemble as E  # noqa: E402

SPECS = E.ensemble_specs(counts=((16, 2), (24, 1), (32, 2)), seed=77)


@pytest.mark.parametrize("precision,tol", [(L.PRECISION_FP64, 1e-12), (L.PRECISION_MIXED, 1e-5)])
def test_batched_solve_equals_separate_solves(precision, tol):
    fwd_opts = [E.box_options(s, precision) for s in SPECS]
    with E.Ensemble(SPECS, workers=3) as ens:
        state = E.EnsembleForward(ens, steps=1, precision=precision)
        ens.solve(state.x, state.m, state.g, fwd_opts)
        batched = [np.stack([a.double().cpu().numpy() for a in g]) for g in state.g]
        for b, s in enumerate(SPECS):
            dt = torch.float64 if precision == L.PRECISION_FP64 else torch.float32
            acc = tuple(torch.empty(s.n, dtype=dt, device="cuda:0") for _ in range(3))
            with TreePmCoordinator(s.mesh) as coord:
                coord.solve_active_set(*state.x[b], state.m[b], ForceAccumulator(acc), fwd_opts[b])
                coord.synchronize()
            ref = np.stack([a.double().cpu().numpy() for a in acc])
            assert rel_l2(batched[b], ref) < tol, (b, s)


def test_batched_forward_equals_box_by_box_steps():
    with E.Ensemble(SPECS, workers=4) as ens:
        fwd = E.EnsembleForward(ens, steps=2, precision=L.PRECISION_FP64)
        fwd.run()
        xs = [np.stack([c.cpu().numpy() for c in x]) for x in fwd.x]
    for b, s in enumerate(SPECS):
        with E.Ensemble([s], workers=1) as one:
            f1 = E.EnsembleForward(one, steps=2, precision=L.PRECISION_FP64)
            f1.run()
            x1 = np.stack([c.cpu().numpy() for c in f1.x[0]])
        d = x1 - xs[b]
        d -= s.box * np.round(d / s.box)
        assert np.abs(d).max() < 1e-9 * s.box


def test_errors_name_the_box():
    with E.Ensemble(SPECS[:2], workers=2) as ens:
        fwd = E.EnsembleForward(ens, steps=1, precision=L.PRECISION_FP64)
        bad = [E.box_options(s, L.PRECISION_FP64) for s in SPECS[:2]]
        bad[1].rcut_cells = 1000.0  # r_cut beyond half the box
        with pytest.raises(ValueError, match="ensemble box 1"):
            ens.solve(fwd.x, fwd.m, fwd.g, bad)


def test_box_gradient_matches_central_differences():
    s = E.ensemble_specs(counts=((16, 1),), seed=5)[0]
    loss, grad, nf, nv = E.box_gradient(s, steps=2)
    assert nf >= 3 and nv == 3
    _, psi = E.box_initial_conditions(s)
    from cosmosim import adjoint

    prob = adjoint.KdkProblem(n_side=s.n_side, box=s.box, mesh=s.mesh,
                              scale_factors=list(np.linspace(s.a_begin, s.a_end, 3)),
                              softening=s.box / s.n_side / 30.0, loss_mesh=s.n_side)
    try:
        from cosmosim import ics

        tdisp = ics.zeldovich_displacement(s.n_side, s.box, s.spectral_index, 1.0,
                                           np.random.default_rng(s.seed + 1))
        target = prob.cic_density(prob.wrap(prob.lattice + torch.tensor(tdisp, dtype=torch.float64,
                                                                        device="cuda:0")))
        p0 = torch.tensor(psi, dtype=torch.float64, device="cuda:0")
        rng = np.random.default_rng(3)
        d = rng.standard_normal(psi.shape)
        d /= np.linalg.norm(d)
        dt = torch.tensor(d, dtype=torch.float64, device="cuda:0")
        h = 1e-6 * s.box / s.n_side
        with torch.no_grad():
            lp = float(adjoint.density_loss(p0 + h * dt, target, prob))
            lm = float(adjoint.density_loss(p0 - h * dt, target, prob))
        fd = (lp - lm) / (2 * h)
        an = float((grad * dt).sum())
        assert abs(fd - an) <= 1e-4 * abs(an), (fd, an)
    finally:
        prob.close()


def test_explicit_adjoint_sweeps_equal_autograd():
    # the sweeps without the autograd engine reproduce torch.autograd's
    # gradient of the same loss (same operations; PM atomics: round-off)
    s = SPECS[0]
    loss_a, grad_a, nf_a, nv_a = E.box_gradient(s, 3)
    loss_e, grad_e, nf_e, nv_e = E.box_gradient(s, 3, explicit=True)
    assert (nf_a, nv_a) == (nf_e, nv_e) == (4, 4)
    assert abs(loss_e - loss_a) <= 1e-12 * abs(loss_a)
    assert rel_l2(grad_e.cpu().numpy(), grad_a.cpu().numpy()) < 1e-11


def test_threaded_ensemble_gradients_equal_serial():
    # boxes differentiated on four host threads, each with its own torch stream
    serial = [E.box_gradient(s, 2, explicit=True) for s in SPECS]
    threaded = E.ensemble_gradients(SPECS, 2, workers=4)
    for (l0, g0, f0, v0), (l1, g1, f1, v1), s in zip(serial, threaded, SPECS):
        assert (f0, v0) == (f1, v1)
        assert abs(l1 - l0) <= 1e-12 * abs(l0), s
        assert rel_l2(g1.cpu().numpy(), g0.cpu().numpy()) < 1e-11, s


def test_launch_shapes_leave_results_unchanged():
    # lines per block of the line passes and waves per block of the timed walk
    # change how the work is cut, not what is computed
    s = SPECS[-1]
    lattice, psi = E.box_initial_conditions(s)
    pos = np.mod(lattice + psi, s.box)
    pos[pos >= s.box] = 0.0
    dev = [torch.from_numpy(np.ascontiguousarray(p)).to("cuda:0") for p in pos]
    m = torch.full((s.n,), 1.0, dtype=torch.float64, device="cuda:0")
    o = E.box_options(s, L.PRECISION_MIXED)
    out = {}
    with TreePmCoordinator(s.mesh) as coord:
        for shape in ((0, 0), (4, 2), (8, 8), (1, 4)):
            coord.set_launch_shape(*shape)
            acc = tuple(torch.empty(s.n, dtype=torch.float32, device="cuda:0") for _ in range(3))
            sr = tuple(torch.zeros(s.n, dtype=torch.float32, device="cuda:0") for _ in range(3))
            coord.solve_active_set(*dev, m, ForceAccumulator(acc, None, None, sr), o)
            coord.synchronize()
            out[shape] = (np.stack([a.cpu().numpy() for a in acc]), np.stack([a.cpu().numpy() for a in sr]))
        with pytest.raises(L.ChuiError):
            coord.set_launch_shape(3, 0)
        with pytest.raises(L.ChuiError):
            coord.set_launch_shape(0, 5)
    ref_acc, ref_sr = out[(0, 0)]
    for shape, (acc, sr) in out.items():
        # per-target walk sums do not depend on the block size; PM by round-off
        assert np.array_equal(sr, ref_sr), shape
        assert rel_l2(acc, ref_acc) < 1e-6, shape


def test_autotune_picks_from_its_candidates():
    s = SPECS[0]
    r = E.autotune(s, rounds=2, repeats=1, leaves=(16, 32), fft_lines=(0, 4), walk_waves=(4, 2))
    assert r["leaf"] in (16, 32) and r["fft_lines"] in (0, 4) and r["walk_waves"] in (4, 2)
    assert [st["knob"] for st in r["steps"]] == ["leaf", "fft_lines", "walk_waves"]
    for st in r["steps"]:
        assert st["noise"] >= 0.02
