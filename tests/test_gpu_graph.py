"""GPU checks of capture-safe solves and their HIP graphs (no reference
counterpart: the reference drives every solve from the host).

- A capture-safe solve (every per-build value kept on the device, nothing read
  back) equals the ordinary solve: the tree-only lanes bit for bit (the tree
  and the walk are deterministic), the PM lanes and totals to the round-off of
  the deposit's atomic summation order, also for an active subset.
- A recorded solve replayed three times, with eager work and synchronisations
  between the replays, reproduces the eager capture-safe solve every time.
- The faults a capture-safe solve reports through the fault word instead of
  throwing: a mass uniformity other than the prepared one (CHUI_ERR_RUNTIME)
  and an active index out of range (CHUI_ERR_INVALID_ARGUMENT), both forced
  from the host inputs; recording on an unprepared context is refused.
- The batched ensemble forward recorded as one graph (boxes as branches)
  equals the eager batched forward after every replay.
- FP64 capture-safe solves equal ordinary ones (tree-only lanes bit for bit),
  also on a clustered set whose walk accepts quadrupole nodes.
- A force VJP recorded with torch.cuda.graph on the context's stream (the
  library detects the recording; the accepted-node list's count stays on the
  device) equals the eager VJP on every replay.
- The gradients of several boxes recorded as one graph (each box a branch:
  its FP64 solves and VJPs, CIC density and transpose and the element-wise
  KDK work) equal the eager gradients after every replay and for new inputs.
"""
from __future__ import annotations

import numpy as np
import pytest

from helpers import empty_dev, options_struct, rel_l2, to_dev

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
if not torch.cuda.is_available():  # pragma: no cover - GPU-only module
    pytest.skip("no HIP device", allow_module_level=True)

from cosmosim import ForceAccumulator, TreePmCoordinator, ics  # noqa: E402
from cosmosim import _lib as L  # noqa: E402
from cosmosim import ensemble as E  # noqa: E402
from test_oracle_pins import clustered  # noqa: E402


def _case(name):
    if name == "clustered":
        x, y, z, m = clustered(1200, 5, 10.0)
        return (x, y, z, m), options_struct(10.0, eps=0.01), 16
    w = ics.config(2, n_side=32)
    o = options_struct(w.box, eps=w.softening)
    return (w.x, w.y, w.z, w.mass), o, w.mesh


def _lanes(n, n_active):
    acc = tuple(empty_dev(n_active, torch.float32) for _ in range(3))
    pm = tuple(empty_dev(n_active, torch.float32) for _ in range(3))
    sr = tuple(torch.zeros(n, dtype=torch.float32, device="cuda:0") for _ in range(3))
    return acc, pm, sr


def _solve(coord, dev, o, active=None):
    n = int(dev[0].shape[0])
    n_active = n if active is None else int(active.shape[0])
    acc, pm, sr = _lanes(n, n_active)
    torch.cuda.synchronize()  # the library's stream does not wait for torch's
    coord.solve_active_set_with_pm_cadence(*dev, ForceAccumulator(acc, active, pm, sr), o, True)
    coord.synchronize()
    return [np.stack([a.cpu().numpy() for a in lane]) for lane in (acc, pm, sr)]


def _assert_same(got, ref):
    acc, pm, sr = got
    acc0, pm0, sr0 = ref
    assert np.array_equal(sr, sr0)  # deterministic tree and walk
    assert rel_l2(pm, pm0) < 1e-6   # deposit atomics: summation order
    assert rel_l2(acc, acc0) < 1e-6


@pytest.mark.parametrize("name", ["clustered", "lattice_halos"])
def test_capture_safe_solves_equal_ordinary(name):
    src, o, mesh = _case(name)
    o.precision = L.PRECISION_MIXED
    dev = to_dev(*src)
    n = int(dev[0].shape[0])
    active = torch.from_numpy(np.arange(0, n, 3, dtype=np.uint32)).to("cuda:0")
    with TreePmCoordinator(mesh) as plain:
        ref = _solve(plain, dev, o)
        ref_active = _solve(plain, dev, o, active)
    with TreePmCoordinator(mesh) as coord:
        coord.set_capture_safe(True)
        _assert_same(_solve(coord, dev, o), ref)  # the preparing (ordinary) solve
        for _ in range(2):                         # capture-safe solves
            _assert_same(_solve(coord, dev, o), ref)
        _assert_same(_solve(coord, dev, o, active), ref_active)


def test_graph_replays_equal_the_eager_solve():
    src, o, mesh = _case("lattice_halos")
    o.precision = L.PRECISION_MIXED
    dev = to_dev(*src)
    n = int(dev[0].shape[0])
    with TreePmCoordinator(mesh) as coord, TreePmCoordinator(mesh) as other:
        coord.set_capture_safe(True)
        _solve(coord, dev, o)
        ref = _solve(coord, dev, o)
        acc, pm, sr = _lanes(n, n)
        torch.cuda.synchronize()
        coord.graph_begin()
        try:
            coord.solve_active_set_with_pm_cadence(*dev, ForceAccumulator(acc, None, pm, sr), o, True)
        finally:
            graph = coord.graph_end()
        try:
            for rep in range(3):
                for lane in acc + pm + sr:
                    lane.fill_(float("nan") if lane.shape[0] == n else 0.0)
                torch.cuda.synchronize()
                graph.launch()
                coord.synchronize()
                got = [np.stack([a.cpu().numpy() for a in lane]) for lane in (acc, pm, sr)]
                _assert_same(got, ref)
                # eager work of the program between replays
                _solve(other, dev, o)
                torch.cuda.synchronize()
        finally:
            graph.close()


def test_capture_safe_reports_changed_mass_uniformity():
    src, o, mesh = _case("lattice_halos")
    o.precision = L.PRECISION_MIXED
    x, y, z, m = src
    dev = to_dev(x, y, z, m)
    with TreePmCoordinator(mesh) as coord:
        coord.set_capture_safe(True)
        _solve(coord, dev, o)  # prepares with uniform masses
        m2 = m.copy()
        m2[5] *= 1.5
        with pytest.raises(L.ChuiError) as err:
            _solve(coord, to_dev(x, y, z, m2), o)
        assert err.value.status == L.CHUI_ERR_RUNTIME
        assert "prepared" in str(err.value)
        _solve(coord, dev, o)  # the context stays usable


def test_capture_safe_reports_active_index_out_of_range():
    src, o, mesh = _case("clustered")
    o.precision = L.PRECISION_MIXED
    dev = to_dev(*src)
    n = int(dev[0].shape[0])
    with TreePmCoordinator(mesh) as coord:
        coord.set_capture_safe(True)
        _solve(coord, dev, o)
        bad = torch.from_numpy(np.array([0, 7, n + 3], dtype=np.uint32)).to("cuda:0")
        with pytest.raises(L.ChuiError) as err:
            _solve(coord, dev, o, bad)
        assert err.value.status == L.CHUI_ERR_INVALID_ARGUMENT


def test_recording_needs_a_prepared_context():
    src, o, mesh = _case("clustered")
    o.precision = L.PRECISION_MIXED
    dev = to_dev(*src)
    n = int(dev[0].shape[0])
    with TreePmCoordinator(mesh) as coord:
        with pytest.raises(L.ChuiError):
            coord.graph_begin()  # not capture-safe
        coord.set_capture_safe(True)
        coord.graph_begin()
        acc, pm, sr = _lanes(n, n)
        with pytest.raises(L.ChuiError) as err:
            coord.solve_active_set_with_pm_cadence(*dev, ForceAccumulator(acc, None, pm, sr), o, True)
        assert "prepared" in str(err.value)
        coord.graph_end().close()
        o.acceptance_policy = 1  # CHUI_ACCEPT_ADAPTIVE: not a capture-safe solve
        with pytest.raises(L.ChuiError) as err:
            coord.solve_active_set_with_pm_cadence(*dev, ForceAccumulator(acc), o, True)
        assert err.value.status == L.CHUI_ERR_UNSUPPORTED


def test_ensemble_graph_equals_eager_forward():
    specs = E.ensemble_specs(counts=((32, 3), (64, 1)))
    with E.Ensemble(specs, workers=4) as ens:
        fwd = E.EnsembleForward(ens, steps=2, precision=L.PRECISION_MIXED)
        start = fwd.state()
        fwd.run()
        ref = fwd.state()
        ref_g = [tuple(t.clone() for t in g) for g in fwd.g]
        fwd.set_state(start)
        graph = fwd.capture()
        try:
            for rep in range(2):
                fwd.set_state(start)
                graph.launch()
                fwd.synchronize()
                got = fwd.state()
                for b, s in enumerate(specs):
                    spacing = s.box / s.n_side
                    for k in range(3):  # positions: round-off of the PM deposit over two steps
                        assert float((got[b][k] - ref[b][k]).abs().max()) < 1e-6 * spacing
                    v = torch.stack(got[b][3:]).cpu().numpy()
                    v0 = torch.stack(ref[b][3:]).cpu().numpy()
                    assert rel_l2(v, v0) < 1e-5
                    g = torch.stack(fwd.g[b]).double().cpu().numpy()
                    g0 = torch.stack(ref_g[b]).double().cpu().numpy()
                    assert rel_l2(g, g0) < 1e-5
                # eager work between replays
                fwd.set_state(start)
                fwd.run()
        finally:
            graph.close()


# ---- FP64 capture-safe solves, recorded VJPs and the gradient graph ----------

def _quadrupole_case():
    # the clustered set of the quadrupole VJP tests: thousands of (target,
    # node) pairs accept a node, so the VJP's accepted-node list is not empty
    from test_gpu_adjoint_quadrupole import BOX as QB, EPS as QE, G as QG, LEAF as QL, MESH as QM, clustered_set

    x, m = clustered_set()
    o = options_struct(QB, G=QG, eps=QE, leaf=QL, quad=True)
    return (x[0], x[1], x[2], m), o, QM


@pytest.mark.parametrize("name", ["clustered", "lattice_halos", "quadrupole"])
def test_fp64_capture_safe_solves_equal_ordinary(name):
    src, o, mesh = _quadrupole_case() if name == "quadrupole" else _case(name)
    o.precision = L.PRECISION_FP64
    dev = to_dev(*src)
    n = int(dev[0].shape[0])

    def solve(coord):
        acc = tuple(empty_dev(n, torch.float64) for _ in range(3))
        sr = tuple(torch.zeros(n, dtype=torch.float64, device="cuda:0") for _ in range(3))
        torch.cuda.synchronize()
        coord.solve_active_set_with_pm_cadence(*dev, ForceAccumulator(acc, None, None, sr), o, True)
        coord.synchronize()
        return np.stack([a.cpu().numpy() for a in acc]), np.stack([a.cpu().numpy() for a in sr])

    with TreePmCoordinator(mesh) as plain:
        acc0, sr0 = solve(plain)
    with TreePmCoordinator(mesh) as coord:
        coord.set_capture_safe(True)
        for _ in range(3):  # the preparing solve, then capture-safe ones
            acc, sr = solve(coord)
            assert np.array_equal(sr, sr0)
            assert rel_l2(acc, acc0) < 1e-12


def test_recorded_force_vjp_equals_eager():
    # the VJP recorded with torch.cuda.graph on the context's stream (the
    # library detects the recording and takes the accepted-node count from
    # the device), replayed three times with eager work between replays
    src, o, mesh = _quadrupole_case()
    dev = to_dev(*src)
    n = int(dev[0].shape[0])
    rng = np.random.default_rng(3)
    lam = [torch.from_numpy(rng.normal(size=n)).to("cuda:0") for _ in range(3)]

    def vjp(coord, recorded=False):
        # eager calls order the library's stream after torch's and wait for
        # it; a recorded call shares torch's stream
        out = tuple(torch.zeros(n, dtype=torch.float64, device="cuda:0") for _ in range(3))
        if not recorded:
            torch.cuda.synchronize()
        coord.force_vjp(*dev, lam, out, o)
        if not recorded:
            coord.synchronize()
        return out

    with TreePmCoordinator(mesh) as plain:
        ref = np.stack([a.cpu().numpy() for a in vjp(plain)])
    with TreePmCoordinator(mesh) as coord, TreePmCoordinator(mesh) as other:
        coord.set_capture_safe(True)
        acc = tuple(empty_dev(n, torch.float64) for _ in range(3))
        torch.cuda.synchronize()
        coord.solve_active_set(*dev, ForceAccumulator(acc), o)  # prepares the context
        coord.synchronize()
        for _ in range(2):
            got = vjp(coord)
            assert rel_l2(np.stack([a.cpu().numpy() for a in got]), ref) < 1e-12
        stream = torch.cuda.Stream()
        coord.set_stream(stream.cuda_stream)
        torch.cuda.synchronize()
        graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(graph, stream=stream):
            out = vjp(coord, recorded=True)
        for rep in range(3):
            for a in out:
                a.fill_(float("nan"))
            torch.cuda.synchronize()
            graph.replay()
            torch.cuda.synchronize()
            coord.fault_check()
            assert rel_l2(np.stack([a.cpu().numpy() for a in out]), ref) < 1e-12, rep
            vjp(other)  # eager work between replays
        del graph


def test_gradient_graph_equals_eager_gradients():
    specs = E.ensemble_specs(counts=((16, 2), (24, 1)), seed=91)
    eager = [E.box_gradient(s, 2, explicit=True) for s in specs]
    gg = E.GradientGraph(specs, 2)
    try:
        for rep in range(3):
            losses, grads = gg.replay()
            for (l0, g0, _, _), l1, g1, s in zip(eager, losses, grads, specs):
                assert abs(l1 - l0) <= 1e-11 * abs(l0), (rep, s)
                assert rel_l2(g1.cpu().numpy(), g0.cpu().numpy()) < 1e-10, (rep, s)
            E.box_gradient(specs[0], 2, explicit=True)  # eager work between replays
            torch.cuda.synchronize()
        # new inputs go through the recorded graph: the gradient at psi + dpsi
        rng = np.random.default_rng(5)
        moved = [p + 0.01 * torch.from_numpy(rng.normal(size=tuple(p.shape))).to(p.device) for p in gg.psi]
        saved = [p.clone() for p in gg.psi]
        _, g_new = gg.replay(moved)
        g_new = [g.clone() for g in g_new]
        for b, s in enumerate(specs):
            prob, _, target = E.box_problem(s, 2)
            try:
                _, g_ref = E.adjoint.loss_and_gradient(moved[b], target, prob)
            finally:
                prob.close()
            assert rel_l2(g_new[b].cpu().numpy(), g_ref.cpu().numpy()) < 1e-10, s
        gg.replay(saved)
    finally:
        gg.close()
