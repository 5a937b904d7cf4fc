"""GPU parity: the HIP path through the C ABI against the CPU oracle.

Bit-exact for integer/topology/moment work (Morton keys, sort order, unwrap
anchor, node membership, FP64 moments); accelerations within the tolerances
written in each test.
"""
from __future__ import annotations

import dataclasses

import numpy as np
import pytest

from helpers import empty_dev, error_metrics, options_struct, rel_l2, to_dev
from oracle import oracle as O

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
if not torch.cuda.is_available():  # pragma: no cover - GPU-only module
    pytest.skip("no HIP device", allow_module_level=True)

from cosmosim import ForceAccumulator, TreePmCoordinator, ics  # noqa: E402
from cosmosim import _lib as L  # noqa: E402


def clustered(n, box, seed, n_halos=5, frac=0.5, scale=0.02):
    rng = np.random.default_rng(seed)
    pos = rng.uniform(0, box, (3, n))
    nh = int(frac * n)
    centres = rng.uniform(0, box, (n_halos, 3))
    which = rng.integers(0, n_halos, nh)
    pos[:, :nh] = (centres[which] + rng.normal(0, scale * box, (nh, 3))).T
    pos = np.mod(pos, box)
    pos[pos >= box] = 0.0
    return pos[0].copy(), pos[1].copy(), pos[2].copy()


def gpu_solve(w_x, w_y, w_z, m, opts, mesh, precision=0, active=None, capture=True, diagnostics=True):
    dt = torch.float64 if precision == 0 else torch.float32
    x, y, z, mm = to_dev(w_x, w_y, w_z, m)
    na = len(w_x) if active is None else len(active)
    acc = [empty_dev(na, dt) for _ in range(3)]
    pm = [empty_dev(na, dt) for _ in range(3)] if capture else None
    sr = [empty_dev(len(w_x), dt) for _ in range(3)] if capture else None  # source-row indexed
    act = to_dev(np.asarray(active, np.uint32).view(np.int32))[0] if active is not None else None
    opts.precision = precision
    with TreePmCoordinator(mesh) as coord:
        res = coord.solve_active_set(x, y, z, mm, ForceAccumulator(tuple(acc), act, tuple(pm) if pm else None,
                                                                 tuple(sr) if sr else None), opts,
                                     want_diagnostics=diagnostics)
        torch.cuda.synchronize()
        tree = coord.tree_arrays()
    host = lambda ts: [t.double().cpu().numpy() for t in ts] if ts else None  # noqa: E731
    return host(acc), host(pm), host(sr), res.diagnostics, tree


# ---------------------------------------------------------------------------
@pytest.mark.parametrize("case", ["uniform", "clustered", "duplicates", "tiny"])
def test_morton_keys_and_order_bit_exact(case):
    rng = np.random.default_rng(7)
    if case == "uniform":
        x, y, z = rng.uniform(-3, 7, (3, 20000))
    elif case == "clustered":
        x, y, z = clustered(30000, 50.0, 3)
    elif case == "duplicates":
        base = rng.uniform(0, 1, (3, 500))
        x, y, z = np.repeat(base, 9, axis=1)
    else:
        x, y, z = rng.uniform(0, 1, (3, 3))
    keys_ref, order_ref = O.morton_order(x, y, z)
    dx, dy, dz = to_dev(x, y, z)
    keys = torch.zeros(len(x), dtype=torch.int64, device="cuda:0")
    order = torch.zeros(len(x), dtype=torch.int32, device="cuda:0")
    with TreePmCoordinator(8) as c:
        c.morton(dx, dy, dz, keys, order)
    assert np.array_equal(keys.cpu().numpy().view(np.uint64), keys_ref)
    assert np.array_equal(order.cpu().numpy().view(np.uint32), order_ref)


@pytest.mark.parametrize("case", ["uniform", "seam_cluster", "two_points", "single", "duplicates", "noncanonical",
                                  "equal_gaps", "dyadic_ties", "bucket_edges", "large_clustered"])
def test_unwrap_anchor_bit_exact(case):
    rng = np.random.default_rng(11)
    L_box = 10.0
    if case == "uniform":
        v = rng.uniform(0, L_box, 10000)
    elif case == "seam_cluster":
        v = np.mod(rng.normal(0.0, 0.3, 5000), L_box)
    elif case == "two_points":
        v = np.array([1.0, 6.0])
    elif case == "single":
        v = np.array([3.3])
    elif case == "duplicates":
        v = np.repeat(rng.uniform(0, L_box, 50), 20)
    elif case == "noncanonical":
        v = rng.uniform(-25, 35, 4000)
    elif case == "dyadic_ties":  # values on a coarse dyadic grid: many equal gaps
        v = rng.integers(0, 64, 5000) * (L_box / 64)
    elif case == "bucket_edges":  # every value on a bucket boundary of the binned search
        v = np.arange(256) * (L_box / 256)
        v = np.delete(v, [17, 18, 100, 101, 200])
    elif case == "large_clustered":
        v = np.mod(np.concatenate([rng.normal(c, 0.05, 40000) for c in (1.0, 4.5, 9.9)]), L_box)
    else:  # a lattice: every gap ties, the smallest anchor wins
        v = (np.arange(64) + 0.5) * (L_box / 64)
        rng.shuffle(v)
    ref_out, ref_anchor = O.unwrap_axis(v, L_box)
    (dv,) = to_dev(v)
    out = torch.zeros_like(dv)
    with TreePmCoordinator(8) as c:
        anchor = c.unwrap(dv, L_box, out)
    assert anchor == ref_anchor
    assert np.array_equal(out.cpu().numpy(), ref_out)


def _tree_cases():
    yield "config0_16", ics.config(0, n_side=16)
    w = ics.config(1, n_side=32)
    yield "config1_32", w
    # 40 particles within ~2^-19 of the box around one point: leaves at depth
    # 18, below the 16 levels the first key sort covers, so the build must
    # fall back to the full-width sort and still match bit for bit.
    d = ics.config(0, n_side=16)
    rng = np.random.default_rng(7)
    x, y, z = d.x.copy(), d.y.copy(), d.z.copy()
    centre = np.array([1.234567, 3.456789, 0.987654])
    tight = centre[:, None] + rng.normal(0.0, d.box * 2.0**-19, (3, 40))
    x[:40], y[:40], z[:40] = tight
    yield "deep_cluster", dataclasses.replace(d, x=x, y=y, z=z)


@pytest.mark.parametrize("name,w", list(_tree_cases()))
def test_tree_topology_and_moments_bit_exact(name, w):
    opts = options_struct(w.box, eps=w.softening)
    _, _, _, diag, tree = gpu_solve(w.x, w.y, w.z, w.mass, opts, w.mesh)
    ux = [O.unwrap_axis(c, w.box)[0] for c in (w.x, w.y, w.z)]
    ref = O.build_tree(*ux, w.mass, 16, True, w.softening)
    assert tree["count"].size == ref["count"].size
    assert np.array_equal(tree["order"], ref["order"])
    assert np.array_equal(tree["begin"], ref["begin"])
    assert np.array_equal(tree["count"], ref["count"])
    assert np.array_equal(tree["child_index"], ref["child_index"])
    assert np.array_equal(tree["geometry"], ref["geometry"])
    assert np.array_equal(tree["moments"], ref["moments"])
    assert diag["node_count"] == ref["count"].size
    assert diag["leaf_count"] == int(np.all(ref["child_index"] == 0xFFFFFFFF, axis=1).sum())


@pytest.mark.parametrize("assignment,deconv", [(0, 0), (0, 1), (1, 1)])
def test_pm_long_range_fp64(assignment, deconv):
    w = ics.config(0, n_side=16)
    opts = options_struct(w.box, eps=w.softening, assignment=assignment, deconv=deconv)
    _, pm, _, _, _ = gpu_solve(w.x, w.y, w.z, w.mass, opts, w.mesh)
    ref = O.pm_accel(w.x, w.y, w.z, w.mass, opts, w.mesh)
    assert rel_l2(pm, ref) < 1e-10


def test_short_range_fp64_matches_oracle():
    w = ics.config(1, n_side=32)
    opts = options_struct(w.box, eps=w.softening)
    _, _, sr, diag, _ = gpu_solve(w.x, w.y, w.z, w.mass, opts, w.mesh)
    ref, cnt, _ = O.short_range(w.x, w.y, w.z, w.mass, opts, w.mesh)
    assert rel_l2(sr, ref) < 1e-12
    assert diag["direct_pair_evaluations"] == cnt["pairs"]
    assert diag["accepted_internal_multipoles"] == cnt["multipoles"]
    assert diag["cutoff_pruned_nodes"] == cnt["pruned"]
    assert diag["visited_nodes"] == cnt["visited"]


def test_total_fp64_within_1e5_relative_rms():
    w = ics.config(0, n_side=32)
    opts = options_struct(w.box, eps=w.softening)
    acc, _, _, _, _ = gpu_solve(w.x, w.y, w.z, w.mass, opts, w.mesh)
    ref = O.solve(w.x, w.y, w.z, w.mass, opts, w.mesh)
    assert rel_l2(acc, ref) < 1e-5


def test_mixed_precision_error_budget():
    w = ics.config(1, n_side=32)
    opts = options_struct(w.box, eps=w.softening)
    acc, pm, sr, _, _ = gpu_solve(w.x, w.y, w.z, w.mass, opts, w.mesh, precision=1)
    ref = O.solve(w.x, w.y, w.z, w.mass, opts, w.mesh)
    m = error_metrics(acc, ref)
    print("mixed vs fp64 oracle:", m)
    assert m["relative_l2"] < 1e-4
    assert m["p99"] < 1e-3


def test_mixed_waves_spanning_the_box():
    # Consecutive tree-order targets can lie far apart (jumps of the octant
    # order), so one wave may hold targets from opposite ends of the box.
    # Every target must still get the exact minimum image: compare the mixed
    # walk with the FP64 walk particle by particle.
    rng = np.random.default_rng(17)
    box = 50.0
    centres = np.array([[1.0, 1.0, 1.0], [49.0, 49.0, 49.0], [25.0, 1.0, 49.0], [1.0, 49.0, 25.0]])
    pts = np.concatenate([c + rng.normal(0.0, 0.4, (40, 3)) for c in centres])
    pts = np.mod(pts, box)
    x, y, z = (pts[:, a].copy() for a in range(3))
    m = np.ones(x.size)
    opts = options_struct(box, eps=0.01)
    _, _, sr64, _, _ = gpu_solve(x, y, z, m, opts, 32, precision=0)
    _, _, sr32, _, _ = gpu_solve(x, y, z, m, opts, 32, precision=1)
    e = np.linalg.norm(np.stack(sr32) - np.stack(sr64), axis=0) / np.linalg.norm(np.stack(sr64), axis=0)
    assert e.max() < 1e-4


def test_active_subset_equals_full_rows():
    # The short-range walk is deterministic per target, so the tree-only lanes
    # (source-row indexed: only the active rows are written) match bit for
    # bit; the PM deposit uses FP64 atomics, so totals agree to round-off only.
    # (the per-lane mapping on both sides: a per-wave walk sums in another order)
    w = ics.config(0, n_side=16)
    opts = options_struct(w.box, eps=w.softening)
    opts.target_mapping = L.TARGETS_PER_LANE
    full, _, full_sr, _, _ = gpu_solve(w.x, w.y, w.z, w.mass, opts, w.mesh)
    rng = np.random.default_rng(5)
    active = rng.choice(w.x.size, 777, replace=False).astype(np.uint32)
    sub, _, sub_sr, _, _ = gpu_solve(w.x, w.y, w.z, w.mass, opts, w.mesh, active=active)
    inactive = np.setdiff1d(np.arange(w.x.size), active)
    for a in range(3):
        assert np.array_equal(sub_sr[a][active], full_sr[a][active])
        assert np.all(sub_sr[a][inactive] == 0.0)  # untouched (the lanes start zeroed)
    assert rel_l2(sub, [f[active] for f in full]) < 1e-13


@pytest.mark.parametrize("precision", [0, 1])
def test_walk_without_diagnostics_matches_counted_walk(precision):
    # Diagnostic solves run a walk variant that also counts. The plain walks
    # classify nodes against each wave's target box and sum the same pairs and
    # multipoles in another order (FP64: with a polynomial table of the
    # Gaussian term instead of erf/exp), so they agree to the round-off of the
    # sums: 1e-12 in FP64, FP32 rounding in mixed precision.
    w = ics.config(1, n_side=32)
    opts = options_struct(w.box, eps=w.softening)
    _, _, sr_d, diag, _ = gpu_solve(w.x, w.y, w.z, w.mass, opts, w.mesh, precision=precision)
    _, _, sr_p, none, _ = gpu_solve(w.x, w.y, w.z, w.mass, opts, w.mesh, precision=precision, diagnostics=False)
    assert diag and not none
    if precision == 0:
        assert rel_l2(sr_p, sr_d) < 1e-12
    else:
        assert rel_l2(sr_p, sr_d) < 2e-6
        per = np.linalg.norm(np.stack(sr_p) - np.stack(sr_d), axis=0) / np.maximum(
            np.linalg.norm(np.stack(sr_d), axis=0), 1e-30)
        assert per.max() < 1e-4


@pytest.mark.parametrize("precision", [0, 1])
@pytest.mark.parametrize("case", ["one", "two", "coincident16", "coincident16_unsoftened", "noncanonical",
                                  "rectangular", "nonuniform_mass"])
def test_edge_cases_match_oracle(case, precision):
    rng = np.random.default_rng(3)
    box = 10.0
    mesh = 16
    if case == "one":
        x, y, z = np.array([1.0]), np.array([2.0]), np.array([3.0])
    elif case == "two":
        x, y, z = np.array([1.0, 1.4]), np.array([2.0, 2.1]), np.array([3.0, 2.9])
    elif case.startswith("coincident16"):
        x, y, z = rng.uniform(0, box, (3, 300))
        x[:16], y[:16], z[:16] = 4.0, 4.0, 4.0
    elif case == "noncanonical":
        x, y, z = rng.uniform(-box, 2 * box, (3, 400))
    elif case == "rectangular":
        x, y, z = rng.uniform(0, 1, (3, 600)) * np.array([[10.0], [12.0], [8.0]])
        box = (10.0, 12.0, 8.0)
    else:
        x, y, z = rng.uniform(0, box, (3, 500))
    m = np.full(x.size, 2.5) if case != "nonuniform_mass" else rng.uniform(0.5, 2.0, x.size)
    opts = options_struct(box, eps=0.0 if case.endswith("unsoftened") else 0.01, G=1.0)
    acc, _, _, _, _ = gpu_solve(x, y, z, m, opts, mesh, precision=precision)
    ref = O.solve(x, y, z, m, opts, mesh)
    assert np.all(np.isfinite(np.stack(acc)))
    # force scale of one particle at the split scale: a lone particle's force is
    # round-off of the PM self-force, so it is compared against that scale.
    h = min(np.atleast_1d(box)) / mesh
    scale = np.sum(m) / (1.25 * h) ** 2
    err = np.sqrt(np.sum((np.stack(acc) - np.stack(ref)) ** 2) / x.size)
    tol = 1e-10 if precision == 0 else 1e-4  # FP64 parity; mixed-precision budget
    assert err < tol * max(np.sqrt(np.sum(np.stack(ref) ** 2) / x.size), scale)


def test_kick_drift_and_reductions():
    rng = np.random.default_rng(9)
    n = 10000
    box = (20.0, 20.0, 20.0)
    x, y, z = rng.uniform(0, 20, (3, n))
    v = rng.normal(0, 100, (3, n))
    a = rng.normal(0, 1e3, (3, n))
    m = rng.uniform(1, 2, n)
    dx, dy, dz, dvx, dvy, dvz, dax, day, daz, dm = to_dev(x, y, z, *v, *a, m)
    with TreePmCoordinator(8) as c:
        c.kick(dvx, dvy, dvz, dax, day, daz, 0.013, 0.97)
        c.drift(dx, dy, dz, dvx, dvy, dvz, 0.0021, box)
        mm = c.mass_momentum(dm, dvx, dvy, dvz)
    hv = [t.cpu().numpy() for t in (dvx, dvy, dvz)]
    hx = [t.cpu().numpy() for t in (dx, dy, dz)]
    rv = [vi.copy() for vi in v]
    O.kick(*rv, *a, 0.013, 0.97)
    rx = [x.copy(), y.copy(), z.copy()]
    O.drift(*rx, *rv, 0.0021, box)
    for i in range(3):
        assert np.array_equal(hv[i], rv[i])
        assert np.array_equal(hx[i], rx[i])
    assert abs(mm[0] - np.sum(m)) < 1e-12 * np.sum(m)
    for i in range(3):
        ref = float(np.sum(m * rv[i]))
        assert abs(mm[1 + i] - ref) < 1e-12 * float(np.sum(np.abs(m * rv[i])))


def test_ewald_gpu_matches_oracle():
    rng = np.random.default_rng(4)
    n = 300
    box = (1.0, 1.0, 1.0)
    x, y, z = rng.uniform(0, 1, (3, n))
    m = rng.uniform(0.5, 1.5, n)
    targets = np.arange(0, n, 7, dtype=np.uint32)
    ref = O.ewald(x, y, z, m, targets, box, 1.0, 2.5, 3, 10, eps=0.0)
    dx, dy, dz, dm = to_dev(x, y, z, m)
    (dt,) = to_dev(targets.view(np.int32))
    out = [empty_dev(targets.size, torch.float64) for _ in range(3)]
    with TreePmCoordinator(8) as c:
        c.ewald(dx, dy, dz, dm, dt, box, 1.0, 11.8, 0, 24, 0.0, *out)
    got = [t.cpu().numpy() for t in out]
    assert rel_l2(got, ref) < 1e-10
