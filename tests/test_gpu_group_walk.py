"""GPU parity of the adaptive acceptance and the group walk against the oracle.

Through chui_treepm_solve_adaptive_device: the traversal counters (visited,
pruned, pairs, accepted nodes and leaves, groups, fallbacks, culling and list
sizes) equal the oracle's exactly, and the short-range accelerations agree to
1e-12 relative L2 (FP64; summation order differs). Cases follow the reference's
group-walk culling test (tests/integration/test_tree_pm_coupling_periodic.cpp:
2140-2230) and add accepted-node-rich clustered sets, both cutoff node policies,
both node expansions, both estimators of the per-target walk, active subsets,
groups above 64 targets and inputs outside the bounded metric.
"""
from __future__ import annotations

import numpy as np
import pytest

from helpers import empty_dev, rel_l2, seam_cluster_options, seam_clusters, to_dev
from oracle import oracle as O

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
if not torch.cuda.is_available():  # pragma: no cover - GPU-only module
    pytest.skip("no HIP device", allow_module_level=True)

from cosmosim import ForceAccumulator, TreePmCoordinator  # noqa: E402
from cosmosim import _lib as L  # noqa: E402
from test_oracle_pins import clustered  # noqa: E402

MESH = 32
COUNTERS = ("visited_nodes", "cutoff_pruned_nodes", "direct_pair_evaluations", "accepted_internal_multipoles",
            "accepted_leaves", "group_walk_groups", "group_walk_grouped_targets", "group_walk_target_group_size_max",
            "group_walk_fallback_targets", "group_walk_fallback_history", "group_walk_fallback_geometry",
            "group_walk_candidate_leaves", "group_walk_culled_leaves", "group_walk_candidate_particles",
            "group_walk_list_entries_max", "group_walk_status")


def gpu_adaptive(x, y, z, m, o, ad, previous=None, active=None, mesh=MESH):
    dx, dy, dz, dm = to_dev(x, y, z, m)
    na = x.size if active is None else len(active)
    acc = [empty_dev(na, torch.float64) for _ in range(3)]
    pm = [empty_dev(na, torch.float64) for _ in range(3)]
    sr = [empty_dev(x.size, torch.float64) for _ in range(3)]  # source-row indexed
    act = to_dev(np.asarray(active, np.uint32).view(np.int32))[0] if active is not None else None
    prev = to_dev(np.asarray(previous, np.float64))[0] if previous is not None else None
    with TreePmCoordinator(mesh) as coord:
        res = coord.solve_active_set(dx, dy, dz, dm, ForceAccumulator(tuple(acc), act, tuple(pm), tuple(sr), prev),
                                     o, want_diagnostics=True, adaptive_options=ad)
        torch.cuda.synchronize()
    host = lambda ts: [t.cpu().numpy() for t in ts]  # noqa: E731
    rows = np.arange(x.size) if active is None else np.asarray(active)
    return host(acc), host(pm), [a[rows] for a in host(sr)], res.diagnostics


def check_against_oracle(x, y, z, m, o, ad, previous=None, active=None, tol=1e-12):
    acc, pm, sr, c = gpu_adaptive(x, y, z, m, o, ad, previous, active)
    ref, rc = O.short_range_adaptive(x, y, z, m, o, ad, MESH, active=active, previous=previous)
    mismatched = {k: (c[k], rc[k]) for k in COUNTERS if c[k] != rc[k]}
    assert not mismatched, mismatched
    assert rel_l2(sr, ref) < tol
    for a, p, s in zip(acc, pm, sr):
        assert np.allclose(a, p + s, rtol=0, atol=1e-12 * np.abs(a).max())
    return sr, c


@pytest.fixture(scope="module")
def seam():
    return seam_clusters()


@pytest.mark.parametrize("size", [0, 32])
def test_seam_clusters_group_walk_matches_oracle(seam, size):
    x, y, z, m = seam
    o, ad = seam_cluster_options(group=True, group_size=size)
    _, c = check_against_oracle(x, y, z, m, o, ad, previous=np.ones(x.size))
    assert c["group_walk_status"] == 1 and c["group_walk_grouped_targets"] == x.size
    assert 0 < c["group_walk_culled_leaves"] < c["group_walk_candidate_leaves"]


def test_seam_clusters_per_target_adaptive_matches_oracle(seam):
    x, y, z, m = seam
    o, ad = seam_cluster_options(group=False)
    check_against_oracle(x, y, z, m, o, ad, previous=np.ones(x.size))


def test_missing_history_falls_back_to_per_target_walk(seam):
    x, y, z, m = seam
    o, ad = seam_cluster_options(group=True, group_size=32)
    sr, c = check_against_oracle(x, y, z, m, o, ad, previous=None)
    assert c["group_walk_fallback_history"] == x.size and c["group_walk_groups"] == 0
    o, ad = seam_cluster_options(group=False)
    sr2, _ = check_against_oracle(x, y, z, m, o, ad, previous=None)
    for a, b in zip(sr, sr2):
        assert np.array_equal(a, b)


@pytest.fixture(scope="module")
def dense():
    x, y, z, m = clustered(3000, 8, 1.0)
    prev = np.random.default_rng(2).uniform(2500.0, 10000.0, x.size)
    return x, y, z, m, prev


@pytest.mark.parametrize("com_cut", [False, True])
@pytest.mark.parametrize("radial", [False, True])
def test_accepting_group_walk_matches_oracle(dense, com_cut, radial):
    x, y, z, m, prev = dense
    o, ad = seam_cluster_options(max_angle=1.0, alpha=0.01, group=True, com_cut=com_cut, radial=radial)
    o.max_leaf_size = 4
    prev = prev.copy()
    prev[::97] = np.nan  # a few groups without history fall back
    _, c = check_against_oracle(x, y, z, m, o, ad, previous=prev)
    assert c["accepted_internal_multipoles"] > 0 and c["group_walk_fallback_history"] > 0


@pytest.mark.parametrize("estimator", [L.ESTIMATOR_RESIDUAL_DERIVATIVE, L.ESTIMATOR_RELATIVE_MONOPOLE])
@pytest.mark.parametrize("lookup", [False, True])
def test_per_target_adaptive_walk_matches_oracle(dense, estimator, lookup):
    x, y, z, m, prev = dense
    o, ad = seam_cluster_options(max_angle=1.0, alpha=0.01, group=False, estimator=estimator, lookup=lookup)
    o.max_leaf_size = 4
    prev = prev.copy()
    prev[::50] = np.nan  # geometric COM-distance fallback for these targets
    _, c = check_against_oracle(x, y, z, m, o, ad, previous=prev)
    assert c["accepted_internal_multipoles"] > 0


def test_active_subset_and_block_size(dense):
    x, y, z, m, prev = dense
    rng = np.random.default_rng(11)
    active = np.sort(rng.choice(x.size, 1100, replace=False))[::-1].copy()  # any slot order
    o, ad = seam_cluster_options(max_angle=1.0, alpha=0.01, group=True, group_size=16)
    o.residual_block_size = 32
    check_against_oracle(x, y, z, m, o, ad, previous=prev[active], active=active)


def test_groups_above_one_wave(dense):
    x, y, z, m, prev = dense
    o, ad = seam_cluster_options(max_angle=1.0, alpha=0.01, group=True, group_size=0)
    o.max_leaf_size = 200
    o.residual_block_size = 256
    _, c = check_against_oracle(x, y, z, m, o, ad, previous=prev)
    assert c["group_walk_target_group_size_max"] > 64


def test_inputs_outside_the_bounded_metric_fall_back(dense):
    x, y, z, m, prev = dense
    x = x.copy()
    x[:40] += 1.0  # the same periodic positions one box length up: the metric is unbounded near them
    o, ad = seam_cluster_options(max_angle=1.0, alpha=0.01, group=True)
    _, c = check_against_oracle(x, y, z, m, o, ad, previous=prev)
    assert c["group_walk_fallback_geometry"] > 0


def test_plain_solve_with_adaptive_policy_is_the_historyless_adaptive_walk(seam):
    x, y, z, m = seam
    o, _ = seam_cluster_options(group=False)
    ad = L.default_adaptive_options()
    acc, pm, sr, _ = gpu_adaptive(x, y, z, m, o, ad, previous=None)
    dx, dy, dz, dm = to_dev(x, y, z, m)
    out = [empty_dev(x.size, torch.float64) for _ in range(3)]
    with TreePmCoordinator(MESH) as coord:
        coord.solve_active_set(dx, dy, dz, dm, ForceAccumulator(tuple(out)), o)
        torch.cuda.synchronize()
    # the PM deposit's atomics leave round-off differences between two solves
    assert rel_l2(acc, [b.cpu().numpy() for b in out]) < 1e-12
