"""CPU pins of the oracle's adaptive acceptance, Gaussian table and group walk.

Checks that do not retype the oracle's code:
- the Hermite table of H(t) against the 50-digit closed form (the absolute
  budget 5e-11 of tests/unit/test_tree_pm_split_kernel.cpp:97-127) and its
  knot slopes against finite differences of that closed form;
- the shared radial coefficients against finite differences of the residual
  kernel in x = r^2, and the radial-coefficient node expansion against the
  reference expansion and its convergence order;
- the group walk of the reference's culling test
  (tests/integration/test_tree_pm_coupling_periodic.cpp:2140-2230): with a
  vanishing opening angle both walks reduce to the table direct sum within
  r_cut (brute force), the pair sets are identical, culling rejects some but
  not all leaves, and a missing history falls back to the per-target walk;
- every node the group list accepts satisfies, for every target of the group,
  the per-target relative-monopole criterion, the angle bound and lies outside
  r_cut-reaching AABB containment; every source inside r_cut of a target is
  in a listed leaf or an accepted node (no pair is lost);
- the exact-self-identity rule of the grouped pair kernel
  (tests/unit/test_tree_pm_split_kernel.cpp:474-519).
"""
from __future__ import annotations

import math

import mpmath
import numpy as np
import pytest

from helpers import options_struct, rel_l2, seam_cluster_options, seam_clusters, adaptive_struct
from oracle import oracle as O
from test_oracle_pins import clustered, ref_H, ref_H_vec

MESH = 32


def ref_residual(x, eps, a):
    return (x + eps * eps) ** -1.5 - ref_H_vec(x / (4 * a * a)) / (8 * a**3)


# ---------------------------------------------------------------------------
# Gaussian Hermite table (tree_pm_split_kernel.hpp:130-183)

def test_gauss_table_within_reference_budget():
    t = np.concatenate([np.linspace(0.0, 64.0, 131073)[::7], [np.nextafter(64.0, 0.0), 63.99, 1e-9, 0.015625]])
    coef, herm = O.gauss_table(t)
    ref = ref_H_vec(t)
    assert np.max(np.abs(coef - ref)) <= 5e-11
    assert np.max(np.abs(herm - ref)) <= 5e-11
    # the per-target coefficient is the series below t = 1/64
    small = t < 0.015625
    assert np.max(np.abs(coef[small] - ref[small]) / ref[small]) < 1e-14


def test_gauss_table_knots_are_closed_form_values_and_slopes():
    _, _, knots = O.gauss_table(np.array([1.0]), want_knots=True)
    step = 64.0 / 4096
    idx = list(range(0, 4097, 97))
    for i in idx:
        v = float(ref_H(i * step))
        assert abs(knots[0, i] - v) <= 1e-15 * max(v, 1.0) + 1e-16
    # slopes: step * dH/dt by central differences of the 50-digit closed form
    mpmath.mp.dps = 50
    for i in idx[1:]:
        t = mpmath.mpf(i) * step
        d = mpmath.mpf("1e-12")
        fd = (ref_H(t + d) - ref_H(t - d)) / (2 * d)
        assert abs(knots[1, i] - float(fd) * step) < 1e-13
    assert abs(knots[1, 0] - step * (-4.0 / (5.0 * math.sqrt(math.pi)))) < 1e-16


# ---------------------------------------------------------------------------
# Radial coefficients and the radial-coefficient node expansion
# (tree_pm_split_kernel.hpp:245-282, tree_pm_coupling.cpp:829-894)

@pytest.mark.parametrize("eps", [0.0, 0.01])
def test_radial_coefficients_are_derivatives_in_r_squared(eps):
    a = 0.4
    x = np.array([1e-4, 0.003, 0.01, 0.05, 0.2, 0.7, 1.9, 4.0]) * a * a
    got = O.residual_radial(x, eps, a)
    f0 = ref_residual(x, eps, a)
    assert np.max(np.abs(got[:, 0] - f0) / np.abs(f0)) < 1e-12
    h = 1e-4 * x
    fp = ref_residual(x + h, eps, a)
    fm = ref_residual(x - h, eps, a)
    d1 = (fp - fm) / (2 * h)
    d2 = (fp - 2 * f0 + fm) / (h * h)
    assert np.max(np.abs(got[:, 1] - 2 * d1) / np.abs(2 * d1)) < 1e-7
    assert np.max(np.abs(got[:, 2] - 4 * d2) / np.abs(4 * d2)) < 1e-4


@pytest.mark.parametrize("quad", [False, True])
def test_radial_node_expansion_equals_reference_expansion(quad):
    rng = np.random.default_rng(4)
    base = rng.normal(size=(3, 20))
    m = rng.uniform(0.5, 1.5, 20)
    for s in (0.05, 0.02):
        p = np.array([0.9, -0.3, 0.5])[:, None] + s * base
        ref = O.cluster_node_accel_radial(*p, m, (0.0, 0.0, 0.0), 1.0, 0.01, 0.6, quad, radial=False)
        got = O.cluster_node_accel_radial(*p, m, (0.0, 0.0, 0.0), 1.0, 0.01, 0.6, quad, radial=True)
        assert np.linalg.norm(got - ref) <= 1e-12 * np.linalg.norm(ref)


def test_radial_node_expansion_convergence_order():
    rng = np.random.default_rng(0)
    base = rng.normal(size=(3, 24))
    m = rng.uniform(0.5, 1.5, 24)
    a, eps, G = 0.8, 0.01, 1.0
    centre = np.array([1.3, 0.4, -0.2])
    errs = []
    for s in (0.04, 0.02, 0.01):
        p = centre[:, None] + s * base
        r2 = (p * p).sum(axis=0)
        ref = (G * m * ref_residual(r2, eps, a) * p).sum(axis=1)
        got = O.cluster_node_accel_radial(*p, m, (0.0, 0.0, 0.0), G, eps, a, True, radial=True)
        errs.append(np.linalg.norm(got - ref) / np.linalg.norm(ref))
    for e0, e1 in zip(errs, errs[1:]):
        assert 0.85 * 8 < e0 / e1 < 1.15 * 8


# ---------------------------------------------------------------------------
# Group walk (tree_pm_group_walk.hpp:78-341, tree_pm_coupling.cpp:4333-4546)

def brute_table_sum(x, y, z, m, eps, a, rcut, G=1.0):
    """Direct residual sum within r_cut over minimum images with the exact H."""
    n = x.size
    pos = np.stack([x, y, z])
    acc = np.zeros((3, n))
    pairs = 0
    for i in range(n):
        d = pos - pos[:, i:i + 1]
        d -= np.round(d)
        r2 = (d * d).sum(axis=0)
        sel = r2 <= rcut * rcut
        sel[i] = False
        pairs += int(sel.sum())
        nz = sel & (r2 > 0)
        acc[:, i] = (G * m[nz] * ref_residual(r2[nz], eps, a) * d[:, nz]).sum(axis=1)
    return acc, pairs


@pytest.fixture(scope="module")
def seam():
    return seam_clusters()


def test_group_walk_without_acceptance_equals_per_target_walk_and_direct_sum(seam):
    x, y, z, m = seam
    n = x.size
    prev = np.ones(n)
    o, ad = seam_cluster_options(group=False)
    ref, rc = O.short_range_adaptive(x, y, z, m, o, ad, MESH, previous=prev)
    assert rc["group_walk_status"] == 0 and rc["accepted_internal_multipoles"] == 0
    h = 1.0 / MESH
    brute, pairs = brute_table_sum(x, y, z, m, 0.002, 1.25 * h, 6.25 * h)
    assert rc["direct_pair_evaluations"] == pairs
    assert rel_l2(ref, brute) < 1e-9
    for size in (0, 32):
        o, ad = seam_cluster_options(group=True, group_size=size)
        got, c = O.short_range_adaptive(x, y, z, m, o, ad, MESH, previous=prev)
        assert c["group_walk_status"] == 1
        assert c["group_walk_grouped_targets"] == n and c["group_walk_fallback_targets"] == 0
        assert 0 < c["group_walk_culled_leaves"] < c["group_walk_candidate_leaves"]
        assert c["group_walk_candidate_particles"] >= c["direct_pair_evaluations"]
        assert c["direct_pair_evaluations"] == pairs
        if size:
            assert 8 < c["group_walk_target_group_size_max"] <= size
        else:
            assert c["group_walk_target_group_size_max"] <= 8
        assert rel_l2(got, ref) < 1e-9


def test_history_fallback_is_the_per_target_walk(seam):
    x, y, z, m = seam
    o, ad = seam_cluster_options(group=True, group_size=32)
    got, c = O.short_range_adaptive(x, y, z, m, o, ad, MESH, previous=None)
    assert c["group_walk_fallback_history"] == x.size == c["group_walk_fallback_targets"]
    assert c["group_walk_groups"] == 0
    o, ad = seam_cluster_options(group=False)
    ref, _ = O.short_range_adaptive(x, y, z, m, o, ad, MESH, previous=None)
    for g, r in zip(got, ref):
        assert np.array_equal(g, r)


def test_sub_floor_history_in_one_target_falls_back_its_group(seam):
    x, y, z, m = seam
    prev = np.ones(x.size)
    prev[5] = 0.0  # not above the floor
    o, ad = seam_cluster_options(group=True, group_size=0)
    _, c = O.short_range_adaptive(x, y, z, m, o, ad, MESH, previous=prev)
    assert 1 <= c["group_walk_fallback_history"] <= 8
    assert c["group_walk_grouped_targets"] + c["group_walk_fallback_targets"] == x.size


def _min_image(d):
    return d - np.round(d)


def test_group_list_acceptance_holds_for_every_target_and_loses_no_pair():
    box = 1.0
    x, y, z, m = clustered(3000, 8, box)
    o, ad = seam_cluster_options(max_angle=1.0, alpha=0.01)
    o.max_leaf_size = 4
    h = 1.0 / MESH
    rcut = 6.25 * h
    # one accepted-rich group: the targets of one leaf in the dense clump
    t = O.build_tree(x, y, z, m, leaf=4, quad=True)
    leaf_nodes = np.nonzero((t["child_index"] == 0xFFFFFFFF).all(axis=1))[0]
    rng = np.random.default_rng(1)
    accel_scale = 5000.0  # |a| large enough that the relative criterion accepts nodes
    checked_nodes = 0
    for leaf in rng.choice(leaf_nodes, 12, replace=False):
        rows = t["order"][t["begin"][leaf]:t["begin"][leaf] + t["count"][leaf]]
        prev = accel_scale * rng.uniform(0.5, 2.0, rows.size)
        nodes, leaves, (centre, half) = O.group_list(x, y, z, m, rows, prev, o, ad, MESH)
        T = O.group_tree(x.size)
        geo, mom = T["geometry"], T["moments"]
        P = np.stack([x[rows], y[rows], z[rows]])
        for nd in nodes:
            hh = geo[nd, 3]
            com = mom[nd, 1:4]
            d = _min_image(com[:, None] - P)
            r2 = (d * d).sum(axis=0)
            # relative monopole with each target's own history, and the angle bound
            assert np.all(mom[nd, 0] * 4 * hh * hh <= o.relative_force_tolerance * prev * r2 * r2)
            assert np.all(2 * hh < ad.maximum_opening_angle * np.sqrt(r2))
            # no target inside the node's cell, and the cell inside r_cut of every target
            dc = np.abs(_min_image(geo[nd, :3, None] - P))
            assert np.all((dc > hh).any(axis=0))
            far = dc + hh
            assert np.all((far * far).sum(axis=0) <= rcut * rcut)
            checked_nodes += 1
        # completeness: each source within r_cut of a target lies in a listed leaf
        # or under an accepted node
        covered = np.zeros(x.size, bool)
        for nd in list(nodes) + list(leaves):
            b, c = T["begin"][nd], T["count"][nd]
            covered[T["order"][b:b + c]] = True
        S = np.stack([x, y, z])
        for k in range(rows.size):
            d = _min_image(S - P[:, k:k + 1])
            near = (d * d).sum(axis=0) <= rcut * rcut
            assert covered[near].all()
    assert checked_nodes > 0


def test_grouped_pairs_exclude_only_the_exact_self():
    # tests/unit/test_tree_pm_split_kernel.cpp:474-519: a distinct source at the
    # target's position counts as a pair (zero force); the target itself does not.
    x = np.array([0.1, 0.1, 0.2])
    y = np.array([0.1, 0.1, 0.1])
    z = np.array([0.1, 0.1, 0.1])
    m = np.ones(3)
    o = options_struct(1.0, G=1.0, eps=0.01, leaf=16, asmth=1.25, rcut_cells=6.25, quad=True)
    o.acceptance_policy = 1
    ad = adaptive_struct(group=True)
    acc, c = O.short_range_adaptive(x, y, z, m, o, ad, 16, previous=np.ones(3))
    assert c["group_walk_groups"] == 1
    assert c["direct_pair_evaluations"] == 6
    assert np.all(np.isfinite(np.stack(acc)))
    a = 1.25 / 16
    expected = 1.0 * ref_residual(np.array([0.01]), 0.01, a)[0] * 0.1
    assert abs(acc[0][0] - expected) < 1e-9 * abs(expected)
    assert acc[1][0] == 0.0 and acc[2][0] == 0.0
