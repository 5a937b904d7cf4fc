"""CPU pins of the oracle (the checker the GPU parity tests compare against).

Nothing here retypes the oracle's code: every check is a brute-force
restatement (bit-by-bit interleave, O(n^2) gap search, direct sums), a closed
form fixed by the mathematics (single Fourier modes, erf/series identities,
lattice symmetries), an invariant (mass conservation, deposit/interpolation
adjointness, Ewald alpha independence) or a known answer from the reference's
own tests (tests/golden/reference_fixtures.json, each entry citing its test).
"""
from __future__ import annotations

import json
import math
from pathlib import Path

import mpmath
import numpy as np
import pytest

from helpers import options_struct, rel_l2
from oracle import oracle as O

FIX = json.loads((Path(__file__).parent / "golden" / "reference_fixtures.json").read_text())
INVALID = 0xFFFFFFFF


# ---------------------------------------------------------------------------
# Morton keys and stable order (tree_ordering.cpp:73-173)

def brute_morton(x, y, z):
    pts = np.stack([x, y, z])
    lo = pts.min(axis=1)
    extent = max(float((pts.max(axis=1) - lo).max()), 1e-12)
    grid = float((1 << 21) - 1)
    q = []
    for a in range(3):
        v = np.clip((pts[a] - lo[a]) / extent, 0.0, 1.0) * grid
        q.append(np.floor(v + 0.5).astype(np.uint64))  # llround of a non-negative value
    keys = np.zeros(len(x), dtype=np.uint64)
    for bit in range(21):
        for a, shift in ((0, 2), (1, 1), (2, 0)):
            keys |= ((q[a] >> np.uint64(bit)) & np.uint64(1)) << np.uint64(3 * bit + shift)
    return keys


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_morton_keys_match_bitwise_interleave(seed):
    rng = np.random.default_rng(seed)
    n = 3000
    x, y, z = rng.uniform(-3.0, 7.0, (3, n))
    if seed == 1:  # heavy duplication and a flat axis
        x = np.round(x, 1)
        z[:] = 0.5
    keys = O.morton_keys(x, y, z)
    assert np.array_equal(keys, brute_morton(x, y, z))
    sk, order = O.morton_order(x, y, z)
    ref_order = np.argsort(keys, kind="stable")
    assert np.array_equal(order, ref_order)
    assert np.array_equal(sk, keys[ref_order])


def test_morton_extremes_and_degenerate_extent():
    x = np.array([0.0, 1.0, 0.5])
    keys = O.morton_keys(x, x, x)
    assert keys[0] == 0
    assert keys[1] == (1 << 63) - 1  # all 63 interleaved bits set
    # all points coincident: extent floor 1e-12, every key 0, order is identity
    same = np.full(17, 3.25)
    k, order = O.morton_order(same, same, same)
    assert np.all(k == 0) and np.array_equal(order, np.arange(17))


# ---------------------------------------------------------------------------
# Periodic unwrap (tree_pm_coupling.cpp:141-218)

def brute_unwrap(v, L):
    """O(n^2): the anchor is the point preceded by the largest circular gap;
    equal gaps go to the smaller anchor; points below the anchor gain +L."""
    best_gap, anchor = -1.0, None
    for c in v:
        below = v[v < c]
        pred = below.max() if below.size else v.max() - L
        gap = c - pred
        if gap > best_gap or (gap == best_gap and c < anchor):
            best_gap, anchor = gap, c
    out = np.where(v < anchor, v + L, v)
    return out, anchor


UNWRAP_CASES = {
    "seam_cluster": (np.array([0.01, 0.02, 9.97, 9.99, 0.03, 9.95]), 10.0),
    "interior_cluster": (np.array([4.0, 4.5, 5.0, 4.25]), 10.0),
    "equal_gaps_lattice": (np.arange(8) * 1.25, 10.0),
    "duplicates": (np.array([2.0, 2.0, 2.0, 7.5, 7.5]), 10.0),
    "single": (np.array([3.5]), 10.0),
    "two_half_box": (np.array([0.0, 5.0]), 10.0),
    "zero_and_near_L": (np.array([0.0, 9.999999, 0.25]), 10.0),
}


@pytest.mark.parametrize("name", sorted(UNWRAP_CASES))
def test_unwrap_matches_gap_search(name):
    v, L = UNWRAP_CASES[name]
    out, anchor = O.unwrap_axis(v, L)
    ref, ref_anchor = brute_unwrap(v, L)
    assert anchor == ref_anchor
    assert np.array_equal(out, ref)


def test_unwrap_random_small_sets():
    rng = np.random.default_rng(11)
    for _ in range(200):
        n = int(rng.integers(1, 12))
        L = float(rng.choice([1.0, 2.5, 100.0]))
        # dyadic grid values so several gaps tie exactly
        v = rng.integers(0, 64, n) * (L / 64.0)
        out, anchor = O.unwrap_axis(v, L)
        ref, ref_anchor = brute_unwrap(v, L)
        assert anchor == ref_anchor and np.array_equal(out, ref)


def test_unwrap_wraps_noncanonical_inputs():
    # dyadic values: any correct wrap into [0, L) is exact
    raw = np.array([-0.25, 1.0, 2.75, -1.5, 0.5])
    out, anchor = O.unwrap_axis(raw, 1.0, canonical=False)
    wrapped = np.array([0.75, 0.0, 0.75, 0.5, 0.5])
    ref, ref_anchor = brute_unwrap(wrapped, 1.0)
    assert anchor == ref_anchor and np.array_equal(out, ref)


# ---------------------------------------------------------------------------
# Octree build and multipoles (tree_gravity.cpp:343-524, 1081-1380)

def clustered(n, seed, box=10.0):
    rng = np.random.default_rng(seed)
    pos = rng.uniform(0, box, (3, n))
    k = n // 2
    pos[:, :k] = np.mod(rng.normal(0.3 * box, 0.02 * box, (3, k)), box)
    m = rng.uniform(0.5, 2.0, n)
    return pos[0].copy(), pos[1].copy(), pos[2].copy(), m


def validate_tree(x, y, z, t, leaf):
    """Structural invariants of the reference octree, checked independently of
    the builder: exact FP64 cube geometry, the x >= centre octant rule, stable
    octant partition of Morton order, the leaf rule and preorder layout."""
    n = x.size
    geo, begin, count, child, order, subtree = (t[k] for k in ("geometry", "begin", "count", "child_index",
                                                                "order", "subtree"))
    nn = geo.shape[0]
    lo = np.array([x.min(), y.min(), z.min()])
    hi = np.array([x.max(), y.max(), z.max()])
    ext = max(float((hi - lo).max()), 1e-10)
    assert np.array_equal(geo[0, :3], 0.5 * (lo + hi))
    assert geo[0, 3] == 0.5 * ext * (1.0 + 1e-8)
    assert begin[0] == 0 and count[0] == n
    assert np.array_equal(np.sort(order), np.arange(n))
    keys = brute_morton(x, y, z)
    rank = np.empty(n, np.int64)
    rank[np.argsort(keys, kind="stable")] = np.arange(n)
    pts = np.stack([x, y, z], axis=1)
    for node in range(nn):
        b, c = int(begin[node]), int(count[node])
        members = order[b:b + c]
        kids = child[node]
        cx, h = geo[node, :3], geo[node, 3]
        if np.all(kids == INVALID):
            assert c <= leaf or h <= 1e-12
            assert subtree[node] == 1
            assert np.all(np.diff(rank[members]) > 0)  # Morton order inside a leaf
            continue
        assert c > leaf and h > 1e-12
        octant = ((pts[members, 0] >= cx[0]) * 4 + (pts[members, 1] >= cx[1]) * 2 + (pts[members, 2] >= cx[2]))
        # members grouped by octant in octant order; together with the child
        # ranges below and Morton order inside leaves this fixes the order of
        # a stable partition completely
        assert np.all(np.diff(octant) >= 0)
        pos, nxt, total = b, node + 1, 1
        for o in range(8):
            k = int(kids[o])
            if k == INVALID:
                assert not np.any(octant == o)
                continue
            assert k == nxt  # preorder: children follow in octant order
            assert begin[k] == pos and count[k] == np.count_nonzero(octant == o) > 0
            sign = np.array([(o >> 2) & 1, (o >> 1) & 1, o & 1]) * 2 - 1
            assert np.array_equal(geo[k, :3], cx + sign * (0.5 * h)) and geo[k, 3] == 0.5 * h
            pos += int(count[k])
            nxt += int(subtree[k])
            total += int(subtree[k])
        assert pos == b + c and subtree[node] == total


def direct_moments(x, y, z, m, members):
    ld = np.longdouble
    px, py, pz, pm = (np.asarray(a[members], dtype=ld) for a in (x, y, z, m))
    M = pm.sum()
    c = np.array([(pm * px).sum() / M, (pm * py).sum() / M, (pm * pz).sum() / M])
    d = np.stack([px - c[0], py - c[1], pz - c[2]])
    d2 = (d * d).sum(axis=0)
    q = {}
    for (i, j), name in zip(((0, 0), (0, 1), (0, 2), (1, 1), (1, 2), (2, 2)), ("xx", "xy", "xz", "yy", "yz", "zz")):
        q[name] = (pm * (3 * d[i] * d[j] - (d2 if i == j else 0))).sum()
    return M, c, q, (pm * d2).sum()


@pytest.mark.parametrize("seed,leaf", [(3, 16), (4, 1), (5, 8)])
def test_tree_invariants_and_direct_moments(seed, leaf):
    x, y, z, m = clustered(1500, seed)
    t = O.build_tree(x, y, z, m, leaf=leaf, quad=True)
    validate_tree(x, y, z, t, leaf)
    mom = t["moments"]
    scale_len = t["geometry"][0, 3]
    for node in range(0, t["geometry"].shape[0], 7):
        members = t["order"][t["begin"][node]:t["begin"][node] + t["count"][node]]
        M, c, q, tr = direct_moments(x, y, z, m, members)
        assert abs(mom[node, 0] - float(M)) <= 1e-13 * float(M)
        assert np.allclose(mom[node, 1:4], np.asarray(c, dtype=float), rtol=0, atol=1e-13 * scale_len)
        qs = float(M) * scale_len**2
        got = dict(zip(("xx", "xy", "xz", "yy", "yz", "zz"), mom[node, 4:10]))
        for k in got:
            assert abs(got[k] - float(q[k])) <= 1e-11 * qs
        assert abs(mom[node, 10] - float(tr)) <= 1e-11 * qs


def test_multipole_accumulation_known_answer():
    f = FIX["multipole_accumulation"]
    t = O.build_tree(np.array(f["x"]), np.array(f["y"]), np.array(f["z"]), np.array(f["mass"]),
                     leaf=f["max_leaf_size"], quad=True)
    mom = t["moments"][0]
    assert abs(mom[0] - f["total_mass"]) < f["tolerance"]
    assert abs(mom[1] - f["com_x"]) < f["tolerance"]
    assert abs(mom[4] + mom[7] + mom[9]) < f["tolerance"]  # traceless
    assert abs(mom[4]) > 0.0


def test_tree_coincident_points_stop_at_half_floor():
    x = np.full(40, 1.0)
    t = O.build_tree(x, x, x, np.ones(40), leaf=4)
    validate_tree(x, x, x, t, 4)


# ---------------------------------------------------------------------------
# Split kernel (tree_pm_split_kernel.hpp:66-228)

def ref_H(t):
    """H(t) = 8 a^3 L(r)/r^3 at t = r^2/(4a^2), from its convergent series
    (t < 1) or the erf closed form, in 50-digit arithmetic."""
    mpmath.mp.dps = 50
    t = mpmath.mpf(t)
    c = 4 / mpmath.sqrt(mpmath.pi)
    if t < 1:
        s, term = mpmath.mpf(0), mpmath.mpf(1)
        for k in range(120):
            s += term / (2 * k + 3)
            term *= -t / (k + 1)
        return c * s
    q = mpmath.sqrt(t)
    return (mpmath.erf(q) - c / 2 * q * mpmath.exp(-t)) / (t * q)


def test_split_zero_and_tiny_radius_known_answers():
    f = FIX["split_zero_and_tiny_radius"]
    limit = 1.0 / (6.0 * math.sqrt(math.pi))
    for a in f["split_scales"]:
        S, Lr, _ = O.split_factors(np.array([0.0]), a, 0.0)
        assert S[0] == 1.0 and Lr[0] == 0.0
        scale = a**3
        lr0 = O.long_range_inv_r3(np.array([0.0]), a)[0]
        assert abs(lr0 * scale - limit) < f["limit_tolerance"] * 10  # double vs long double limit
        for ratio in f["radius_ratios"]:
            r = a * ratio
            q2 = (r / (2 * a)) ** 2
            expected = float(ref_H(q2)) / 8.0
            got = O.long_range_inv_r3(np.array([r]), a)[0]
            assert math.isfinite(got)
            assert abs(got * scale - expected) < f["coefficient_relative_tolerance"] * expected
        for er in f["softening_ratios"]:
            eps = a * er
            zero = O.residual_inv_r3(np.array([0.0]), eps, a)[0]
            assert abs((zero - (eps**-3 - limit / scale)) * scale) < f["coincident_tolerance"]
            r = a * 1e-12
            tiny = O.residual_inv_r3(np.array([r * r]), eps, a)[0]
            assert abs((tiny - zero) * scale) < f["coincident_tolerance"]


def test_split_long_range_coefficient_across_interval():
    # H(t) over the table interval (tests/unit/test_tree_pm_split_kernel.cpp:97-127),
    # analytic branch only: absolute budget 5e-11 on H.
    a = 0.37
    for t in np.linspace(0.0, 64.0, 801):
        r = 2 * a * math.sqrt(t)
        got = O.long_range_inv_r3(np.array([r]), a)[0] * 8 * a**3
        assert abs(got - float(ref_H(t))) <= 5e-11


def test_split_complementarity_and_composition():
    f = FIX["split_complementarity"]
    a = f["asmth_cells"] * f["mesh_spacing"]
    S, Lr, _ = O.split_factors(np.array(f["radii"]), a, 0.0)
    assert np.all(S >= 0) and np.all(S <= 1 + 1e-12) and np.all(Lr >= -1e-12) and np.all(Lr <= 1)
    assert np.all(np.abs(S + Lr - 1) < f["tolerance"])
    for r in f["radii"]:  # S(r) against the erfc closed form
        q = mpmath.mpf(r) / (2 * a)
        s_ref = mpmath.erfc(q) + 2 / mpmath.sqrt(mpmath.pi) * q * mpmath.exp(-q * q)
        assert abs(O.split_factors(np.array([r]), a, 0.0)[0][0] - float(s_ref)) < 1e-15
    g = FIX["split_composition"]
    a = g["split_scale"]
    for er in g["softening_over_split"]:
        eps = er * a
        for r in g["radii"]:
            _, Lr, R = O.split_factors(np.array([r]), a, eps)
            pm_long = Lr[0] / r**3
            expected = (r * r + eps * eps) ** -1.5
            assert abs(R[0] + pm_long - expected) / expected < g["tolerance"]


# ---------------------------------------------------------------------------
# Short-range walk (tree_pm_coupling.cpp:4026-4272): with monopoles the strict
# acceptance never accepts an internal node, so the walk is the exact direct sum
# of the residual within r_cut over minimum images.

def ref_H_vec(t):
    """Vectorised H(t): 40-term series below t = 1, erf closed form above."""
    from scipy.special import erf

    t = np.asarray(t, dtype=float)
    out = np.empty_like(t)
    lo = t < 1.0
    s = np.zeros(np.count_nonzero(lo))
    term = np.ones_like(s)
    for k in range(40):
        s += term / (2 * k + 3)
        term *= -t[lo] / (k + 1)
    out[lo] = 4 / math.sqrt(math.pi) * s
    q = np.sqrt(t[~lo])
    out[~lo] = (erf(q) - 2 / math.sqrt(math.pi) * q * np.exp(-t[~lo])) / (t[~lo] * q)
    return out


def test_vectorised_H_matches_50_digit_reference():
    t = np.concatenate([np.linspace(0, 2, 41), np.linspace(2, 64, 32)])
    ref = np.array([float(ref_H(v)) for v in t])
    assert np.max(np.abs(ref_H_vec(t) - ref) / ref) < 1e-14


def ref_residual(r2, eps, a):
    return (r2 + eps * eps) ** -1.5 - ref_H_vec(r2 / (4 * a * a)) / (8 * a**3)


def brute_short_range(x, y, z, m, box, G, eps, a, rcut):
    n = x.size
    acc = np.zeros((3, n))
    pairs = 0
    pos = np.stack([x, y, z])
    for i in range(n):
        d = pos - pos[:, i:i + 1]
        d -= box * np.round(d / box)
        r2 = (d * d).sum(axis=0)
        sel = (r2 <= rcut * rcut)
        sel[i] = False
        pairs += int(sel.sum())
        nz = sel & (r2 > 0)
        f = G * m[nz] * ref_residual(r2[nz], eps, a)
        acc[:, i] = (f * d[:, nz]).sum(axis=1)
    return acc, pairs


@pytest.mark.parametrize("seed", [21, 22])
def test_strict_monopole_walk_equals_direct_sum(seed):
    box = 10.0
    x, y, z, m = clustered(400, seed, box)
    mesh = 16
    opts = options_struct(box, G=1.3, eps=0.02, quad=False, leaf=4)
    a = 1.25 * box / mesh
    rcut = 4.5 * 1.25 * box / mesh
    acc, counters, _ = O.short_range(x, y, z, m, opts, mesh)
    ref, pairs = brute_short_range(x, y, z, m, box, 1.3, 0.02, a, rcut)
    assert counters["multipoles"] == 0
    assert counters["pairs"] == pairs
    assert rel_l2(acc, ref) < 1e-12


def test_strict_quadrupole_walk_close_to_direct_sum():
    box = 10.0
    x, y, z, m = clustered(1200, 31, box)
    mesh = 8  # wide cutoff: distant compact nodes pass 2h/r < 0.08
    opts = options_struct(box, G=1.0, eps=0.01, quad=True, leaf=2)
    a = 1.25 * box / mesh
    acc, counters, _ = O.short_range(x, y, z, m, opts, mesh)
    ref, _ = brute_short_range(x, y, z, m, box, 1.0, 0.01, a, 4.5 * a)
    assert counters["multipoles"] > 0
    # accepted nodes satisfy 2h/r < 0.08: the third-order truncation of the
    # quadrupole expansion bounds the error well below 1e-4
    assert rel_l2(acc, ref) < 1e-4


@pytest.mark.parametrize("quad,order", [(False, 2), (True, 3)])
def test_accepted_node_kernel_convergence_order(quad, order):
    # Shrinking a source cluster by 2 must shrink the error of its node
    # expansion (residual kernel, softened) by 2^2 for a monopole and 2^3 with
    # the quadrupole term: a wrong sign or factor in any quadrupole term leaves
    # a second-order error.
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
        got = O.cluster_node_accel(*p, m, (0.0, 0.0, 0.0), G, eps, a, quad)
        errs.append(np.linalg.norm(got - ref) / np.linalg.norm(ref))
    for e0, e1 in zip(errs, errs[1:]):
        assert 0.85 * 2**order < e0 / e1 < 1.15 * 2**order


def test_walk_translation_invariance_across_seam():
    # A compact cluster straddling the x seam gives the same short-range forces
    # as the same cluster moved to the middle of the box
    # (tests/integration/test_tree_pm_coupling_periodic.cpp:630-720).
    box = 1.0
    rng = np.random.default_rng(7)
    off = rng.uniform(-0.02, 0.02, (3, 12))
    seam = np.mod(off + np.array([[0.0], [0.5], [0.5]]), box)
    mid = off + 0.5
    opts = options_struct((1.0, 1.0, 1.0), G=1.0, eps=0.001, quad=True, leaf=2)
    a1, _, _ = O.short_range(*seam, np.ones(12), opts, 8)
    a2, _, _ = O.short_range(*mid, np.ones(12), opts, 8)
    assert rel_l2(a1, a2) < 1e-10


# ---------------------------------------------------------------------------
# Particle-mesh part (pm_solver.cpp)

def test_pm_mass_conservation_known_answers():
    for key, scheme in (("pm_cic_mass_conservation", 0), ("pm_tsc_mass_conservation_wrap", 1)):
        f = FIX[key]
        mesh = tuple(f["mesh"])
        opts = options_struct(f["box"], assignment=scheme)
        m = np.array(f["mass"])
        _, dens = O.pm_accel(np.array(f["x"]), np.array(f["y"]), np.array(f["z"]), m, opts, mesh, want_density=True)
        cell_volume = f["box"] ** 3 / np.prod(mesh)
        assert abs(dens.sum() * cell_volume - m.sum()) < f["tolerance"]


def test_pm_cic_weights_brute_force():
    # one particle at grid position (1.25, 2.5, 3.75) on an 8^3 mesh of a unit box
    opts = options_struct(1.0)
    g = np.array([1.25, 2.5, 3.75])
    _, dens = O.pm_accel(*(g[:, None] / 8.0), np.array([1.0]), opts, 8, want_density=True)
    dens = dens.reshape(8, 8, 8) / 8**3  # back to mass per cell
    expected = np.zeros((8, 8, 8))
    for i, wi in ((1, 0.75), (2, 0.25)):
        for j, wj in ((2, 0.5), (3, 0.5)):
            for k, wk in ((3, 0.25), (4, 0.75)):
                expected[i, j, k] = wi * wj * wk
    assert np.allclose(dens, expected, atol=1e-15)


def mode_density(mesh, fn):
    nx, ny, nz = mesh
    ix = np.arange(nx)[:, None, None]
    iy = np.arange(ny)[None, :, None]
    iz = np.arange(nz)[None, None, :]
    return np.broadcast_to(fn(ix, iy, iz), mesh).astype(float).copy()


@pytest.mark.parametrize("split,deconv", [(0.0, 0), (0.07, 0), (0.07, 1)])
def test_pm_single_mode_closed_form(split, deconv):
    f = FIX["pm_analytic_mode"]
    mesh = tuple(f["mesh"])
    L, G, A = f["box"], f["G"], f["amplitude"]
    k = 2 * math.pi / L
    x = lambda ix: (ix + 0.5) / mesh[0] * L  # noqa: E731
    rho = mode_density(mesh, lambda ix, iy, iz: A * np.sin(k * x(ix)) + 0 * iy + 0 * iz)
    opts = options_struct(L, G=G, deconv=deconv)
    opts.scale_factor = f["scale_factor"]
    fx, fy, fz = O.pm_mesh(rho, opts, mesh, split)
    amp = 4 * math.pi * G * A / k * math.exp(-(k * split) ** 2)
    if deconv:
        amp /= (math.sin(0.5 * k * L / mesh[0]) / (0.5 * k * L / mesh[0])) ** 4  # CIC window squared
    expected = mode_density(mesh, lambda ix, iy, iz: amp * np.cos(k * x(ix)) + 0 * iy + 0 * iz)
    assert np.max(np.abs(fx - expected)) < 1e-12 * amp
    assert np.max(np.abs(fy)) < 1e-12 * amp and np.max(np.abs(fz)) < 1e-12 * amp
    cos_sim = np.sum(fx * expected) / np.sqrt(np.sum(fx**2) * np.sum(expected**2))
    assert cos_sim > f["reference_cosine_similarity_min"]


def test_pm_nyquist_contract():
    f = FIX["pm_nyquist_contract"]
    mesh = tuple(f["mesh"])
    opts = options_struct(f["box"], G=f["G"])
    nyq = mode_density(mesh, lambda ix, iy, iz: np.where(ix % 2 == 0, 1.0, -1.0) + 0 * iy + 0 * iz)
    fx, _, _ = O.pm_mesh(nyq, opts, mesh, 0.0)
    assert np.max(np.abs(fx)) < f["pure_nyquist_force_x_max"]
    mixed = mode_density(mesh, lambda ix, iy, iz: np.where(ix % 2 == 0, 1.0, -1.0)
                         * np.sin(2 * np.pi * iy / mesh[1]) + 0 * iz)
    fx, fy, _ = O.pm_mesh(mixed, opts, mesh, 0.0)
    assert np.max(np.abs(fy)) > f["mixed_force_y_min"]
    assert np.max(np.abs(fx)) < 1e-10 * np.max(np.abs(fy)) + 1e-10
    near = mode_density(mesh, lambda ix, iy, iz: np.sin(2 * np.pi * (mesh[0] // 2 - 1) * ix / mesh[0])
                        + 0 * iy + 0 * iz)
    fx, _, _ = O.pm_mesh(near, opts, mesh, 0.0)
    assert np.max(np.abs(fx)) > f["near_nyquist_force_x_min"]


def test_pm_uniform_density_has_no_force():
    opts = options_struct(1.0, G=1.0)
    fx, fy, fz = O.pm_mesh(np.full((8, 8, 8), 3.0), opts, (8, 8, 8), 0.1)
    assert max(np.abs(fx).max(), np.abs(fy).max(), np.abs(fz).max()) < 1e-12


def test_pm_mesh_fft_against_numpy():
    # numpy's FFT pins the oracle's own transforms, index order and k layout
    # on a rectangular mesh/box; the operator itself is pinned by the closed
    # forms above.
    rng = np.random.default_rng(3)
    mesh = (8, 16, 4)
    box = (1.0, 2.0, 0.5)
    rho = rng.normal(size=mesh)
    split, G = 0.09, 0.7
    for deconv, scheme in ((0, 0), (1, 0), (1, 1)):
        opts = options_struct(box, G=G, deconv=deconv, assignment=scheme)
        got = O.pm_mesh(rho, opts, mesh, split)
        spec = np.fft.fftn(rho - rho.mean())
        ks = [2 * np.pi * np.fft.fftfreq(n, d=L / n) for n, L in zip(mesh, box)]
        KX, KY, KZ = np.meshgrid(*ks, indexing="ij")
        k2 = KX**2 + KY**2 + KZ**2
        k2[0, 0, 0] = 1.0
        green = -4 * np.pi * G / k2 * np.exp(-k2 * split**2)
        if deconv:
            p = 2 if scheme == 0 else 3
            W = np.ones(mesh)
            for K, n, L, sl in ((KX, mesh[0], box[0], 0), (KY, mesh[1], box[1], 1), (KZ, mesh[2], box[2], 2)):
                W = W * np.sinc(K * L / n / (2 * np.pi)) ** p
            green = green / np.maximum(W * W, 1e-12)
        green[0, 0, 0] = 0.0
        phi = spec * green
        for K, n, ax, g in ((KX, mesh[0], 0, got[0]), (KY, mesh[1], 1, got[1]), (KZ, mesh[2], 2, got[2])):
            Kd = K.copy()
            idx = [slice(None)] * 3
            idx[ax] = n // 2
            Kd[tuple(idx)] = 0.0  # zero-Nyquist first derivative
            ref = np.fft.ifftn(-1j * Kd * phi).real
            assert np.max(np.abs(g - ref)) < 1e-12 * np.max(np.abs(ref))


@pytest.mark.parametrize("scheme", [0, 1])
def test_pm_interpolation_is_adjoint_of_deposit(scheme):
    rng = np.random.default_rng(9 + scheme)
    mesh = (8, 4, 16)
    box = (2.0, 1.0, 4.0)
    n = 300
    x, y, z = (rng.uniform(0, L, n) for L in box)
    m = rng.uniform(0.5, 1.5, n)
    opts = options_struct(box, assignment=scheme)
    _, dens = O.pm_accel(x, y, z, m, opts, mesh, want_density=True)
    fields = [rng.normal(size=mesh) for _ in range(3)]
    got = O.pm_interpolate(fields, x, y, z, opts, mesh)
    cell_volume = np.prod(box) / np.prod(mesh)
    for a in range(3):
        lhs = np.sum(m * got[a])
        rhs = cell_volume * np.sum(dens.reshape(mesh) * fields[a])
        assert abs(lhs - rhs) < 1e-12 * np.sum(np.abs(m * got[a]))


@pytest.mark.parametrize("scheme", [0, 1])
def test_pm_interpolation_reproduces_linear_field(scheme):
    mesh = (16, 8, 8)
    opts = options_struct(1.0, assignment=scheme)
    lin = mode_density(mesh, lambda ix, iy, iz: ix + 0.0 * iy + 0.0 * iz)
    g = np.array([2.0, 2.3, 5.5, 7.75, 12.125])  # grid positions away from the wrap
    x = g / mesh[0]
    got = O.pm_interpolate([lin, lin * 0, lin * 0], x, np.full(5, 0.3), np.full(5, 0.6), opts, mesh)
    assert np.allclose(got[0], g, atol=1e-13)


def test_pm_momentum_conservation():
    rng = np.random.default_rng(12)
    x, y, z = rng.uniform(0, 1, (3, 500))
    m = rng.uniform(0.5, 1.5, 500)
    opts = options_struct(1.0, G=1.0)
    acc = O.pm_accel(x, y, z, m, opts, 16)
    net = [abs(np.sum(m * a)) for a in acc]
    scale = np.sum(m * np.sqrt(sum(a * a for a in acc)))
    assert max(net) < 1e-12 * scale


# ---------------------------------------------------------------------------
# FP64 Ewald (tests/support/periodic_ewald_reference.cpp semantics)

def ewald_at_sources(f, box, G, alpha, images, modes, shift=None, eps=0.0):
    x, y, z = (np.array(f[k], dtype=float) for k in ("x", "y", "z"))
    if shift is not None:
        s = np.array(shift, dtype=float)
        x, y, z = x + s[:, 0] * box[0], y + s[:, 1] * box[1], z + s[:, 2] * box[2]
    m = np.array(f["mass"])
    return np.stack(O.ewald(x, y, z, m, np.arange(m.size), box, G, alpha, images, modes, eps))


def test_ewald_convergence_and_alpha_independence():
    f = FIX["ewald_convergence"]
    box = tuple(f["box"])
    lv = {k: ewald_at_sources(f, box, f["G"], f["alpha"], *v) for k, v in f["levels"].items()}
    alt = ewald_at_sources(f, box, f["G"], f["alternate_alpha"], f["levels"]["stricter"][0], f["alternate_modes"])
    l2m = rel_l2(lv["loose"], lv["medium"])
    m2s = rel_l2(lv["medium"], lv["strict"])
    s2s = rel_l2(lv["strict"], lv["stricter"])
    assert l2m > f["loose_to_medium_min"]
    assert m2s < f["contraction"] * l2m
    assert s2s < f["contraction"] * m2s
    assert s2s < f["strict_to_stricter_max"]
    assert rel_l2(lv["stricter"], alt) < f["alpha_independence_max"]


def test_ewald_integer_box_translation_and_net_force():
    f = FIX["ewald_convergence"]
    box = tuple(f["box"])
    images, modes = f["levels"]["strict"]
    ref = ewald_at_sources(f, box, f["G"], f["alpha"], images, modes)
    moved = ewald_at_sources(f, box, f["G"], f["alpha"], images, modes, shift=FIX["ewald_translation_shifts"]["shifts"])
    assert rel_l2(moved, ref) < FIX["ewald_translation_shifts"]["tolerance"]
    m = np.array(f["mass"])
    net = np.abs((m * ref).sum(axis=1)).max()
    assert net <= FIX["ewald_net_force"]["tolerance"] * max(np.sum(m * np.linalg.norm(ref, axis=0)), 1.0)


def test_ewald_self_force_symmetry_and_sign():
    f = FIX["ewald_self_and_symmetry"]
    box = tuple(f["box"])
    s = f["single"]
    a = O.ewald(np.array([s["x"]]), np.array([s["y"]]), np.array([s["z"]]), np.array([s["mass"]]),
                np.array([0]), box, f["G"], f["alpha"], f["real_images"], f["modes"])
    assert np.linalg.norm([v[0] for v in a]) < f["self_tolerance"]
    # target at the centre between two equal sources: a zero-mass tracer
    p = f["pair"]
    c = f["centre"]
    xs = np.array(p["x"] + [c[0]]), np.array(p["y"] + [c[1]]), np.array(p["z"] + [c[2]])
    a = O.ewald(*xs, np.array(p["mass"] + [0.0]), np.array([2]), box, f["G"], f["alpha"], f["real_images"],
                f["modes"])
    assert np.linalg.norm([v[0] for v in a]) < f["symmetric_tolerance"]
    g = FIX["ewald_g_normalisation"]
    sp, tp = g["source_position"], g["target_position"]
    xs = np.array([sp[0], tp[0]]), np.array([sp[1], tp[1]]), np.array([sp[2], tp[2]])
    unit = O.ewald(*xs, np.array([g["source_mass"], 0.0]), np.array([1]), tuple(g["box"]), 1.0, g["alpha"],
                   g["real_images"], g["modes"])
    scaled = O.ewald(*xs, np.array([g["source_mass"], 0.0]), np.array([1]), tuple(g["box"]), g["G_scaled"],
                     g["alpha"], g["real_images"], g["modes"])
    u = np.array([v[0] for v in unit])
    s_ = np.array([v[0] for v in scaled])
    assert u[0] > 0.0  # attraction towards the source at larger x
    assert np.linalg.norm(s_ - g["G_scaled"] * u) / np.linalg.norm(g["G_scaled"] * u) < g["tolerance"]


def test_ewald_newtonian_limit_for_close_pair():
    # separation << box: the periodic force tends to G m / r^2 plus the
    # O(r / L^3) lattice correction.
    box = (10.0, 10.0, 10.0)
    r = 1e-3
    a = O.ewald(np.array([5.0, 5.0 + r]), np.array([5.0, 5.0]), np.array([5.0, 5.0]), np.array([1.0, 0.0]),
                np.array([1]), box, 1.0, 0.35, 3, 10)
    assert abs(-a[0][0] / (1.0 / r**2) - 1.0) < 1e-6


# ---------------------------------------------------------------------------
# Full TreePM on the reference's 48-particle periodic fixture

def fixture48():
    i = np.arange(48, dtype=float)
    x = np.fmod((17.0 * i + 3.0) * 0.017, 1.0)
    y = np.fmod((29.0 * i + 5.0) * 0.013, 1.0)
    z = np.fmod((41.0 * i + 7.0) * 0.011, 1.0)
    m = 0.8 + 0.01 * (np.arange(48) % 11)
    return x, y, z, m


def test_treepm_periodic_48_against_direct_and_ewald():
    f = FIX["treepm_periodic_48"]
    x, y, z, m = fixture48()
    opts = options_struct(1.0, G=1.0, eps=f["softening"], theta=f["theta"], leaf=f["leaf"],
                          rcut_cells=f["rcut_cells"], asmth=f["asmth_cells"], deconv=f["window_deconvolution"])
    acc = np.stack(O.solve(x, y, z, m, opts, f["mesh"]))
    # minimum-image softened direct sum, as the reference test's comparison
    pos = np.stack([x, y, z])
    direct = np.zeros((3, 48))
    for i in range(48):
        d = pos - pos[:, i:i + 1]
        d -= np.round(d)
        r2 = (d * d).sum(axis=0) + f["softening"] ** 2
        w = m * r2**-1.5
        w[i] = 0.0
        direct[:, i] = (w * d).sum(axis=1)
    assert rel_l2(acc, direct) < f["max_rel_l2_vs_min_image_direct"]
    ew = np.stack(O.ewald(x, y, z, m, np.arange(48), (1.0, 1.0, 1.0), 1.0, 5.0, 2, 12, f["softening"]))
    assert rel_l2(acc, ew) < 2e-2
    _, counters, _ = O.short_range(x, y, z, m, opts, f["mesh"])
    assert counters["skipped_pairs"] > 0
