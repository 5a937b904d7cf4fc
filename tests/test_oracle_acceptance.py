"""Pins of the oracle's node acceptance, independent of its own code (CPU).

The strict policy (tree_pm_acceptance.hpp:54-66) accepts an internal node as
a quadrupole only when the target is outside its cell, the whole cell lies
inside r_cut, 2h/r < 0.08 and the opening criterion (2h [+ |c - COM|]) / r < theta
holds, with r the distance from the target to the node's COM. Here:

- known answers on one compact node: a probe target swept across each
  boundary (2h/r = 0.08, (2h + offset)/r = theta for both criteria, the
  farthest-corner distance = r_cut, a target inside the cell), accepted on the
  reference's side of each boundary only;
- a node-by-node validator on clustered sources: every accepted node meets
  every criterion, every opened internal node fails one, every pruned node's
  cell lies beyond r_cut, every summed leaf's does not, and the visit sets are
  closed (children of opened nodes visited, descendants of accepted or pruned
  nodes not). Distances are evaluated in 40-digit mpmath from the node
  geometry, not with the oracle's arithmetic; the validator is shown to
  reject an envelope constant of 0.07 or a criterion without the COM offset;
- residualAccuracyProxy (tree_pm_acceptance.hpp:14-51): its first and second
  radial derivatives of the residual kernel (r^2 + eps^2)^{-3/2} - L(r)/r^3
  against mpmath numerical derivatives of the closed form, on both branches
  of L (series below q = 1/8, erf above) and both branches of its maximum;
- the Ewald sum's Plummer correction (periodic_ewald_reference.cpp:300-330):
  the oracle's softened minus unsoftened Ewald force equals a brute-force
  image sum of (Plummer - Newton) pair forces over +-12 images.
"""
from __future__ import annotations

import math

import mpmath as mp
import numpy as np
import pytest

from helpers import options_struct
from oracle import oracle as O
from test_oracle_pins import clustered

mp.mp.dps = 40

BOX = 100.0
MESH = 64  # mesh spacing 1.5625: r_s = 1.95, r_cut = 8.79


def _cluster(k=40, seed=3, spread=0.1, centre=(50.2, 49.7, 50.4)):
    rng = np.random.default_rng(seed)
    pos = np.asarray(centre)[:, None] + rng.uniform(-spread, spread, (3, k))
    m = rng.uniform(0.5, 2.0, k)
    return pos[0].copy(), pos[1].copy(), pos[2].copy(), m


def _root(x, y, z, m, leaf=8):
    t = O.build_tree(x, y, z, m, leaf=leaf, quad=True)
    g, mo = t["geometry"][0], t["moments"][0]
    return [mp.mpf(float(v)) for v in g[:3]], mp.mpf(float(g[3])), [mp.mpf(float(v)) for v in mo[1:4]]


def _probe(src, o, target):
    return O.probe_walk(*src, o, MESH, [float(v) for v in target])[1]


def _along(com, direction, r):
    u = [mp.mpf(d) for d in direction]
    nrm = mp.sqrt(sum(v * v for v in u))
    return [com[a] + r * u[a] / nrm for a in range(3)]


def _offset(c, com):
    return mp.sqrt(sum((c[a] - com[a]) ** 2 for a in range(3)))


DIRECTION = (1.0, 0.31, -0.22)


@pytest.mark.parametrize("side,expect", [(1 + 1e-7, True), (1 - 1e-7, False)])
def test_envelope_boundary(side, expect):
    src = _cluster()
    c, h, com = _root(*src)
    o = options_struct(BOX, G=1.0, eps=0.01, theta=0.5, leaf=8, quad=True)
    r_b = 2 * h / mp.mpf("0.08")  # 2h/r < 0.08 <=> r > 25 h
    assert (2 * h + _offset(c, com)) / r_b < mp.mpf("0.5")  # the opening criterion holds on both sides
    tr = _probe(src, o, _along(com, DIRECTION, r_b * side))
    assert (0 in tr["accepted"]) == expect
    assert (0 in tr["opened"]) == (not expect)


@pytest.mark.parametrize("criterion", [0, 1])
@pytest.mark.parametrize("side,expect", [(1 + 1e-7, True), (1 - 1e-7, False)])
def test_opening_criterion_boundary(criterion, side, expect):
    src = _cluster()
    c, h, com = _root(*src)
    theta = mp.mpf("0.05")  # r_b ~ 4.4 < r_cut - 2h: the cell stays contained
    o = options_struct(BOX, G=1.0, eps=0.01, theta=float(theta), leaf=8, quad=True, criterion=criterion)
    lead = 2 * h if criterion == 0 else 2 * h + _offset(c, com)
    r_b = lead / theta
    assert 2 * h / r_b < mp.mpf("0.08")  # the envelope holds on both sides
    tr = _probe(src, o, _along(com, DIRECTION, r_b * side))
    assert (0 in tr["accepted"]) == expect


@pytest.mark.parametrize("side,expect", [(1 + 1e-7, True), (1 - 1e-7, False)])
def test_containment_boundary(side, expect):
    src = _cluster()
    c, h, com = _root(*src)
    target = _along(com, DIRECTION, 30 * h)
    far = mp.sqrt(sum((abs(c[a] - target[a]) + h) ** 2 for a in range(3)))  # farthest corner of the cell
    o = options_struct(BOX, G=1.0, eps=0.01, theta=0.5, leaf=8, quad=True)
    o.rcut_cells = float(far * side / (mp.mpf(BOX) / MESH))
    tr = _probe(src, o, target)
    assert (0 in tr["accepted"]) == expect


def test_target_inside_the_cell_never_accepts():
    src = _cluster()
    c, h, com = _root(*src)
    o = options_struct(BOX, G=1.0, eps=0.01, theta=10.0, leaf=8, quad=True)
    tr = _probe(src, o, [c[0] + 0.9 * h, c[1] - 0.5 * h, c[2]])
    assert 0 in tr["opened"] and tr["accepted"].size == 0


def test_monopoles_are_never_accepted():
    src = _cluster()
    c, h, com = _root(*src)
    o = options_struct(BOX, G=1.0, eps=0.01, theta=0.5, leaf=8, quad=False)
    tr = _probe(src, o, _along(com, DIRECTION, 60 * h))
    assert tr["accepted"].size == 0 and tr["leaves"].size > 0


# ---- node-by-node validator ---------------------------------------------------
def _mi(d, L):
    return d - L * mp.nint(d / L)


def _criteria(tree, n, target, L, theta, criterion, rcut, envelope="0.08", offset_term=True):
    """Each criterion of the strict acceptance as (value, bound): the node is
    accepted when every value < bound (inside: the target is in the cell)."""
    g, mo = tree["geometry"][n], tree["moments"][n]
    c = [mp.mpf(float(v)) for v in g[:3]]
    h = mp.mpf(float(g[3]))
    com = [mp.mpf(float(v)) for v in mo[1:4]]
    dc = [abs(_mi(c[a] - target[a], L)) for a in range(3)]
    inside = all(dc[a] <= h for a in range(3))
    far = mp.sqrt(sum((dc[a] + h) ** 2 for a in range(3)))
    r = mp.sqrt(sum(_mi(com[a] - target[a], L) ** 2 for a in range(3)))
    lead = 2 * h + (_offset(c, com) if (criterion != 0 and offset_term) else 0)
    return inside, [(far, rcut), (2 * h / r, mp.mpf(envelope)), (lead / r, theta)]


def _min_distance(tree, n, target, L):
    g = tree["geometry"][n]
    h = mp.mpf(float(g[3]))
    e = [max(mp.mpf(0), abs(_mi(mp.mpf(float(g[a])) - target[a], L)) - h) for a in range(3)]
    return mp.sqrt(sum(v * v for v in e))


def _validate(tree, trace, target, L, theta, criterion, rcut, **kw):
    """Violations of the strict policy in one traced walk; near-ties within
    1e-12 (relative) of a bound are skipped, since FP64 may decide them either way."""
    bad = []

    def tie(v, b):
        return abs(v - b) <= mp.mpf("1e-12") * abs(b)

    for n in trace["accepted"]:
        inside, crit = _criteria(tree, n, target, L, theta, criterion, rcut, **kw)
        if any(tie(v, b) for v, b in crit):
            continue
        if inside or not all(v < b for v, b in crit):
            bad.append(("accepted", int(n)))
    for n in trace["opened"]:
        inside, crit = _criteria(tree, n, target, L, theta, criterion, rcut, **kw)
        if any(tie(v, b) for v, b in crit):
            continue
        if not inside and all(v < b for v, b in crit):
            bad.append(("opened", int(n)))
    for n in trace["pruned"]:
        d = _min_distance(tree, n, target, L)
        if not tie(d, rcut) and not d > rcut:
            bad.append(("pruned", int(n)))
    for n in trace["leaves"]:
        d = _min_distance(tree, n, target, L)
        if not tie(d, rcut) and d > rcut:
            bad.append(("leaf", int(n)))
    return bad


def _closed(tree, trace):
    child = tree["child_index"]
    subtree = tree["subtree"].astype(np.int64)
    visited = np.concatenate([trace[k] for k in ("accepted", "opened", "leaves", "pruned")]).astype(np.int64)
    assert np.unique(visited).size == visited.size
    vis = set(visited.tolist())
    assert 0 in vis
    for n in trace["opened"]:
        for c in child[n]:
            if c != 0xFFFFFFFF:
                assert int(c) in vis
    for key in ("accepted", "pruned"):
        for n in trace[key]:
            inner = set(range(int(n) + 1, int(n) + int(subtree[n])))
            assert not (inner & vis)


@pytest.fixture(scope="module")
def clustered_walks():
    box = 10.0
    x, y, z, m = clustered(6000, 11, box)
    rng = np.random.default_rng(5)
    # probe targets inside the dense cluster, in the field and near the seam
    targets = [(x[i], y[i], z[i]) for i in rng.choice(x.size // 2, 10, replace=False)]
    targets += [tuple(rng.uniform(0, box, 3)) for _ in range(6)] + [(0.01, 5.0, 9.99), (9.995, 0.002, 3.0)]
    out = {}
    for theta, criterion in ((0.5, 1), (0.05, 1), (0.05, 0)):
        o = options_struct(box, G=1.0, eps=0.005, theta=theta, leaf=8, quad=True, criterion=criterion)
        o.rcut_cells = 12.0
        mesh = 16
        rcut = mp.mpf(12.0) * mp.mpf(box) / mesh
        walks = []
        for tgt in targets:
            _, tr = O.probe_walk(x, y, z, m, o, mesh, tgt)
            walks.append(([mp.mpf(float(v)) for v in tgt], tr))
        tree = O.group_tree(x.size)  # the tree the probe walks built (unwrapped frame)
        out[(theta, criterion)] = (tree, walks, mp.mpf(theta), criterion, rcut)
    return out, mp.mpf(box)


@pytest.mark.parametrize("setting", [(0.5, 1), (0.05, 1), (0.05, 0)])
def test_validator_accepts_the_oracle(clustered_walks, setting):
    data, L = clustered_walks
    tree, walks, theta, criterion, rcut = data[setting]
    accepted = opened_near = 0
    for target, tr in walks:
        assert _validate(tree, tr, target, L, theta, criterion, rcut) == []
        _closed(tree, tr)
        accepted += tr["accepted"].size
        for n in tr["opened"]:
            inside, crit = _criteria(tree, n, target, L, theta, criterion, rcut)
            if not inside and crit[0][0] < crit[0][1]:
                opened_near += 1
    # the walks exercise the boundaries: many accepted nodes, and opened nodes
    # that were contained and outside (rejected by the envelope or the MAC)
    assert accepted > 100 and opened_near > 50


def test_validator_rejects_plausible_mistakes(clustered_walks):
    data, L = clustered_walks
    tree, walks, theta, criterion, rcut = data[(0.5, 1)]
    # an envelope of 0.07 instead of 0.08: some accepted nodes violate it
    assert any(_validate(tree, tr, t, L, theta, criterion, rcut, envelope="0.07") for t, tr in walks)
    tree, walks, theta, criterion, rcut = data[(0.05, 1)]
    # the COM-distance criterion without its offset term: some opened nodes
    # would have passed it, so the validator flags them
    assert any(_validate(tree, tr, t, L, theta, criterion, rcut, offset_term=False) for t, tr in walks)


# ---- residualAccuracyProxy ------------------------------------------------------
def _residual(r, eps, a):
    q = r / (2 * a)
    lf = mp.erf(q) - 2 * q / mp.sqrt(mp.pi) * mp.exp(-q * q)
    return (r * r + eps * eps) ** mp.mpf(-1.5) - lf / r**3


@pytest.mark.parametrize("eps", [0.0, 0.03])
def test_residual_proxy_derivatives(eps):
    a, G, M = 0.4, 1.7, 2.5
    h_small, h_large = 1e-4, 0.05
    rs = np.concatenate([np.geomspace(0.01, 0.099, 6), np.geomspace(0.11, 4.0, 14)])  # q < 1/8 and above
    checked = {"width": 0, "derivative": 0}
    for r in rs:
        f1 = mp.diff(lambda v: _residual(v, mp.mpf(eps), mp.mpf(a)), mp.mpf(r), 1)
        f2 = mp.diff(lambda v: _residual(v, mp.mpf(eps), mp.mpf(a)), mp.mpf(r), 2)
        for h, offset in ((h_small, 0.0), (h_large, 0.02)):
            width = 2 * h
            rho = math.sqrt(3.0) * h + offset
            term_w = mp.mpf(width) ** 2 / mp.mpf(r) ** 4
            term_d = mp.mpf(0.5) * rho**2 * (3 * abs(f1) + abs(mp.mpf(r) * f2 - f1))
            expect = G * M * max(term_w, term_d)
            got = O.residual_proxy(h, offset, M, r * r, eps, a, G)[()]
            assert abs(got - float(expect)) <= 1e-9 * float(expect)
            checked["width" if term_w >= term_d else "derivative"] += 1
    assert checked["width"] > 0 and checked["derivative"] > 0


def test_residual_proxy_invalid_inputs_are_infinite():
    out = O.residual_proxy([0.1, 0.1], [0.0, 0.0], [1.0, 1.0], [0.0, np.inf], 0.01, 0.4, 1.0)
    assert np.all(np.isinf(out))


# ---- Ewald Plummer correction ---------------------------------------------------
def _plummer_correction(pos, m, box, eps, lim):
    ii = np.arange(-lim, lim + 1, dtype=np.float64)
    img = np.stack(np.meshgrid(ii, ii, ii, indexing="ij"), axis=-1).reshape(-1, 3) * box
    out = np.zeros((3, pos.shape[0]))
    for i in range(pos.shape[0]):
        for j in range(pos.shape[0]):
            d = pos[j] - pos[i]
            d -= box * np.round(d / box)
            s = d[None, :] + img
            if i == j:
                s = s[np.any(img != 0.0, axis=1)]
            r2 = np.sum(s * s, axis=1)
            f = m[j] * ((r2 + eps * eps) ** -1.5 - r2**-1.5)
            out[:, i] += np.sum(f[:, None] * s, axis=0)
    return out


def test_ewald_softening_correction_against_an_image_sum():
    """The oracle keeps the reference's default +-5 correction images
    (periodic_ewald_reference.hpp:56): it equals an independent image sum
    over the same range, and the truncation is below the reference's own
    convergence bound (relative L2 < 1e-5 against +-7 images,
    test_periodic_ewald_reference.cpp:149-191) against +-12 images."""
    box = 1.0
    pos = np.array([[0.1, 0.2, 0.3], [0.12, 0.21, 0.27], [0.8, 0.7, 0.05], [0.5, 0.95, 0.5]])
    m = np.array([1.0, 2.0, 0.5, 1.5])
    x, y, z = (pos[:, a].copy() for a in range(3))
    tg = np.arange(4, dtype=np.uint32)
    eps = 0.03
    soft = np.stack(O.ewald(x, y, z, m, tg, (box,) * 3, 1.0, 2.5, 3, 10, eps))
    hard = np.stack(O.ewald(x, y, z, m, tg, (box,) * 3, 1.0, 2.5, 3, 10, 0.0))
    corr = soft - hard
    ref5 = _plummer_correction(pos, m, box, eps, 5)
    ref12 = _plummer_correction(pos, m, box, eps, 12)
    assert np.max(np.abs(corr - ref5)) <= 1e-10 * np.max(np.abs(soft))
    assert np.sqrt(np.sum((ref5 - ref12) ** 2) / np.sum(soft**2)) < 1e-5
    # the correction is the near-field softening: the close pair dominates
    assert np.abs(ref5[:, 0]).max() > 100 * np.abs(ref5[:, 3]).max()
