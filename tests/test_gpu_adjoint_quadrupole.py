"""The force VJP with quadrupole moments (multipole_order 2) against autograd.

There is no reference counterpart (the reference has no adjoint). The check
is independent of the hand-written adjoint: a PyTorch FP64 model restates the
strict walk's force from the oracle's own decisions. For every target the
oracle's probe walk (tree_pm_coupling.cpp:4026-4272 restated) lists the leaves
it sums and the internal nodes it accepts; the model then writes the force as
  PM(x)_i + sum over the summed leaves' particles k of G m_k F(r^2) r (r <= r_cut)
          + sum over accepted nodes N of the monopole + quadrupole expansion
            (tree_pm_coupling.cpp:743-827) about the node's centre of mass,
with the node's mass, centre of mass and quadrupole computed from the
particle positions in torch, so torch.autograd differentiates through them.
The GPU VJP must equal autograd of sum_i lam_i . a_i to 1e-11 relative L2
(measured 6e-16), on a clustered set where 9213 (target, node) pairs accept a
node and the accepted-node terms move the VJP by 3.9e-5; and directional
central differences of the GPU forces must agree to 1e-5 (median of four
random directions).
"""
from __future__ import annotations

import math

import numpy as np
import pytest

from helpers import options_struct, rel_l2

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
if not torch.cuda.is_available():  # pragma: no cover - GPU-only module
    pytest.skip("no HIP device", allow_module_level=True)

import adjoint_reference as R  # noqa: E402
from cosmosim import ForceAccumulator, TreePmCoordinator  # noqa: E402
from cosmosim import _lib as L  # noqa: E402
from oracle import oracle as O  # noqa: E402

INV_SQRT_PI = 1.0 / math.sqrt(math.pi)
BOX, MESH, EPS, LEAF = 10.0, 32, 0.01, 4
G = 43.0091


def clustered_set(seed=7):
    """Six tight clumps about one r_cut apart plus a uniform background, away
    from the box faces (so the tree's unwrapped frame is the input frame)."""
    rng = np.random.default_rng(seed)
    centres = np.array([[4.0, 4.0, 4.0], [5.1, 4.1, 4.0], [4.2, 5.2, 4.3], [4.1, 4.3, 5.2], [5.3, 5.2, 5.1],
                        [6.2, 4.6, 4.8]])
    parts = [c + rng.normal(scale=0.012, size=(150, 3)) for c in centres]
    parts.append(rng.uniform(2.5, 7.5, size=(300, 3)))
    x = np.concatenate(parts).T.copy()
    m = rng.uniform(0.5, 1.5, size=x.shape[1])
    return x, m


def short_factor(r, a):
    q = 0.5 * r / a
    return torch.special.erfc(q) + 2.0 * INV_SQRT_PI * q * torch.exp(-q * q)


def node_accel(mass, q, d, a, eps):
    """monopolePlusQuadrupoleAccelPeriodicReference for offsets d (3, k) =
    COM - target, node masses (k,) and stored quadrupoles q (7, k): qxx, qxy,
    qxz, qyy, qyz, qzz, trace."""
    dx, dy, dz = d
    r2 = dx * dx + dy * dy + dz * dz
    safe_r2 = torch.clamp(r2, min=1e-30)
    r = torch.sqrt(safe_r2)
    denom = torch.clamp(r2 + eps * eps, min=1e-30)
    soft = 1.0 / (denom * torch.sqrt(denom))
    split = short_factor(r, a)
    newton = 1.0 / (safe_r2 * r)
    radial = soft + (split - 1.0) * newton
    c = [G * mass * radial * dx, G * mass * radial * dy, G * mass * radial * dz]
    qxx, qxy, qxz, qyy, qyz, qzz, tr = q
    mxx, mxy, mxz = (qxx + tr) / 3.0, qxy / 3.0, qxz / 3.0
    myy, myz, mzz = (qyy + tr) / 3.0, qyz / 3.0, (qzz + tr) / 3.0
    mdx = mxx * dx + mxy * dy + mxz * dz
    mdy = mxy * dx + myy * dy + myz * dz
    mdz = mxz * dx + myz * dy + mzz * dz
    dmd = dx * mdx + dy * mdy + dz * mdz
    a2 = a * a
    ex = torch.exp(-r2 / (4.0 * a2))
    s1 = -0.5 * INV_SQRT_PI * r2 * ex / (a * a2)
    s2 = INV_SQRT_PI * ex * (-r / (a * a2) + 0.25 * r * r2 / (a * a2 * a2))
    f1 = -3.0 * r / (denom * denom * torch.sqrt(denom))
    f2 = -3.0 / (denom * denom * torch.sqrt(denom)) + 15.0 * r2 / (denom * denom * denom * torch.sqrt(denom))
    n1 = -3.0 / (safe_r2 * safe_r2)
    n2 = 12.0 / (safe_r2 * safe_r2 * r)
    rf = f1 + s1 * newton + (split - 1.0) * n1
    rs = f2 + s2 * newton + 2.0 * s1 * n1 + (split - 1.0) * n2
    fo = rf / r
    cr = rs / r2 - rf / (r2 * r)
    c[0] = c[0] + G * (mdx * fo + 0.5 * tr * dx * fo + 0.5 * dx * dmd * cr)
    c[1] = c[1] + G * (mdy * fo + 0.5 * tr * dy * fo + 0.5 * dy * dmd * cr)
    c[2] = c[2] + G * (mdz * fo + 0.5 * tr * dz * fo + 0.5 * dz * dmd * cr)
    return torch.stack(c)


@pytest.fixture(scope="module")
def decisions():
    """The oracle's strict walk for every particle: summed pairs and accepted nodes."""
    x, m = clustered_set()
    n = x.shape[1]
    opts = options_struct(BOX, G=G, eps=EPS, leaf=LEAF, quad=True)
    leaves_of, acc_t, acc_node = [], [], []
    for i in range(n):
        _, lists = O.probe_walk(x[0], x[1], x[2], m, opts, MESH, x[:, i])
        leaves_of.append(lists["leaves"])
        for node in lists["accepted"]:
            acc_t.append(i)
            acc_node.append(int(node))
    # the probe's tree (its node numbers): the same sources, leaf size,
    # moments and softening; the inputs lie away from the box faces, so the
    # unwrapped frame is the input frame
    built = O.build_tree(x[0], x[1], x[2], m, leaf=LEAF, quad=True, eps=EPS)
    rows_i, rows_k = [], []
    for i in range(n):
        for leaf in leaves_of[i]:
            b, c = int(built["begin"][leaf]), int(built["count"][leaf])
            ks = built["order"][b:b + c].astype(np.int64)
            rows_i.append(np.full(ks.size, i))
            rows_k.append(ks)
    return dict(x=x, m=m, opts=opts, tree=built, pair_i=np.concatenate(rows_i), pair_k=np.concatenate(rows_k),
                acc_t=np.array(acc_t, np.int64), acc_node=np.array(acc_node, np.int64))


def model_accel(xt, m, dec, with_nodes=True):
    """Strict-walk force of the model (3, N) at positions xt (3, N), torch FP64."""
    n = xt.shape[1]
    h = BOX / MESH
    a_split = 1.25 * h
    rcut = 4.5 * 1.25 * h
    acc = R.pm_accel(xt, m, BOX, MESH, G, a_split)
    i = torch.from_numpy(dec["pair_i"])
    k = torch.from_numpy(dec["pair_k"])
    d = xt[:, k] - xt[:, i]
    d = d - BOX * torch.round(d / BOX)
    r2 = (d * d).sum(0)
    inside = (r2 <= rcut * rcut) & (r2 > 0)
    r2s = torch.where(inside, r2, torch.ones_like(r2))
    f = torch.where(inside, R.residual_inv_r3(r2s, EPS, a_split), torch.zeros_like(r2))
    acc = acc + torch.zeros(3, n, dtype=xt.dtype).index_add(1, i, G * f * m[k] * d)
    if with_nodes and dec["acc_t"].size:
        tree = dec["tree"]
        nodes = np.unique(dec["acc_node"])
        slot = {int(v): j for j, v in enumerate(nodes)}
        mass, com, quad = [], [], []
        for node in nodes:
            b, c = int(tree["begin"][node]), int(tree["count"][node])
            rows = torch.from_numpy(tree["order"][b:b + c].astype(np.int64))
            mm = m[rows]
            mt = mm.sum()
            cm = (xt[:, rows] * mm).sum(1) / mt
            dd = xt[:, rows] - cm[:, None]
            dx, dy, dz = dd
            d2 = (dd * dd).sum(0)
            q = torch.stack([(mm * (3 * dx * dx - d2)).sum(), (mm * 3 * dx * dy).sum(), (mm * 3 * dx * dz).sum(),
                             (mm * (3 * dy * dy - d2)).sum(), (mm * 3 * dy * dz).sum(),
                             (mm * (3 * dz * dz - d2)).sum(), (mm * d2).sum()])
            mass.append(mt)
            com.append(cm)
            quad.append(q)
        mass, com, quad = torch.stack(mass), torch.stack(com, 1), torch.stack(quad, 1)
        js = torch.tensor([slot[int(v)] for v in dec["acc_node"]])
        ts = torch.from_numpy(dec["acc_t"])
        off = com[:, js] - xt[:, ts]
        off = off - BOX * torch.round(off / BOX)
        contrib = node_accel(mass[js], quad[:, js], off, a_split, EPS)
        acc = acc + torch.zeros(3, n, dtype=xt.dtype).index_add(1, ts, contrib)
    return acc


def gpu_fields(dec, lam, order):
    x, m = dec["x"], dec["m"]
    dev = [torch.from_numpy(np.ascontiguousarray(v)).to("cuda:0") for v in (x[0], x[1], x[2], m)]
    lam_d = [torch.from_numpy(np.ascontiguousarray(v)).to("cuda:0") for v in lam]
    n = x.shape[1]
    opts = options_struct(BOX, G=G, eps=EPS, leaf=LEAF, quad=order == 2)
    with TreePmCoordinator(MESH) as coord:
        acc = tuple(torch.zeros(n, dtype=torch.float64, device="cuda:0") for _ in range(3))
        coord.solve_active_set(*dev, ForceAccumulator(acc), opts)
        out = tuple(torch.zeros(n, dtype=torch.float64, device="cuda:0") for _ in range(3))
        coord.force_vjp(*dev, lam_d, out, opts)
        coord.synchronize()
    return (np.stack([a.cpu().numpy() for a in acc]), np.stack([o.cpu().numpy() for o in out]))


def test_quadrupole_nodes_are_accepted(decisions):
    n = decisions["x"].shape[1]
    print(f"{decisions['acc_t'].size} accepted (target, node) pairs over {n} targets, "
          f"{np.unique(decisions['acc_node']).size} distinct nodes")
    assert decisions["acc_t"].size > n // 4


def test_forward_equals_model_with_accepted_nodes(decisions):
    xt = torch.from_numpy(decisions["x"])
    m = torch.from_numpy(decisions["m"])
    ref = model_accel(xt, m, decisions).numpy()
    rng = np.random.default_rng(3)
    lam = rng.normal(size=decisions["x"].shape)
    acc, _ = gpu_fields(decisions, lam, 2)
    err = rel_l2(acc, ref)
    print(f"forward vs model: {err:.3e}")
    assert err < 1e-9


def test_force_vjp_with_quadrupoles_equals_autograd(decisions):
    m = torch.from_numpy(decisions["m"])
    rng = np.random.default_rng(5)
    lam = rng.normal(size=decisions["x"].shape)
    xt = torch.from_numpy(decisions["x"]).clone().requires_grad_(True)
    loss = (torch.from_numpy(lam) * model_accel(xt, m, decisions)).sum()
    (ref,) = torch.autograd.grad(loss, xt)
    x0 = torch.from_numpy(decisions["x"]).clone().requires_grad_(True)
    loss0 = (torch.from_numpy(lam) * model_accel(x0, m, decisions, with_nodes=False)).sum()
    (ref_no_nodes,) = torch.autograd.grad(loss0, x0)
    _, vjp = gpu_fields(decisions, lam, 2)
    err = rel_l2(vjp, ref.numpy())
    node_part = rel_l2(ref_no_nodes.numpy(), ref.numpy())
    print(f"quadrupole VJP vs autograd: {err:.3e}; the accepted-node terms move the VJP by {node_part:.3e}")
    assert node_part > 1e-5
    assert err < 1e-11


def test_force_vjp_with_quadrupoles_matches_central_differences(decisions):
    # lam . (a(x + h v) - a(x - h v)) / 2h of the GPU forces against vjp . v,
    # with h small enough that no pair crosses r_cut and no node changes its
    # acceptance for these seeds
    x, m = decisions["x"], decisions["m"]
    n = x.shape[1]
    rng = np.random.default_rng(11)
    lam = rng.normal(size=x.shape)
    _, vjp = gpu_fields(decisions, lam, 2)
    opts = options_struct(BOX, G=G, eps=EPS, leaf=LEAF, quad=True)
    h = 1e-7
    errs = []
    with TreePmCoordinator(MESH) as coord:
        def forces(xx):
            dev = [torch.from_numpy(np.ascontiguousarray(v)).to("cuda:0") for v in (xx[0], xx[1], xx[2], m)]
            acc = tuple(torch.zeros(n, dtype=torch.float64, device="cuda:0") for _ in range(3))
            coord.solve_active_set(*dev, ForceAccumulator(acc), opts)
            coord.synchronize()
            return np.stack([a.cpu().numpy() for a in acc])
        for _ in range(4):
            v = rng.normal(size=x.shape)
            v /= np.linalg.norm(v)
            fd = (np.sum(lam * forces(x + h * v)) - np.sum(lam * forces(x - h * v))) / (2 * h)
            an = float(np.sum(vjp * v))
            errs.append(abs(fd - an) / abs(an))
    print("central differences, relative errors:", errs)
    assert np.median(errs) < 1e-5
