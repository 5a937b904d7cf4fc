"""GPU checks of the differentiable forward: the force VJP and the KDK adjoint.

There is no reference counterpart (the reference has no adjoint). The bars:
- the forward accelerations of the monopole strict walk equal the FP64
  direct-sum model of tests/adjoint_reference.py (relative L2 < 1e-9);
- the force VJP equals torch.autograd on that model (relative L2 < 1e-8) and
  the directional central difference of the GPU forces (relative < 1e-5);
- the gradient of the density loss after several KDK steps equals the
  direct-sum model's autograd gradient (relative L2 < 1e-7) and central
  differences along random unit directions (relative < 1e-4).
The residual force is cut at r_cut where it is still 1.75e-2 of Newtonian, so
the forward jumps whenever a pair crosses r_cut; central differences use steps
of 1e-6 mean spacings, small enough that no pair crosses for these seeds
(steps of 1e-5 to 1e-3 spacings already straddle such crossings, and the
direct-sum model shows the same jumps).
"""
from __future__ import annotations

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
if not torch.cuda.is_available():  # pragma: no cover - GPU-only module
    pytest.skip("no HIP device", allow_module_level=True)

import adjoint_reference as R  # noqa: E402
from cosmosim import adjoint  # noqa: E402


def rel_l2(a, b):
    a = np.asarray(a, dtype=np.float64).ravel()
    b = np.asarray(b, dtype=np.float64).ravel()
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def problem(n_side, steps, a_end=0.3, checkpoint_every=2):
    return adjoint.adjoint_problem(n_side=n_side, steps=steps, a_end=a_end, checkpoint_every=checkpoint_every)


def positions(prob, psi):
    return prob.wrap(prob.lattice + psi.to(torch.float64))


def test_forward_accelerations_equal_direct_sum_model():
    prob, psi, _ = problem(8, 1)
    try:
        x = positions(prob, psi)
        g = prob.forces(x)
        ref = R.reference_accel(prob, x.cpu())
        assert rel_l2(g.cpu().numpy(), ref.numpy()) < 1e-9
    finally:
        prob.close()


def test_force_vjp_equals_autograd_of_direct_sum_model():
    prob, psi, _ = problem(8, 1)
    try:
        x = positions(prob, psi)
        lam = torch.tensor(np.random.default_rng(5).standard_normal(x.shape), dtype=torch.float64,
                           device=x.device)
        ours = prob.force_vjp(x, lam)
        xc = x.cpu().clone().requires_grad_(True)
        (ref,) = torch.autograd.grad((R.reference_accel(prob, xc) * lam.cpu()).sum(), xc)
        assert rel_l2(ours.cpu().numpy(), ref.numpy()) < 1e-8
    finally:
        prob.close()


def test_force_vjp_matches_central_differences_of_gpu_forces():
    prob, psi, _ = problem(16, 1)
    try:
        x = positions(prob, psi)
        rng = np.random.default_rng(9)
        lam = torch.tensor(rng.standard_normal(x.shape), dtype=torch.float64, device=x.device)
        s = prob.force_vjp(x, lam)
        h = 1e-6 * prob.box / prob.n_side
        for _ in range(3):
            d = rng.standard_normal(x.shape)
            d /= np.linalg.norm(d)
            dt = torch.tensor(d, dtype=torch.float64, device=x.device)
            gp = prob.forces(prob.wrap(x + h * dt))
            gm = prob.forces(prob.wrap(x - h * dt))
            fd = float(((gp - gm) * lam).sum()) / (2 * h)
            an = float((s * dt).sum())
            assert abs(fd - an) <= 1e-5 * abs(an), (fd, an)
    finally:
        prob.close()


def test_kdk_gradient_equals_direct_sum_autograd():
    prob, psi, target = problem(8, 3)
    try:
        psi64 = psi.to(torch.float64).requires_grad_(True)
        loss = adjoint.density_loss(psi64, target, prob)
        (grad,) = torch.autograd.grad(loss, psi64)
        ref_loss, ref_grad = R.reference_gradient(prob, psi.to(torch.float64).cpu(), target.cpu())
        assert abs(float(loss) - ref_loss) <= 1e-9 * abs(ref_loss)
        assert rel_l2(grad.cpu().numpy(), ref_grad.numpy()) < 1e-7
        # the backward recomputed one position inside each two-step segment
        assert prob.vjp_evaluations == prob.steps + 1
    finally:
        prob.close()


def test_kdk_gradient_matches_central_differences():
    prob, psi, target = problem(16, 4)
    try:
        psi64 = psi.to(torch.float64).requires_grad_(True)
        loss = adjoint.density_loss(psi64, target, prob)
        (grad,) = torch.autograd.grad(loss, psi64)
        rng = np.random.default_rng(17)
        h = 1e-6 * prob.box / prob.n_side
        for _ in range(3):
            d = rng.standard_normal(psi.shape)
            d /= np.linalg.norm(d)
            dt = torch.tensor(d, dtype=torch.float64, device=psi.device)
            with torch.no_grad():
                lp = float(adjoint.density_loss(psi64 + h * dt, target, prob))
                lm = float(adjoint.density_loss(psi64 - h * dt, target, prob))
            fd = (lp - lm) / (2 * h)
            an = float((grad * dt).sum())
            assert abs(fd - an) <= 1e-4 * max(abs(an), 1e-12), (fd, an)
    finally:
        prob.close()


def test_cic_density_vjp_is_the_transpose():
    prob, psi, _ = problem(8, 1)
    try:
        x = positions(prob, psi)
        m = prob.loss_mesh
        g = torch.tensor(np.random.default_rng(3).standard_normal((m, m, m)), dtype=torch.float64,
                         device=x.device)
        out = prob.cic_density_vjp(x, g)
        xc = x.cpu().clone().requires_grad_(True)
        idx, w = R.cic_weights(xc, prob.box, m)
        count = torch.zeros(m**3, dtype=torch.float64).index_add(0, idx.reshape(-1), w.reshape(-1))
        delta = count * (m**3 / xc.shape[1]) - 1.0
        (ref,) = torch.autograd.grad((delta * g.cpu().reshape(-1)).sum(), xc)
        assert rel_l2(out.cpu().numpy(), ref.numpy()) < 1e-12
        mine = prob.cic_density(x)
        assert rel_l2(mine.cpu().numpy().ravel() + 1.0, delta.detach().numpy() + 1.0) < 1e-13
    finally:
        prob.close()


@pytest.mark.parametrize("field,value", [("acceptance_policy", 1), ("precision", 1)])
def test_force_vjp_rejects_unsupported_options(field, value):
    prob, psi, _ = problem(8, 1)
    try:
        x = positions(prob, psi)
        setattr(prob.options, field, value)
        with pytest.raises(NotImplementedError):
            prob.force_vjp(x, torch.zeros_like(x))
    finally:
        prob.close()
