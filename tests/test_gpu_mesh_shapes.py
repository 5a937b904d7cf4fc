"""GPU parity of the PM part on meshes that are not powers of two.

Mesh axes of the form 2^a 3^b (a >= 1) take the mixed-radix LDS line passes
(radix 3 then 8/4/2). The PM accelerations equal the oracle's (its own FFT,
a direct DFT for these lengths) to 1e-10 relative L2 in FP64, including
rectangular meshes mixing power-of-two and other axes; at a 192^3 mesh, too
large for the oracle's DFT, they equal an independent FP64 torch.fft pipeline
(CIC deposit, Green's function, spectral gradient, CIC interpolation;
tests/adjoint_reference.py). Mixed precision keeps its error budget.
"""
from __future__ import annotations

import numpy as np
import pytest

from helpers import error_metrics, options_struct, rel_l2
from oracle import oracle as O

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
if not torch.cuda.is_available():  # pragma: no cover - GPU-only module
    pytest.skip("no HIP device", allow_module_level=True)

import adjoint_reference as R  # noqa: E402
from test_gpu_parity import gpu_solve  # noqa: E402


def particles(n, box, seed):
    rng = np.random.default_rng(seed)
    box = np.broadcast_to(np.asarray(box, dtype=float), (3,))
    pos = rng.uniform(0.0, 1.0, (3, n)) * box[:, None]
    m = rng.uniform(0.5, 1.5, n)
    return pos[0].copy(), pos[1].copy(), pos[2].copy(), m


@pytest.mark.parametrize("mesh,box,assignment,deconv", [
    ((24, 24, 24), 10.0, 0, 0),
    ((48, 48, 48), 10.0, 1, 1),
    ((12, 32, 96), (6.0, 8.0, 12.0), 0, 1),
    ((96, 6, 18), (16.0, 6.0, 9.0), 1, 0),
])
def test_pm_fp64_non_power_of_two_matches_oracle(mesh, box, assignment, deconv):
    x, y, z, m = particles(3000, box, 3)
    opts = options_struct(box, eps=0.0, assignment=assignment, deconv=deconv, rcut_cells=2.0)
    _, pm, _, _, _ = gpu_solve(x, y, z, m, opts, mesh)
    ref = O.pm_accel(x, y, z, m, opts, mesh)
    assert rel_l2(pm, ref) < 1e-10


def test_total_fp64_non_power_of_two_matches_oracle():
    x, y, z, m = particles(4000, 20.0, 8)
    opts = options_struct(20.0, eps=0.01)
    acc, _, _, _, _ = gpu_solve(x, y, z, m, opts, 48)
    ref = O.solve(x, y, z, m, opts, 48)
    assert rel_l2(acc, ref) < 1e-10


def test_mixed_non_power_of_two_error_budget():
    x, y, z, m = particles(6000, 30.0, 9)
    opts = options_struct(30.0, eps=0.01)
    acc, _, _, _, _ = gpu_solve(x, y, z, m, opts, 48, precision=1)
    ref = O.solve(x, y, z, m, opts, 48)
    e = error_metrics(acc, ref)
    assert e["relative_l2"] < 1e-4 and e["p99"] < 1e-3


@pytest.mark.parametrize("precision,tol", [(0, 1e-10), (1, 2e-5)])
def test_pm_192_cubed_matches_torch_fft_pipeline(precision, tol):
    box, n_mesh = 50.0, 192
    x, y, z, m = particles(40000, box, 12)
    opts = options_struct(box, eps=0.0, rcut_cells=4.5 * 1.25)
    _, pm, _, _, _ = gpu_solve(x, y, z, m, opts, n_mesh, precision=precision)
    pos = torch.tensor(np.stack([x, y, z]), dtype=torch.float64)
    rs = 1.25 * box / n_mesh
    ref = R.pm_accel(pos, torch.tensor(m, dtype=torch.float64), box, n_mesh, opts.gravitational_constant, rs)
    assert rel_l2(pm, ref.numpy()) < tol
