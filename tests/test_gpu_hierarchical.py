"""GPU hierarchical KDK block (cosmosim.hierarchical.HierarchicalKdk) against its
CPU restatement with the oracle's forces (tests/hierarchical_reference.py).

Same bins, same number of short-range-only fine-tick solves; FP64 positions and
velocities agree to 1e-9 of the box and of the velocity scale; mixed precision
stays within 1e-4 of the mean spacing.
"""
from __future__ import annotations

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
if not torch.cuda.is_available():  # pragma: no cover - GPU-only module
    pytest.skip("no HIP device", allow_module_level=True)

import hierarchical_reference as HR  # noqa: E402
from cosmosim import _lib as L  # noqa: E402
from cosmosim import hierarchical as H  # noqa: E402
from cosmosim import ics  # noqa: E402
from helpers import options_struct  # noqa: E402

COSMO = H.Cosmology(ics.H0_CODE, ics.OMEGA_M, 1.0 - ics.OMEGA_M)


@pytest.mark.parametrize("n_side,max_bin", [(8, 2), (12, 3)])
@pytest.mark.parametrize("precision", [L.PRECISION_FP64, L.PRECISION_MIXED])
def test_block_matches_cpu_restatement(n_side, max_bin, precision):
    w = ics.config(0, n_side=n_side)
    x = np.stack([w.x, w.y, w.z])
    u = np.stack([w.vx, w.vy, w.vz]) / w.scale_factor
    o = options_struct(w.box, eps=w.softening)
    xr, ur, plan_r, n_sr = HR.reference_block(x, u, w.mass, o, w.mesh, COSMO, max_bin, 0.04, max_span=2e-3)
    assert len(set(plan_r.bins.tolist())) > 1
    o.precision = precision
    dev = lambda a: [torch.tensor(c, dtype=torch.float64, device="cuda:0") for c in a]  # noqa: E731
    kdk = H.HierarchicalKdk(w.mesh, dev(x), dev(u), torch.tensor(w.mass, device="cuda:0"), o, COSMO, max_bin)
    try:
        # mixed-precision forces could move a particle across a bin boundary: use the reference's bins
        plan = kdk.block(0.04, max_span=2e-3, bins=None if precision == L.PRECISION_FP64 else plan_r.bins)
        xg = np.stack([c.cpu().numpy() for c in kdk.x])
        ug = np.stack([c.cpu().numpy() for c in kdk.u])
        assert np.array_equal(plan.bins.cpu().numpy(), plan_r.bins)
        assert kdk.short_range_solves == n_sr and kdk.solves == 2
    finally:
        kdk.close()
    d = xg - xr
    d -= w.box * np.round(d / w.box)
    if precision == L.PRECISION_FP64:
        assert np.abs(d).max() < 1e-9 * w.box
        assert np.abs(ug - ur).max() < 1e-9 * np.abs(ur).max()
    else:
        assert np.abs(d).max() < 1e-4 * w.box / n_side
