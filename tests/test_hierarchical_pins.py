"""CPU pins of the hierarchical KDK block schedule (cosmosim.hierarchical.plan_block)
and its CPU restatement (tests/hierarchical_reference.py), with the oracle's forces.

- the half-tick grid is at equal cosmic time (bisection of the 64-sample
  midpoint integral, advanceScaleFactorByCosmicTime);
- free expansion (no force): x = x0 + a0 u0 int dt/a^2 and u = u0 a0/a over
  the whole block, against scipy quadrature;
- every particle in the top bin gives the block of one tick over the same
  span (the opening/closing half-kicks cover the same scale-factor ranges);
- the bins follow the criterion eta sqrt(eps/(|A|/a^3)) with quantum
  span / 2^M (docs/time_integration.md:392-470).
"""
from __future__ import annotations

import numpy as np
import pytest
from scipy import integrate

pytest.importorskip("torch")

import hierarchical_reference as HR  # noqa: E402
from cosmosim import hierarchical as H  # noqa: E402
from cosmosim import ics  # noqa: E402
from helpers import options_struct  # noqa: E402

COSMO = H.Cosmology(ics.H0_CODE, ics.OMEGA_M, 1.0 - ics.OMEGA_M)


def small_box(n_side=8, seed=3):
    w = ics.config(0, n_side=n_side)
    u = np.stack([w.vx, w.vy, w.vz]) / w.scale_factor  # u = p / a from u_pec = a dx/dt
    return w, np.stack([w.x, w.y, w.z]), u


def test_half_tick_grid_is_at_equal_cosmic_time():
    plan = H.plan_block(COSMO, 0.1, np.array([1.0, 4.0, 16.0]), 0.01, 3, max_span=2e-4)
    for h in range(1, len(plan.scale)):
        # each half tick is advanced from the block start (time_integration.cpp:1395-1400)
        dt = COSMO.cosmic_time(plan.scale[0], plan.scale[h])
        assert abs(dt - 0.5 * h * plan.quantum) < 1e-12 * plan.quantum
        t_ref, _ = integrate.quad(lambda a: 1.0 / (a * COSMO.hubble(a)), plan.scale[0], plan.scale[h],
                                  epsabs=0, epsrel=1e-13)
        assert abs(t_ref - 0.5 * h * plan.quantum) < 1e-5 * 0.5 * h * plan.quantum  # midpoint-rule error
    # bins: largest b <= M with quantum 2^b <= dt_i
    dt = 0.2 * np.sqrt(0.01 / (np.array([1.0, 4.0, 16.0]) / 0.1**3))
    for d, b in zip(dt, plan.bins):
        assert plan.quantum * 2**b <= d * (1 + 1e-12) and (b == 3 or plan.quantum * 2 ** (b + 1) > d)


def test_free_expansion_closed_form():
    w, x, u = small_box()
    o = options_struct(w.box, G=1e-300, eps=w.softening)
    m = w.mass
    a0 = 0.04
    x1, u1, plan, _ = HR.reference_block(x, u, m, o, w.mesh, COSMO, 2, a0, max_span=5e-4,
                                         bins=np.full(m.size, 1))
    a1 = plan.scale[-1]
    drift, _ = integrate.quad(lambda a: 1.0 / (a**3 * COSMO.hubble(a)), a0, a1, epsabs=0, epsrel=1e-13)
    disp = a0 * u * drift
    xe = np.mod(x + disp, w.box)
    d = x1 - xe
    d -= w.box * np.round(d / w.box)
    # the per-tick 64-sample midpoint integrals against the exact integral
    assert np.abs(d).max() < 2e-5 * np.abs(disp).max()
    assert np.allclose(u1, u * a0 / a1, rtol=1e-12, atol=0)


def test_top_bin_block_equals_one_tick_block():
    w, x, u = small_box()
    o = options_struct(w.box, eps=w.softening)
    m = w.mass
    pm, sr = HR.forces(x, m, o, w.mesh)
    a0 = 0.04
    dt_min = (0.2 * np.sqrt(w.softening / (np.sqrt(((pm + sr) ** 2).sum(axis=0)) / a0**3))).min()
    span = 0.5 * dt_min  # below both blocks' own limits, so both span exactly this
    xa, ua, pa, na = HR.reference_block(x, u, m, o, w.mesh, COSMO, 2, 0.04, max_span=span, bins=np.full(m.size, 2))
    xb, ub, pb, nb = HR.reference_block(x, u, m, o, w.mesh, COSMO, 0, 0.04, max_span=span)
    assert na == 0 and nb == 0  # no fine-tick solve: nobody closes before the block end
    assert abs(pa.scale[-1] - pb.scale[-1]) < 1e-12 * pb.scale[-1]
    d = xa - xb
    d -= w.box * np.round(d / w.box)
    assert np.abs(d).max() < 1e-9 * w.box
    assert np.abs(ua - ub).max() < 1e-7 * np.abs(ub).max()


def test_mixed_bins_differ_from_the_coarsest_bin_by_a_small_splitting_error():
    w, x, u = small_box()
    o = options_struct(w.box, eps=w.softening)
    m = w.mass
    x1, u1, plan, n_sr = HR.reference_block(x, u, m, o, w.mesh, COSMO, 2, 0.04, eta=0.2, max_span=2e-3)
    assert len(set(plan.bins.tolist())) > 1 and n_sr >= 1
    # the same block span with every particle on the coarsest interval
    x0, u0, p0, _ = HR.reference_block(x, u, m, o, w.mesh, COSMO, 2, 0.04, eta=0.2, max_span=2e-3,
                                       bins=np.full(m.size, 2))
    assert p0.quantum == plan.quantum
    d = x1 - x0
    d -= w.box * np.round(d / w.box)
    spacing = w.box / 8
    assert 0 < np.abs(d).max() < 0.05 * spacing
