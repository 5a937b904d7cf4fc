"""Host-side checks of the ensemble workload law (no GPU)."""
from __future__ import annotations

import numpy as np
import pytest

pytest.importorskip("torch")

from cosmosim import ensemble as E  # noqa: E402


def test_ensemble_law_counts_sizes_and_ranges():
    specs = E.ensemble_specs()
    assert len(specs) == 48
    sizes = [s.n_side for s in specs]
    assert [sizes.count(k) for k in (32, 64, 96, 128)] == [16, 16, 12, 4]
    assert all(s.mesh == 2 * s.n_side for s in specs)
    assert all(abs(s.box - 50.0 * s.n_side / 64) < 1e-12 for s in specs)
    assert all(-2.8 <= s.spectral_index <= -2.0 for s in specs)
    assert all(0.05 <= s.a_begin <= 0.1 and 0.25 <= s.a_end - s.a_begin <= 0.45 for s in specs)
    assert len({s.seed for s in specs}) == 48
    assert sum(s.n for s in specs) == 16 * 32**3 + 16 * 64**3 + 12 * 96**3 + 4 * 128**3


def test_box_initial_conditions_are_seeded_and_scaled():
    s = E.ensemble_specs(counts=((16, 2),))[0]
    lat, psi = E.box_initial_conditions(s)
    lat2, psi2 = E.box_initial_conditions(s)
    assert np.array_equal(psi, psi2) and np.array_equal(lat, lat2)
    spacing = s.box / s.n_side
    assert abs(np.sqrt(np.mean(psi**2)) - 0.1 * spacing) < 1e-12 * spacing
    assert lat.shape == (3, s.n) and lat.min() > 0 and lat.max() < s.box
