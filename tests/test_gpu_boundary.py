"""GPU checks of the coordinator boundary's contracts.

- PM cadence reuse keyed by field identity (tree_pm_coupling.cpp:2483-2532):
  a refresh stores the caller's source generation, PM field version and force
  epoch with the mesh-level options; a reuse (refresh 0) is admitted only when
  they match (the stale window skips the source generation and epoch, never
  the field version) and otherwise fails with std::runtime_error, here
  CHUI_ERR_RUNTIME. Reuse with the same identity reproduces the refresh's PM
  lanes bit for bit.
- The walks' traversal-fault word is reported as CHUI_ERR_LOGIC by the next
  synchronising call (set here from the host, so no kernel misbehaves).
- Split diagnostics (computeTreePmDiagnostics, tree_pm_coupling.cpp:6276-6312)
  against the GPU walk: the short-range lane of an isolated pair at r_s and
  just inside r_cut is S(r) G m / r^2 with the S the diagnostics report, and a
  pair just beyond r_cut contributes nothing.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

from helpers import empty_dev, options_struct, to_dev

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
if not torch.cuda.is_available():  # pragma: no cover - GPU-only module
    pytest.skip("no HIP device", allow_module_level=True)

from cosmosim import ForceAccumulator, TreePmCoordinator, split_diagnostics  # noqa: E402
from cosmosim import _lib as L  # noqa: E402
from test_oracle_pins import clustered  # noqa: E402

BOX = 10.0
MESH = 16


def _identity(o, generation, version=1, epoch=(1, 1.0), stale=0):
    o.source_generation = generation
    o.pm_field_version = version
    o.force_epoch_sequence = epoch[0]
    o.force_epoch_scale_factor = epoch[1]
    o.allow_stale_long_range_field = stale
    return o


def _solve(coord, dev, o, refresh, precision):
    o.precision = precision
    dt = torch.float64 if precision == 0 else torch.float32
    n = int(dev[0].shape[0])
    acc = tuple(empty_dev(n, dt) for _ in range(3))
    pm = tuple(empty_dev(n, dt) for _ in range(3))
    coord.solve_active_set_with_pm_cadence(*dev, ForceAccumulator(acc, None, pm), o, refresh)
    coord.synchronize()
    return np.stack([a.cpu().numpy() for a in acc]), np.stack([a.cpu().numpy() for a in pm])


@pytest.fixture(scope="module")
def sources():
    x, y, z, m = clustered(1200, 5, BOX)
    return x, y, z, m


@pytest.mark.parametrize("precision", [0, 1])
def test_reuse_same_identity_is_bit_identical(sources, precision):
    dev = to_dev(*sources)
    o = _identity(options_struct(BOX, eps=0.01), generation=7, version=2, epoch=(3, 0.5))
    with TreePmCoordinator(MESH) as coord:
        acc_r, pm_r = _solve(coord, dev, o, True, precision)
        acc_u, pm_u = _solve(coord, dev, o, False, precision)
    assert np.array_equal(pm_r, pm_u)
    assert np.array_equal(acc_r, acc_u)


@pytest.mark.parametrize("field,value", [
    ("source_generation", 8), ("force_epoch_sequence", 4), ("force_epoch_scale_factor", 0.51),
    ("pm_field_version", 3),
])
def test_reuse_rejected_when_identity_changes(sources, field, value):
    dev = to_dev(*sources)
    o = _identity(options_struct(BOX, eps=0.01), generation=7, version=2, epoch=(3, 0.5))
    with TreePmCoordinator(MESH) as coord:
        _solve(coord, dev, o, True, 0)
        setattr(o, field, value)
        with pytest.raises(L.ChuiRuntimeError, match="without a compatible PM field") as err:
            _solve(coord, dev, o, False, 0)
        assert err.value.status == L.CHUI_ERR_RUNTIME
        # the field stays installed: the original identity still reuses it
        o2 = _identity(options_struct(BOX, eps=0.01), generation=7, version=2, epoch=(3, 0.5))
        _solve(coord, dev, o2, False, 0)


@pytest.mark.parametrize("field,value", [
    ("scale_factor", 0.5), ("gravitational_constant", 2.0), ("asmth_cells", 1.5), ("assignment_scheme", 1),
    ("window_deconvolution", 1),
])
def test_reuse_rejected_after_changed_options(sources, field, value):
    dev = to_dev(*sources)
    o = _identity(options_struct(BOX, eps=0.01), generation=1)
    with TreePmCoordinator(MESH) as coord:
        _solve(coord, dev, o, True, 0)
        setattr(o, field, value)
        with pytest.raises(L.ChuiRuntimeError):
            _solve(coord, dev, o, False, 0)


def test_reuse_without_a_field_and_per_precision(sources):
    dev = to_dev(*sources)
    o = _identity(options_struct(BOX, eps=0.01), generation=1)
    with TreePmCoordinator(MESH) as coord:
        with pytest.raises(L.ChuiRuntimeError, match="unavailable for reuse"):
            _solve(coord, dev, o, False, 0)
        _solve(coord, dev, o, True, 0)
        # the mixed form keeps its own (FP32) field: none installed yet
        with pytest.raises(L.ChuiRuntimeError, match="unavailable for reuse"):
            _solve(coord, dev, o, False, 1)


def test_stale_window_admits_a_new_generation_not_a_new_version(sources):
    x, y, z, m = sources
    dev = to_dev(x, y, z, m)
    o = _identity(options_struct(BOX, eps=0.01), generation=1, version=5)
    moved = to_dev((x + 0.05) % BOX, y, z, m)
    with TreePmCoordinator(MESH) as coord:
        _, pm_old = _solve(coord, dev, o, True, 0)
        _identity(o, generation=2, version=5, epoch=(2, 1.0), stale=1)
        _, pm_stale = _solve(coord, moved, o, False, 0)
        # the old field interpolated at the moved targets, not a new field:
        # it differs from the refresh at the moved sources, by about the
        # field's variation over 0.05 length units
        _, pm_fresh = _solve(coord, moved, _identity(o, generation=2, version=5, epoch=(2, 1.0)), True, 0)
        assert not np.array_equal(pm_stale, pm_fresh) and not np.array_equal(pm_stale, pm_old)
        assert np.sqrt(np.sum((pm_stale - pm_fresh) ** 2) / np.sum(pm_fresh**2)) < 0.2
        _identity(o, generation=3, version=6, epoch=(3, 1.0), stale=1)
        with pytest.raises(L.ChuiRuntimeError):
            _solve(coord, moved, o, False, 0)


@pytest.mark.parametrize("precision", [0, 1])
def test_walk_fault_is_reported_as_logic_error(sources, precision):
    x, y, z, m = sources
    o = options_struct(BOX, eps=0.01, precision=precision)
    lib = L.load()
    n = x.size
    with TreePmCoordinator(MESH) as coord:
        # host-pointer solve: synchronises and reports before it returns
        out = [np.zeros(n) for _ in range(3)]
        L.check(lib.chui_treepm_debug_inject_walk_fault(coord.handle))
        with pytest.raises(L.ChuiLogicError, match="lost progress") as err:
            coord.solve_active_set(x, y, z, m, ForceAccumulator(tuple(out)), o)
        assert err.value.status == L.CHUI_ERR_LOGIC
        coord.solve_active_set(x, y, z, m, ForceAccumulator(tuple(out)), o)  # the word was cleared
        # device-pointer solve: reported by the next synchronising call
        dev = to_dev(x, y, z, m)
        dt = torch.float64 if precision == 0 else torch.float32
        acc = tuple(empty_dev(n, dt) for _ in range(3))
        L.check(lib.chui_treepm_debug_inject_walk_fault(coord.handle))
        coord.solve_active_set(*dev, ForceAccumulator(acc), o)
        with pytest.raises(L.ChuiLogicError):
            coord.synchronize()
        coord.solve_active_set(*dev, ForceAccumulator(acc), o)
        coord.synchronize()
        # a solve with diagnostics checks before it returns
        L.check(lib.chui_treepm_debug_inject_walk_fault(coord.handle))
        with pytest.raises(L.ChuiLogicError):
            coord.solve_active_set(*dev, ForceAccumulator(acc), o, want_diagnostics=True)


@pytest.mark.parametrize("precision", [0, 1])
def test_split_diagnostics_against_the_walk(precision):
    box, mesh = 64.0, 32
    h = box / mesh
    d = split_diagnostics(1.25, 4.5 * 1.25, h)
    rs, rc = d["split_scale_comoving"], d["cutoff_radius_comoving"]
    o = options_struct(box, G=1.0, eps=0.0, quad=False, precision=precision)
    c = 0.5 * box + 0.123
    # S falls by 9.2 times the relative step of r at r_cut (d ln S / d ln r at
    # q = 2.25), so the pair just inside r_cut sits 1e-8 (FP64) or 2e-6 (mixed,
    # whose r^2 rounds by ~1e-7) inside and the tolerance allows for the step
    inside = 1e-8 if precision == 0 else 2e-6
    tol_split = 1e-12 if precision == 0 else 2e-6
    tol_cut = 2e-7 if precision == 0 else 5e-5
    cases = [(rs, d["short_range_factor_at_split"], tol_split),
             (rc * (1.0 - inside), d["short_range_factor_at_cutoff"], tol_cut),
             (rc * (1.0 + 1e-5), 0.0, 0.0)]
    dt = torch.float64 if precision == 0 else torch.float32
    for r, s_expected, tol in cases:
        x = np.array([c, c + r])
        y = np.array([c, c])
        z = np.array([c, c])
        m = np.array([1.0, 1.0])
        dev = to_dev(x, y, z, m)
        acc = tuple(empty_dev(2, dt) for _ in range(3))
        sr = tuple(empty_dev(2, dt) for _ in range(3))
        with TreePmCoordinator(mesh) as coord:
            diag = coord.solve_active_set(*dev, ForceAccumulator(acc, None, None, sr), o,
                                          want_diagnostics=True).diagnostics
        assert diag["split_scale_comoving"] == rs and diag["cutoff_radius_comoving"] == rc
        ax = float(sr[0][0].cpu())
        newton = 1.0 / (r * r)
        if s_expected == 0.0:
            assert ax == 0.0 and float(sr[0][1].cpu()) == 0.0
        else:
            assert abs(ax / newton - s_expected) <= tol * s_expected
            assert float(sr[0][1].cpu()) == pytest.approx(-ax, rel=1e-12 if precision == 0 else 1e-6)
    assert abs(d["short_range_factor_at_split"] + d["long_range_factor_at_split"] - 1.0) <= 1e-15


def test_options_struct_size_matches_header():
    o = L.default_options()
    assert o.struct_size == C.sizeof(L.Options)
