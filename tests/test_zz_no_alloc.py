"""GPU check of the allocation-free reduction entry point.

chui_treepm_mass_momentum_noalloc_device keeps its partial sums in the
library's per-GPU device array instead of a context buffer and writes the
result to caller-owned device memory. It must equal
chui_treepm_mass_momentum_device bit for bit (same grid, same summation tree),
also on a second call with a smaller input right after the first (the array
keeps the first call's partials), and it must reject null lanes.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
if not torch.cuda.is_available():  # pragma: no cover - GPU-only module
    pytest.skip("no HIP device", allow_module_level=True)

from cosmosim import TreePmCoordinator  # noqa: E402
from cosmosim import _lib as L  # noqa: E402


def _lanes(n, seed):
    rng = np.random.default_rng(seed)
    host = [rng.uniform(0.5, 2.0, n)] + [rng.normal(0.0, 300.0, n) for _ in range(3)]
    return [torch.from_numpy(a).to("cuda:0") for a in host], host


def _existing(lib, h, lanes, n):
    torch.cuda.synchronize()
    out = (C.c_double * 4)()
    L.check(lib.chui_treepm_mass_momentum_device(h, *[t.data_ptr() for t in lanes], n, out))
    return np.array(list(out))


def _noalloc(lib, h, lanes, n, out_dev):
    torch.cuda.synchronize()  # inputs and the output's fill come from torch's stream
    L.check(lib.chui_treepm_mass_momentum_noalloc_device(h, *[t.data_ptr() for t in lanes], n,
                                                          out_dev.data_ptr()))
    L.check(lib.chui_treepm_synchronize(h))
    return out_dev.cpu().numpy().copy()


def test_noalloc_reduction_equals_the_allocating_entry_point():
    lib = L.load()
    with TreePmCoordinator(8) as coord:
        h = coord._h
        out_dev = torch.full((4,), float("nan"), dtype=torch.float64, device="cuda:0")
        for n, seed in ((1_000_003, 1), (4_097, 2), (0, 3)):
            lanes, host = _lanes(n, seed)
            got = _noalloc(lib, h, lanes, n, out_dev)
            want = _existing(lib, h, lanes, n)
            assert np.array_equal(got, want), (n, got, want)
            m = host[0]
            ref = [m.sum()] + [np.sum(m * v) for v in host[1:]]
            scale = [m.sum()] + [np.sum(np.abs(m * v)) for v in host[1:]]
            for k in range(4):
                assert abs(got[k] - ref[k]) <= 1e-12 * max(scale[k], 1.0)


def test_noalloc_reduction_rejects_null_lanes():
    lib = L.load()
    with TreePmCoordinator(8) as coord:
        lanes, _ = _lanes(16, 4)
        out_dev = torch.zeros(4, dtype=torch.float64, device="cuda:0")
        st = lib.chui_treepm_mass_momentum_noalloc_device(coord._h, lanes[0].data_ptr(), None, lanes[2].data_ptr(),
                                                          lanes[3].data_ptr(), 16, out_dev.data_ptr())
        assert st == L.CHUI_ERR_INVALID_ARGUMENT
        st = lib.chui_treepm_mass_momentum_noalloc_device(coord._h, *[t.data_ptr() for t in lanes], 16, None)
        assert st == L.CHUI_ERR_INVALID_ARGUMENT
