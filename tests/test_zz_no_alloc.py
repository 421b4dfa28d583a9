"""mmpbsa_gb_energy_noalloc: the GB energy with every temporary in the
library's fixed per-device scratch array (no device allocation in the call),
called through the C ABI with caller-allocated device inputs and outputs.
Its Born radii and energy equal mmpbsa_gb_energy's bit for bit (same
kernels, same partial-sum layout) on two calls in a row, the second on a
smaller complex (the array keeps the first call's contents), with and
without the radii output; a system whose temporaries exceed the array is
refused with MMPBSA_ERR_TOO_LARGE."""
import ctypes as C

import numpy as np
import pytest

import rust_simulation_tools as rst
from rust_simulation_tools import _abi, synthetic

pytestmark = pytest.mark.gpu


def _noalloc(top, x_nm, gp, want_radii):
    lib = _abi.load_library()
    c = rst._ctx(top)
    n = top.n_atoms
    x = np.ascontiguousarray(np.asarray(x_nm, np.float64) * 10.0)  # as the library sees nm input
    d_x, d_r, d_e = C.c_void_p(), C.c_void_p(), C.c_void_p()
    assert lib.mmpbsa_device_alloc(0, 24 * n, C.byref(d_x)) == 0
    assert lib.mmpbsa_device_alloc(0, 8 * n, C.byref(d_r)) == 0
    assert lib.mmpbsa_device_alloc(0, 8, C.byref(d_e)) == 0
    try:
        assert lib.mmpbsa_device_upload(d_x, x.ctypes.data, 24 * n) == 0
        p = gp._c
        rc = lib.mmpbsa_gb_energy_noalloc(c.h, d_x, C.byref(p), d_r if want_radii else None, d_e)
        assert rc == 0, lib.mmpbsa_last_error(c.h)
        e = np.zeros(1)
        R = np.zeros(n)
        assert lib.mmpbsa_device_download(e.ctypes.data, d_e, 8) == 0
        if want_radii:
            assert lib.mmpbsa_device_download(R.ctypes.data, d_r, 8 * n) == 0
        return float(e[0]), R
    finally:
        for d in (d_x, d_r, d_e):
            lib.mmpbsa_device_free(d)


@pytest.mark.parametrize("want_radii", [True, False])
def test_noalloc_matches_gb_energy(cfg0, want_radii):
    top, x, _, _, _ = cfg0
    small, xs, _, _ = synthetic.build_complex(600, 64, seed=11)
    for model in (2, 0):
        gp = rst.GbParams(model=model)
        for t, xx in ((top, x / 10.0), (small, xs / 10.0)):   # second call: fewer atoms
            ref = rst.compute_gb_energy(t, xx, gp)
            e, R = _noalloc(t, xx, gp, want_radii)
            assert e == ref.total, (model, t.n_atoms, e, ref.total)
            if want_radii:
                assert np.array_equal(R, ref.born_radii())


def test_noalloc_refuses_above_scratch():
    from _topo import make_topology
    n = 260000
    big = make_topology(np.zeros(n), np.full(n, 1.5), residue_pointers=(0,))
    lib = _abi.load_library()
    c = rst._ctx(big)
    d = C.c_void_p()
    assert lib.mmpbsa_device_alloc(0, 16, C.byref(d)) == 0
    try:
        rc = lib.mmpbsa_gb_energy_noalloc(c.h, d, C.byref(rst.GbParams()._c), None, d)
        assert rc == _abi.MMPBSA_ERR_TOO_LARGE
        assert b"scratch" in lib.mmpbsa_last_error(c.h)
    finally:
        lib.mmpbsa_device_free(d)
