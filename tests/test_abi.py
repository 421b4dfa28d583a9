"""The C-ABI library loads on the CPU and exports every symbol that
include/mmpbsa_hip.h declares (no compute call without a GPU)."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from rust_simulation_tools import _abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "mmpbsa_hip.h")


def declared_symbols():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(mmpbsa_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_header_symbols():
    lib = _abi.load_library()
    syms = declared_symbols()
    assert len(syms) >= 19
    for s in syms:
        assert hasattr(lib, s), s
    assert set(syms) == set(_abi.SIGNATURES), set(syms) ^ set(_abi.SIGNATURES)


def test_struct_layouts():
    # field offsets the C compiler would use (natural alignment on x86-64)
    assert C.sizeof(_abi.GbParamsC) == 64
    assert C.sizeof(_abi.SaParamsC) == 32
    assert C.sizeof(_abi.MmTermsC) == 56
    assert _abi.PbParamsC.fine_dims.offset == 96
    assert C.sizeof(_abi.FrameTermsC) == 9 * 8 + 4 * 8


def test_defaults_match_reference():
    import rust_simulation_tools as rst
    g = rst.GbParams()
    assert (g._c.model, g.solute_dielectric, g.solvent_dielectric, g.salt_concentration,
            g.temperature, g._c.offset, g._c.rgbmax, g.cutoff) == (1, 1.0, 78.3, 0.15, 300.0,
                                                                    0.09, 25.0, 9999.0)
    p = rst.PbParams()
    assert (p.grid_spacing, p._c.grid_buffer, p.solute_dielectric, p.solvent_dielectric,
            p.salt_concentration, p._c.probe_radius, p._c.ion_radius, p._c.tolerance,
            p._c.max_iterations, p._c.fillratio, p._c.fscale) == (0.5, 10.0, 1.0, 80.0, 0.15,
                                                                  1.4, 2.0, 1e-6, 10000, 4.0, 8)
    s = rst.SaParams()
    assert (s.surface_tension, s.offset, s.probe_radius, s.n_sphere_points) == (0.0072, 0.0, 1.4,
                                                                                480)


def test_errors_without_device_or_bad_input():
    import rust_simulation_tools as rst
    from _topo import make_topology
    top = make_topology([1.0], [1.5])
    with pytest.raises((RuntimeError, ValueError)):
        rst.compute_gb_energy(top, np.zeros((2, 3)))   # wrong atom count
    lib = _abi.load_library()
    h = C.c_void_p()
    rc = lib.mmpbsa_ctx_create(None, 0, C.byref(h))
    assert rc == 1  # invalid argument


def test_bad_residue_is_value_error():
    from rust_simulation_tools.topology import AmberTopology
    from _topo import make_topology
    top = make_topology([0.0, 0.0], [1.5, 1.5], residue_pointers=(0, 1))
    with pytest.raises(ValueError):
        top.build_selection([5])
