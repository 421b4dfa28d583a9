"""Solute stripping and the GB radii cache on the GPU.

- compute_pb_energy_solvated (pb_energy.rs:442-482) and the reference's
  tests/test_mmpbsa.py:419-458 (test_solvated_matches_stripped): a topology
  with residues outside the receptor and ligand gives the same binding
  energy as the stripped complex, and a PB energy with solute_residues the
  same as the stripped topology's (and the oracle's within 0.01 kcal/mol).
- GbCache / compute_gb_energy_cached (gb_energy.rs:66-120, :372-387): a
  miss computes radii equal to the oracle's (1e-5 relative); a frame within
  the RMSD threshold reuses the stored radii unchanged, and its energy is
  compute_gb_energy_from_radii (gb_energy.rs:283-355) of those radii at the
  new coordinates; a frame beyond the threshold computes and stores again.
"""
import ctypes as C

import numpy as np
import pytest

import rust_simulation_tools as rst
from rust_simulation_tools import synthetic
from rust_simulation_tools.topology import extract_subtopology

pytestmark = pytest.mark.gpu

TOL = 0.01
PB = dict(grid_spacing=1.0, grid_buffer=4.0, solute_dielectric=4.0)


@pytest.fixture(scope="module")
def solvated():
    # 50 residues of 16 atoms; the first 30 are the solute (receptor 0-14,
    # ligand 15-29), the other 20 play the solvent
    top, x, _, _ = synthetic.build_complex(800, 64, seed=103)
    x_nm = x / 10.0
    n_test = 30
    rec, lig = list(range(0, n_test // 2)), list(range(n_test // 2, n_test))
    sel = top.build_selection(list(range(n_test)))
    stripped = extract_subtopology(top, sel)
    return top, x_nm, rec, lig, stripped, x_nm[sel]


@pytest.mark.parametrize("method", ["gb", "pb"])
def test_solvated_matches_stripped(solvated, method):
    top, x_nm, rec, lig, stripped, xs_nm = solvated
    kw = dict(pb_params=rst.PbParams(**PB)) if method == "pb" else {}
    f_solv = rst.compute_binding_energy_single_frame(top, x_nm, rec, lig, **kw)
    f_strip = rst.compute_binding_energy_single_frame(stripped, xs_nm, rec, lig, **kw)
    for k in ("delta_total", "delta_mm", "delta_polar", "delta_sa", "complex_polar",
              "receptor_polar", "ligand_polar"):
        assert abs(getattr(f_solv, k) - getattr(f_strip, k)) <= 1e-4, k


def test_pb_solute_residues(solvated, oracle):
    top, x_nm, rec, lig, stripped, xs_nm = solvated
    solute = rec + lig
    r = rst.compute_pb_energy(top, x_nm, rst.PbParams(**PB), solute_residues=solute)
    s = rst.compute_pb_energy(stripped, xs_nm, rst.PbParams(**PB))
    assert (r.total, r.e_solv, r.e_ref, r.iterations) == (s.total, s.e_solv, s.e_ref, s.iterations)
    o = oracle.pb_energy(stripped, xs_nm * 10.0, oracle.pb_params(**PB))
    assert abs(r.total - o.total) < TOL and abs(r.e_solv - o.solv.energy) < TOL
    # the full topology without stripping is a different (larger) solute
    full = rst.compute_pb_energy(top, x_nm, rst.PbParams(**PB))
    assert abs(full.total - r.total) > 1.0
    # a second call with the same residues reuses the resident solute
    again = rst.compute_pb_energy(top, x_nm, rst.PbParams(**PB), solute_residues=solute)
    assert again.total == r.total
    with pytest.raises(ValueError, match="out of range"):
        rst.compute_pb_energy(top, x_nm, rst.PbParams(**PB), solute_residues=[0, 50])


def test_gb_cache_and_energy_from_radii(cfg0, oracle):
    top, x, _, _, seed = cfg0
    x0 = (x / 10.0) * 10.0
    gp = dict(model=2)
    cache = rst.GbCache(0.5)
    e0 = rst.compute_gb_energy_cached(top, x0 / 10.0, rst.GbParams(**gp), cache)
    assert not cache.last_reused
    oe0, oR0 = oracle.gb_energy(top, x0, oracle.gb_params(**gp))
    assert np.max(np.abs(e0.born_radii() - oR0) / oR0) < 1e-5 and abs(e0.total - oe0) < TOL
    # within the threshold (RMSD 0.17 A): stored radii, energy from them
    rng = np.random.default_rng(4)
    x1 = ((x0 + rng.normal(0, 0.1, x0.shape)) / 10.0) * 10.0
    e1 = rst.compute_gb_energy_cached(top, x1 / 10.0, rst.GbParams(**gp), cache)
    assert cache.last_reused
    assert np.array_equal(e1.born_radii(), e0.born_radii())
    L = oracle.lib()
    v = top.view()
    R0 = np.ascontiguousarray(e0.born_radii())
    x1c = np.ascontiguousarray(x1)
    oe1 = L.orc_gb_energy_from_radii(C.byref(v), x1c.ctypes.data_as(C.POINTER(C.c_double)),
                                     C.byref(oracle.gb_params(**gp)),
                                     R0.ctypes.data_as(C.POINTER(C.c_double)))
    assert abs(e1.total - oe1) < TOL
    direct = rst.compute_gb_energy_from_radii(top, x1 / 10.0, e0.born_radii(), rst.GbParams(**gp))
    assert direct == e1.total
    # beyond the threshold (RMSD 1.2 A): recomputed and stored
    x2 = ((x0 + rng.normal(0, 0.7, x0.shape)) / 10.0) * 10.0
    e2 = rst.compute_gb_energy_cached(top, x2 / 10.0, rst.GbParams(**gp), cache)
    assert not cache.last_reused
    oe2, oR2 = oracle.gb_energy(top, x2, oracle.gb_params(**gp))
    assert np.max(np.abs(e2.born_radii() - oR2) / oR2) < 1e-5 and abs(e2.total - oe2) < TOL
    e3 = rst.compute_gb_energy_cached(top, x2 / 10.0, rst.GbParams(**gp), cache)
    assert cache.last_reused and np.array_equal(e3.born_radii(), e2.born_radii())
    # a cache holding another atom count never hits
    t_small, xs, _, _ = synthetic.build_complex(100, 16, seed=5)
    rst.compute_gb_energy_cached(t_small, xs / 10.0, rst.GbParams(**gp), cache)
    assert not cache.last_reused
