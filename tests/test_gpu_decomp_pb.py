"""Per-residue PB terms and explicit coarse-to-fine focusing grids on the GPU
against the FP64 oracle (the large-decomposition workload at a size the
oracle finishes in seconds: 1,200 atoms, a 33^3 coarse grid at 2 A centred on
the complex, a 65^3 fine grid at 0.5 A centred on the interface).

Bars: each residue's PB term and dG_PB within 0.01 kcal/mol of the oracle;
the residue terms sum to the library's own dG_PB; vdW, elec and SA columns
identical to the GB decomposition's; PB cycle counts equal to the oracle's.
"""
import numpy as np
import pytest

import rust_simulation_tools as rst
from rust_simulation_tools import synthetic

pytestmark = pytest.mark.gpu

TOL = 0.01
PB_BASE = dict(solute_dielectric=4.0, solvent_dielectric=80.0, salt_concentration=0.15)


@pytest.fixture(scope="module")
def system():
    top, x, rec, lig = synthetic.build_complex(1200, 120, seed=synthetic.seed_for(4, 7))
    ra, la = top.build_selection(rec), top.build_selection(lig)
    # Angstrom coordinates exactly as the library sees them from nm input
    x = (x / 10.0) * 10.0
    grids = synthetic.focusing_grids(x, ra, la, coarse_dims=33, coarse_spacing=2.0,
                                     fine_dims=65, fine_spacing=0.5)
    return top, x, rec, lig, dict(PB_BASE, **grids)


def test_focusing_grids_pb_energy(system, oracle):
    top, x, _, _, kw = system
    r = rst.compute_pb_energy(top, x / 10.0, rst.PbParams(**kw))
    o = oracle.pb_energy(top, x, oracle.pb_params(**kw))
    print(f"dG_PB {r.total:.6f} vs {o.total:.6f}; cycles {r.iterations} vs "
          f"({o.solv.iterations}, {o.ref.iterations})")
    assert r.fine_dims == (65, 65, 65) == tuple(o.fine_dims)
    assert r.coarse_dims == (33, 33, 33) == tuple(o.coarse_dims)
    assert r.iterations == (o.solv.iterations, o.ref.iterations)
    assert abs(r.e_solv - o.solv.energy) < TOL and abs(r.e_ref - o.ref.energy) < TOL
    assert abs(r.total - o.total) < TOL


def test_pb_residue_terms(system, oracle):
    top, x, rec, lig, kw = system
    d, dpb = rst.decompose_binding_energy_pb(top, x / 10.0, rec, lig,
                                             pb_params=rst.PbParams(**kw))
    orr, oll, od = oracle.pb_residue_terms(top, x, rec, lig, oracle.pb_params(**kw))
    got_r = np.array([c.polar for c in d.receptor_residues()])
    got_l = np.array([c.polar for c in d.ligand_residues()])
    assert [c.residue_index for c in d.receptor_residues()] == [int(r) for r in rec]
    assert [c.residue_index for c in d.ligand_residues()] == [int(r) for r in lig]
    err = max(np.abs(got_r - orr).max(), np.abs(got_l - oll).max())
    print(f"dG_PB {dpb:.6f} vs {od:.6f}; worst residue error {err:.2e}; "
          f"largest term {max(np.abs(orr).max(), np.abs(oll).max()):.3f}")
    assert err < TOL
    assert abs(dpb - od) < TOL
    # the residue terms are a partition of dG_PB
    assert abs(got_r.sum() + got_l.sum() - dpb) < 1e-8 * max(1.0, abs(dpb))
    # vdW / elec / SA come from the same kernels as the GB decomposition
    g = rst.decompose_binding_energy(top, x / 10.0, rec, lig)
    for a, b in zip(d.receptor_residues() + d.ligand_residues(),
                    g.receptor_residues() + g.ligand_residues()):
        assert (a.vdw, a.elec, a.sa) == (b.vdw, b.elec, b.sa)


def test_pb_residue_terms_request_order(system):
    """The order of the requested residues only permutes the output."""
    top, x, rec, lig, kw = system
    p = rst.PbParams(**kw)
    d1, e1 = rst.decompose_binding_energy_pb(top, x / 10.0, rec, lig, pb_params=p)
    d2, e2 = rst.decompose_binding_energy_pb(top, x / 10.0, list(rec)[::-1], lig, pb_params=p)
    a = np.array([c.polar for c in d1.receptor_residues()])
    b = np.array([c.polar for c in d2.receptor_residues()])[::-1]
    assert np.allclose(a, b, rtol=0, atol=1e-6) and abs(e1 - e2) < 1e-6
