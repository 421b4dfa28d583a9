"""Frame MM and GB terms when the complex is exactly the two partners: the
complex's pair terms come from the receptor's and ligand's own passes plus the
receptor x ligand pairs, and its Born radii from the partners' psi sums plus
the cross sums, so its intra-partner pairs are evaluated once.  Each
subsystem's frame MM term must equal that subsystem's own full pass
(compute_mm_energy on the extracted topology, every pair of it summed
directly) to a few FP64 ulps, also on a frame whose MM total is ~1e14
kcal/mol; each GB term must equal compute_gb_energy on the extracted topology
to rounding (a complex atom's psi adds the same pair terms in another order).
A split with a bond between the partners must take the full MM pass, which a
sum of the partners would miss the bond of."""
import numpy as np
import pytest

import rust_simulation_tools as rst
from rust_simulation_tools import synthetic
from rust_simulation_tools.topology import extract_subtopology

pytestmark = pytest.mark.gpu


def _direct(top, x_nm, rec, lig):
    ra, la = top.build_selection(rec), top.build_selection(lig)
    ca = np.sort(np.concatenate([ra, la]))
    return [rst.compute_mm_energy(extract_subtopology(top, sel), x_nm[sel]).total
            for sel in (ca, ra, la)]


def _check(top, x_nm, rec, lig, ulps):
    f = rst.compute_binding_energy_single_frame(top, x_nm, rec, lig)
    got = (f.complex_mm, f.receptor_mm, f.ligand_mm)
    want = _direct(top, x_nm, rec, lig)
    for name, g, w in zip(("complex", "receptor", "ligand"), got, want):
        err = abs(g - w) / np.spacing(abs(w))
        print(f"{name}: frame {g!r} direct {w!r} ({err:.1f} ulp)")
        assert err <= ulps, name
    return got, want


def test_partner_sums_2500_atoms(cfg0):
    top, x, rec, lig, _ = cfg0
    _check(top, np.asarray(x) / 10.0, rec, lig, ulps=4)


def test_partner_sums_overlapping_frame():
    # 6,660 atoms (6,404 + 256), frame 34 of the Poisson-serving draw: a few
    # pairs within ~0.1 A put the complex and receptor MM totals near 2.7e14
    seed = synthetic.seed_for(3, 7)
    top, base, rec, lig = synthetic.build_complex(6660, 256, None, seed)
    x = synthetic.noisy_frame(base, seed, 34) / 10.0
    got, want = _check(top, x, rec, lig, ulps=4)
    assert abs(want[0]) > 1e13
    # the delta term, where the intra-partner pairs cancel: within the
    # rounding of the two large totals it is formed from
    d_got = got[0] - got[1] - got[2]
    d_want = want[0] - want[1] - want[2]
    assert abs(d_got - d_want) <= 8 * np.spacing(abs(want[0]))


def test_bond_across_partners_takes_full_pass(cfg0):
    top, x, _, _, _ = cfg0
    # residues 0-9 against residue 10: the synthetic chains bond consecutive
    # residues of a segment
    rp = np.append(np.asarray(top.residue_pointers), top.n_atoms)
    res_of = np.searchsorted(rp, np.arange(top.n_atoms), side="right") - 1
    b = np.asarray(top.bonds)
    assert any(res_of[i] == 9 and res_of[j] == 10 for i, j, _ in b)
    _check(top, np.asarray(x) / 10.0, list(range(10)), [10], ulps=4)


@pytest.mark.parametrize("model", [0, 1, 2])
def test_gb_from_partner_psi(cfg0, model):
    top, x, rec, lig, _ = cfg0
    x_nm = np.asarray(x) / 10.0
    gb = rst.GbParams(model=model)
    f = rst.compute_binding_energy_single_frame(top, x_nm, rec, lig, gb_params=gb)
    ra, la = top.build_selection(rec), top.build_selection(lig)
    ca = np.sort(np.concatenate([ra, la]))
    for name, got, sel in (("complex", f.complex_polar, ca), ("receptor", f.receptor_polar, ra),
                           ("ligand", f.ligand_polar, la)):
        want = rst.compute_gb_energy(extract_subtopology(top, sel), x_nm[sel], gb).total
        print(f"model {model} {name}: frame {got!r} direct {want!r}")
        assert abs(got - want) <= 1e-10 * abs(want), name


def test_gb_from_partner_psi_with_cutoff(cfg0):
    top, x, rec, lig, _ = cfg0
    x_nm = np.asarray(x) / 10.0
    gb = rst.GbParams(model=2, cutoff=12.0, salt_concentration=0.0)
    f = rst.compute_binding_energy_single_frame(top, x_nm, rec, lig, gb_params=gb)
    ca = np.sort(np.concatenate([top.build_selection(rec), top.build_selection(lig)]))
    want = rst.compute_gb_energy(extract_subtopology(top, ca), x_nm[ca], gb).total
    assert abs(f.complex_polar - want) <= 1e-10 * abs(want)
