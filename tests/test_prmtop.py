"""Topology ingest: the library's prmtop reader (prmtop.rs:436-733 rules) on a
hand-written file with the edge cases, a write/read round trip of a synthetic
topology, and the reference's error messages.  Host only, no device."""
import os

import numpy as np
import pytest

import rust_simulation_tools as rst
from rust_simulation_tools import synthetic

HERE = os.path.dirname(os.path.abspath(__file__))
SMALL = os.path.join(HERE, "golden", "small.prmtop")


def test_small_file_rules():
    t = rst.read_prmtop(SMALL)
    assert (t.n_atoms, t.n_residues, t.n_types) == (5, 2, 2)
    assert t.atom_names == ["N", "H1", "CA", "CB", "O"]
    assert t.residue_labels == ["ALA", "GLY"]
    assert list(t.residue_pointers) == [0, 3] and list(t.residue_pointers()) == [0, 3]
    assert list(t.atom_type_indices) == [0, 1, 0, 0, 1]
    amber = np.array([-5.46668999, 3.64445999, 1.82223, -0.911115, 0.911115])
    assert np.array_equal(t.charges_amber, amber)
    assert np.array_equal(t.charges(), amber / 18.2223)
    # bonds: H section first; a negative index and a type < 1 are dropped; /3
    assert t.bonds.tolist() == [[0, 1, 1], [0, 2, 0], [2, 3, 0]]
    assert t.angles.tolist() == [[1, 0, 2, 0], [0, 2, 3, 0]]
    # |index| / 3; negative third or fourth index -> ignore_14; type < 1 dropped
    assert t.dihedrals.tolist() == [[1, 0, 2, 3, 0, 0], [0, 2, 3, 4, 1, 1], [0, 2, 3, 4, 0, 1]]
    assert list(t.num_excluded_atoms) == [3, 2, 2, 1, 1]
    assert list(t.excluded_atoms_list) == [1, 2, 3, 2, 3, 3, 4, 4, -1]
    assert (t.scee_scale_factor, t.scnb_scale_factor) == (1.2, 2.0)
    assert np.array_equal(t.radii, np.zeros(5))          # no RADII section
    assert np.allclose(t.screen, [0.79, 0.85, 0.72, 0.72, 0.85])
    assert np.allclose(t.masses, [14.01, 1.008, 12.01, 12.01, 16.0])
    assert list(t.nb_parm_index) == [1, 2, 2, 3]
    assert t.bond_force_constants.tolist() == [300.0, 434.0]
    # per-type LJ from the self terms: sigma = (A/B)^(1/6) nm/10, eps = B^2/4A kJ
    a, b = 9.44293233e5, 8.01323529e2
    assert t.sigmas()[0] == pytest.approx((a / b) ** (1 / 6) * 0.1, rel=1e-15)
    assert t.epsilons()[0] == pytest.approx(b * b / (4 * a) * 4.184, rel=1e-15)


def _write(path, top):
    def block(flag, fmt, vals, kind):
        lines = [f"%FLAG {flag}", f"%FORMAT({fmt})"]
        per = {"i": 10, "f": 5, "s": 20}[kind]
        vals = list(vals)
        for k in range(0, max(len(vals), 1), per):
            chunk = vals[k:k + per]
            if kind == "i":
                lines.append("".join(f"{int(v):8d}" for v in chunk))
            elif kind == "f":
                lines.append(" ".join(repr(float(v)) for v in chunk))
            else:
                lines.append("".join(f"{str(v)[:4]:<4s}" for v in chunk))
        return lines

    n, nr, nt = top.n_atoms, top.n_residues, int(top.n_types)
    ptr = [n, nt, 0, 0, 0, 0, 0, 0, 0, 0, len(top.excluded_atoms_list), nr] + [0] * 19
    out = ["%VERSION  VERSION_STAMP = V0001.000"]
    out += block("POINTERS", "10I8", ptr, "i")
    out += block("ATOM_NAME", "20a4", top.atom_names, "s")
    out += block("CHARGE", "5E16.8", top.charges_amber, "f")
    out += block("ATOM_TYPE_INDEX", "10I8", top.atom_type_indices + 1, "i")
    out += block("NUMBER_EXCLUDED_ATOMS", "10I8", top.num_excluded_atoms, "i")
    out += block("EXCLUDED_ATOMS_LIST", "10I8", top.excluded_atoms_list + 1, "i")
    out += block("NONBONDED_PARM_INDEX", "10I8", top.nb_parm_index, "i")
    out += block("RESIDUE_LABEL", "20a4", top.residue_labels, "s")
    out += block("RESIDUE_POINTER", "10I8", top.residue_pointers + 1, "i")
    for flag, v in (("LENNARD_JONES_ACOEF", top.lj_acoef), ("LENNARD_JONES_BCOEF", top.lj_bcoef),
                    ("RADII", top.radii), ("SCREEN", top.screen), ("MASS", top.masses),
                    ("BOND_FORCE_CONSTANT", top.bond_force_constants),
                    ("BOND_EQUIL_VALUE", top.bond_equil_values),
                    ("ANGLE_FORCE_CONSTANT", top.angle_force_constants),
                    ("ANGLE_EQUIL_VALUE", top.angle_equil_values),
                    ("DIHEDRAL_FORCE_CONSTANT", top.dihedral_force_constants),
                    ("DIHEDRAL_PERIODICITY", top.dihedral_periodicities),
                    ("DIHEDRAL_PHASE", top.dihedral_phases),
                    ("SCEE_SCALE_FACTOR", [top.scee_scale_factor]),
                    ("SCNB_SCALE_FACTOR", [top.scnb_scale_factor])):
        out += block(flag, "5E16.8", v, "f")
    b = top.bonds.copy()
    out += block("BONDS_WITHOUT_HYDROGEN", "10I8",
                 np.stack([b[:, 0] * 3, b[:, 1] * 3, b[:, 2] + 1], 1).ravel(), "i")
    a = top.angles
    out += block("ANGLES_WITHOUT_HYDROGEN", "10I8",
                 np.stack([a[:, 0] * 3, a[:, 1] * 3, a[:, 2] * 3, a[:, 3] + 1], 1).ravel(), "i")
    d = top.dihedrals
    sign = np.where(d[:, 5] != 0, -1, 1)
    out += block("DIHEDRALS_WITHOUT_HYDROGEN", "10I8",
                 np.stack([d[:, 0] * 3, d[:, 1] * 3, d[:, 2] * 3, sign * d[:, 3] * 3, d[:, 4] + 1],
                          1).ravel(), "i")
    with open(path, "w") as fh:
        fh.write("\n".join(out) + "\n")


def test_round_trip_synthetic(tmp_path):
    top, _, _, _ = synthetic.build_complex(600, 48, None, 99)
    p = tmp_path / "syn.prmtop"
    _write(p, top)
    t = rst.read_prmtop(p)
    for k in ("charges_amber", "radii", "screen", "atom_type_indices", "nb_parm_index",
              "lj_acoef", "lj_bcoef", "residue_pointers", "bonds", "angles",
              "num_excluded_atoms", "excluded_atoms_list", "bond_force_constants",
              "dihedral_phases"):
        assert np.array_equal(np.asarray(getattr(t, k)), np.asarray(getattr(top, k))), k
    # dihedrals round-trip except where the fourth index is 0 (no sign to carry
    # ignore_14), which the synthetic chains never produce
    assert np.array_equal(t.dihedrals, top.dihedrals)
    assert np.array_equal(t.charges, top.charges_amber / 18.2223)
    assert t.residue_labels == [s[:4] for s in top.residue_labels]


@pytest.mark.parametrize("text, msg", [
    ("%FLAG TITLE\nx\n", "Missing section: POINTERS"),
    ("%FLAG POINTERS\n1 2 3\n", "POINTERS section too short"),
    ("%FLAG POINTERS\n" + " ".join(["1"] * 12).replace("1", "x", 1) + "\n",
     "Failed to parse integer in POINTERS"),
    ("%FLAG POINTERS\n-1 1 0 0 0 0 0 0 0 0 0 1\n", "POINTERS contains negative values"),
])
def test_errors(tmp_path, text, msg):
    p = tmp_path / "bad.prmtop"
    p.write_text(text)
    with pytest.raises(OSError, match=msg):
        rst.read_prmtop(p)
    with pytest.raises(OSError, match="Failed to open prmtop file"):
        rst.read_prmtop(tmp_path / "missing.prmtop")
