"""The SES exterior flood fill on the GPU around an enclosed cavity.

A hollow shell of atoms around one charged atom: closed, the cavity is not
reachable by the probe and is solute (eps_in); with a cap removed, the
probe enters and the cavity is solvent (eps_out).  Both PB energies must
match the FP64 oracle (its flood fill is pinned against a naive BFS) within
0.01 kcal/mol with the oracle's cycle counts, and the two shells must differ
by far more than that, so a flood that leaked into the closed cavity or
missed the open one would fail.
"""
import numpy as np
import pytest

import rust_simulation_tools as rst

from _topo import hollow_shell as shell

pytestmark = pytest.mark.gpu

TOL = 0.01
PB = dict(grid_spacing=0.5, solute_dielectric=4.0, solvent_dielectric=80.0,
          salt_concentration=0.15, fine_dims=(65, 65, 65))


@pytest.mark.parametrize("open_cap", [False, True])
def test_shell_pb_energy(oracle, open_cap):
    top, x = shell(open_cap)
    x = (x / 10.0) * 10.0  # Angstrom coordinates exactly as the library sees them
    r = rst.compute_pb_energy(top, x / 10.0, rst.PbParams(**PB))
    o = oracle.pb_energy(top, x, oracle.pb_params(**PB))
    print(f"open={open_cap}: dG_PB {r.total:.6f} vs {o.total:.6f}; cycles {r.iterations} vs "
          f"({o.solv.iterations}, {o.ref.iterations})")
    assert r.iterations == (o.solv.iterations, o.ref.iterations)
    assert abs(r.e_solv - o.solv.energy) < TOL
    assert abs(r.e_ref - o.ref.energy) < TOL
    assert abs(r.total - o.total) < TOL


def test_cavity_matters():
    totals = []
    for open_cap in (False, True):
        top, x = shell(open_cap)
        totals.append(rst.compute_pb_energy(top, x / 10.0, rst.PbParams(**PB)).total)
    print(f"closed {totals[0]:.4f}, open {totals[1]:.4f}")
    assert abs(totals[0] - totals[1]) > 1.0
