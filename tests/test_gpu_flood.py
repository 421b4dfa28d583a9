"""The SES exterior flood on bits (32 grid lines per word, a wavefront's scan
along each y and z column of words) against the byte flood (one thread group
per line of bytes), which MMPBSA_FLOOD=bytes selects.  Both reach the one
fixpoint, the open nodes connected to an open boundary node, so the PB
energies and cycle counts must be equal bit for bit: on the closed and the
open hollow shell (a cavity the flood must not enter, and one it must), and on
the 2,500-atom complex with a cubic grid and with an uneven one whose lines
take one, two and three words per lane (y 97, z 161) and two words along x."""
import numpy as np
import pytest

import rust_simulation_tools as rst

from _topo import hollow_shell as shell

pytestmark = pytest.mark.gpu


def _pb(top, x_nm, dims, h):
    p = rst.PbParams(grid_spacing=h, solute_dielectric=4.0, solvent_dielectric=80.0,
                     salt_concentration=0.15, fine_dims=dims)
    e = rst.compute_pb_energy(top, x_nm, p)
    return (e.total, e.e_solv, e.e_ref, e.iterations)


def _case(name, cfg0):
    if name.startswith("shell"):
        top, x = shell(name == "shell_open")
        return top, x / 10.0, (65, 65, 65), 0.5
    top, x, _, _, _ = cfg0
    if name == "cfg0_cubic":
        return top, np.asarray(x) / 10.0, (65, 65, 65), 0.6
    return top, np.asarray(x) / 10.0, (49, 97, 161), 1.0


@pytest.mark.parametrize("name", ["shell_closed", "shell_open", "cfg0_cubic", "cfg0_uneven"])
def test_bit_flood_equals_byte_flood(cfg0, monkeypatch, name):
    top, x_nm, dims, h = _case(name, cfg0)
    monkeypatch.delenv("MMPBSA_FLOOD", raising=False)
    bits = _pb(top, x_nm, dims, h)
    monkeypatch.setenv("MMPBSA_FLOOD", "bytes")
    byts = _pb(top, x_nm, dims, h)
    print(f"{name}: bits {bits!r}\n   bytes {byts!r}")
    assert bits == byts
