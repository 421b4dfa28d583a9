"""Face dielectrics in two launches (the faces with both nodes on one side per
node, then a wavefront per mixed edge) against the one-thread-per-node pass
over every face.  MMPBSA_FACES_LIST_CAP caps the mixed-edge list: at 0 every
map overflows it and takes the per-node pass, at 2,000 the coarse and fine
maps of the 2,500-atom complex overflow it in part.  The PB energies and
binding terms must equal the default run's bit for bit: both paths clip the
same atom intervals in the same scan order and merge them in the same order."""
import numpy as np
import pytest

import rust_simulation_tools as rst

pytestmark = pytest.mark.gpu

PB = dict(grid_spacing=0.6, solute_dielectric=4.0, solvent_dielectric=80.0,
          fine_dims=(65, 65, 65))


def _run(top, x_nm, rec, lig):
    e = rst.compute_pb_energy(top, x_nm, rst.PbParams(**PB))
    f = rst.compute_binding_energy_single_frame(top, x_nm, rec, lig,
                                                pb_params=rst.PbParams(**PB))
    return (e.total, e.iterations, f.complex_polar, f.receptor_polar, f.ligand_polar,
            f.delta_total)


@pytest.mark.parametrize("cap", ["0", "2000"])
def test_split_faces_equal_per_node_faces(cfg0, monkeypatch, cap):
    top, x, rec, lig, _ = cfg0
    x_nm = np.asarray(x) / 10.0
    monkeypatch.delenv("MMPBSA_FACES_LIST_CAP", raising=False)
    split = _run(top, x_nm, rec, lig)
    monkeypatch.setenv("MMPBSA_FACES_LIST_CAP", cap)
    node = _run(top, x_nm, rec, lig)
    print(f"cap {cap}: split {split!r}\n        per node {node!r}")
    assert split == node
