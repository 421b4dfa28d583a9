"""compute_binding_energy over trajectory files (DCD, mdcrd) equals the
in-memory batched path on the frames the readers return; stride / start / end
select frames as the reference does; a periodic box triggers reimaging."""
import os
import sys

import numpy as np
import pytest

import rust_simulation_tools as rst
from rust_simulation_tools import synthetic

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
from test_trajectory import _write_dcd, _write_mdcrd  # noqa: E402
from test_prmtop import _write as _write_prmtop  # noqa: E402

pytestmark = pytest.mark.gpu

PB = dict(grid_spacing=0.8, grid_buffer=4.0, solute_dielectric=4.0, solvent_dielectric=80.0,
          salt_concentration=0.15)
KEYS = ("complex_mm", "complex_polar", "receptor_polar", "ligand_polar", "complex_sa",
        "delta_total")


@pytest.fixture(scope="module")
def system():
    seed = synthetic.seed_for(5, 0)
    top, base, rec, lig = synthetic.build_complex(700, 64, None, seed)
    frames = np.stack([synthetic.noisy_frame(base, seed, k) for k in range(5)])  # A
    return top, frames, rec, lig


def test_dcd_trajectory(system, tmp_path):
    top, frames, rec, lig = system
    p = tmp_path / "t.dcd"
    _write_dcd(p, frames, cells=[[500.0, 500.0, 500.0, 90.0, 90.0, 90.0]] * 5)
    pb = rst.PbParams(**PB)
    res = rst.compute_binding_energy(top, p, rec, lig, pb_params=pb, trajectory_format="dcd",
                                     stride=2, start_frame=1)
    traj, _ = rst.DcdReader(p).read_all()  # nm, as the reader returns them
    ref = rst.compute_binding_energy_frames(top, traj[[1, 3]], rec, lig, pb_params=pb)
    assert len(res.frames()) == 2
    for a, b in zip(res.frames(), ref.frames()):
        for k in KEYS:
            assert getattr(a, k) == getattr(b, k), k
    assert res.mean_delta_total == ref.mean_delta_total
    gb = rst.compute_binding_energy(top, p, rec, lig, trajectory_format="dcd", end_frame=2)
    assert len(gb.frames()) == 2


def test_mdcrd_trajectory(system, tmp_path):
    top, frames, rec, lig = system
    p = tmp_path / "t.mdcrd"
    _write_mdcrd(p, np.round(frames[:3], 3))
    res = rst.compute_binding_energy(top, p, rec, lig)          # GB, every frame
    rd = rst.MdcrdReader(p, top.n_atoms, False)
    fr = np.stack([rd.read_frame() for _ in range(3)]) / 10.0
    ref = rst.compute_binding_energy_frames(top, fr, rec, lig)
    for a, b in zip(res.frames(), ref.frames()):
        for k in KEYS:
            # the in-memory path divides by 10 and multiplies back: coordinates
            # differ by an ulp, energies (MM up to 1e7 here) by ~1e-15 relative
            assert getattr(a, k) == pytest.approx(getattr(b, k), rel=1e-12, abs=1e-9), k
    with pytest.raises(ValueError, match="No frames found in trajectory"):
        rst.compute_binding_energy(top, p, rec, lig, start_frame=10)
    with pytest.raises(ValueError, match="Failed to open mdcrd"):
        rst.compute_binding_energy(top, tmp_path / "none.mdcrd", rec, lig)


def test_dcd_reimaging_in_a_real_box(system, oracle, tmp_path):
    """Frames wrapped into a 40 A box (the 700-atom complex spans ~28 A and
    sits on a box corner, so its chains cross the faces): the trajectory
    driver reimages each frame (binding.rs:234-308, :630-645) before the
    batched PB path.  Its energies equal those of the same frames reimaged
    by the oracle's restatement and run in memory, and the complex
    coordinates it kept of the last frame equal the oracle's bit for bit."""
    import ctypes as C
    top, frames, rec, lig = system
    L = 40.0
    c = 0.5 * (frames[0].min(0) + frames[0].max(0))
    wrapped = np.mod(frames - c, L)
    p = tmp_path / "box.dcd"
    _write_dcd(p, wrapped, cells=[[L, L, L, 90.0, 90.0, 90.0]] * len(frames))
    pb = rst.PbParams(**PB)
    res = rst.compute_binding_energy(top, p, rec, lig, pb_params=pb, trajectory_format="dcd",
                                     stride=2)
    traj, _ = rst.DcdReader(p).read_all()          # nm, float32-rounded positions
    bonds = np.asarray(top.bonds)[:, :2]
    box = np.array([L, L, L])
    fixed = np.stack([oracle.reimage(traj[k] * 10.0, bonds, box) for k in (0, 2, 4)])
    assert np.abs(fixed - traj[[0, 2, 4]] * 10.0).max() > 10.0   # reimaging moved atoms
    ref = rst.compute_binding_energy_frames(top, fixed / 10.0, rec, lig, pb_params=pb)
    assert len(res.frames()) == 3
    for a, b in zip(res.frames(), ref.frames()):
        for k in KEYS:
            assert getattr(a, k) == pytest.approx(getattr(b, k), rel=1e-12, abs=1e-9), k
    # the complex coordinates the driver kept of the last frame it processed
    s = rst._ctx(top).split(rec, lig)
    lib = rst._abi.load_library()
    out = (rst._abi.FrameTermsC * 3)()
    last = np.zeros((top.n_atoms, 3))
    assert lib.mmpbsa_trajectory_frames(s, out, 3, last.ctypes.data_as(rst._abi.c_double_p)) == 0
    assert np.array_equal(last, fixed[-1])


def test_prmtop_and_dcd_against_oracle(system, oracle, tmp_path):
    """Serving from files end to end: the complex written as a prmtop and read
    back by read_prmtop (the library's parser), its frames as a DCD, through
    compute_binding_energy with GB and with PB; every frame term is within
    0.01 kcal/mol of the FP64 oracle run on the read topology and the frames
    the DCD reader returns."""
    top, frames, rec, lig = system
    pt = tmp_path / "complex.prmtop"
    _write_prmtop(pt, top)
    t = rst.read_prmtop(pt)
    assert t.n_atoms == top.n_atoms and t.n_residues == top.n_residues
    d = tmp_path / "t.dcd"
    _write_dcd(d, frames[:2])
    traj, _ = rst.DcdReader(d).read_all()  # nm
    for method in ("gb", "pb"):
        kw = dict(pb_params=rst.PbParams(**PB)) if method == "pb" else \
            dict(gb_params=rst.GbParams(model=2))
        res = rst.compute_binding_energy(t, d, rec, lig, trajectory_format="dcd", **kw)
        assert len(res.frames()) == 2
        for k, f in enumerate(res.frames()):
            o = oracle.frame_energy(t, traj[k] * 10.0, rec, lig,
                                    method=1 if method == "pb" else 0,
                                    gb=oracle.gb_params(model=2), pb=oracle.pb_params(**PB))
            for part in ("complex", "receptor", "ligand"):
                for term in ("mm", "polar", "sa"):
                    a, b = getattr(f, f"{part}_{term}"), o[part][term]
                    assert abs(a - b) < 0.01, (method, k, part, term, a, b)
            assert abs(f.delta_total - o["delta_total"]) < 0.01, (method, k)
