"""Trajectory ingest (host only): the DCD reader against a fixture cut from the
reference's data/trajectory.dcd and decoded independently with numpy, DCD and
mdcrd write/read round trips (byte orders, unit cells, boxes), the
reference's error messages, and reimage_complex's closed-form cases."""
import os
import struct

import numpy as np
import pytest

import rust_simulation_tools as rst

HERE = os.path.dirname(os.path.abspath(__file__))
DCD = os.path.join(HERE, "golden", "trajectory_2frames.dcd")
EXP = os.path.join(HERE, "golden", "trajectory_2frames_expected.npz")


def test_dcd_fixture_matches_independent_decode():
    e = np.load(EXP)
    h = rst.read_dcd_header(DCD)
    assert h["n_frames"] == 2 and h["n_atoms"] == int(e["n_atoms"]) == 12209
    assert h["has_unit_cell"] and h["is_charmm"] and not h["is_big_endian"]
    assert h["first_frame_offset"] == int(e["first_frame_offset"])
    assert h["frame_size"] == int(e["frame_size"])
    assert len(h["titles"]) == int(e["n_titles"])
    r = rst.DcdReader(DCD)
    traj, box = r.read_all()
    assert traj.shape == (2, 12209, 3) and box.shape == (2, 6)
    assert np.array_equal(traj[:, :200], e["head"]) and np.array_equal(traj[:, -200:], e["tail"])
    assert np.array_equal(traj.sum(axis=1), e["sums"])
    assert np.array_equal(box, e["cells"])
    assert r.read_frame() is None                         # at the end
    x, b = r.read_frame_at(1)
    assert np.array_equal(x, traj[1]) and r.current_frame == 2
    with pytest.raises(OSError, match="out of range"):
        r.seek(2)


def _write_dcd(path, frames_A, cells=None, big=False):
    """A CHARMM-style DCD (positions in A, optional unit cells a, b, c,
    alpha, beta, gamma), record markers in the given byte order; coordinates
    and cells little-endian as the reader expects."""
    e = ">" if big else "<"
    nf, n, _ = frames_A.shape
    hdr = b"CORD" + struct.pack(e + "9i", nf, 0, 1, nf, 0, 0, 0, 0, 0)
    hdr += struct.pack("<f", 0.5) + struct.pack(e + "i", 1 if cells is not None else 0)
    hdr += b"\0" * 32 + struct.pack(e + "i", 24)
    out = struct.pack(e + "i", 84) + hdr + struct.pack(e + "i", 84)
    title = b"synthetic".ljust(80)
    out += struct.pack(e + "2i", 84, 1) + title + struct.pack(e + "i", 84)
    out += struct.pack(e + "3i", 4, n, 4)
    for f in range(nf):
        if cells is not None:
            a, b, c, al, be, g = cells[f]
            out += struct.pack(e + "i", 48) + struct.pack("<6d", a, g, b, be, al, c) + struct.pack(e + "i", 48)
        for d in range(3):
            out += struct.pack(e + "i", 4 * n) + frames_A[f, :, d].astype("<f4").tobytes()
            out += struct.pack(e + "i", 4 * n)
    open(path, "wb").write(out)


@pytest.mark.parametrize("big", [False, True])
def test_dcd_round_trip(tmp_path, big):
    rng = np.random.default_rng(5)
    fr = rng.uniform(-30, 30, (3, 17, 3))
    cells = [[40.0 + k, 41.0, 42.0, 90.0, 91.0, 92.0] for k in range(3)]
    p = tmp_path / "t.dcd"
    _write_dcd(p, fr, cells, big)
    h = rst.read_dcd_header(p)
    assert h["is_big_endian"] == big and h["n_frames"] == 3 and h["titles"] == ["synthetic"]
    traj, box = rst.DcdReader(p).read_all()
    assert np.array_equal(traj, fr.astype(np.float32).astype(np.float64) * 0.1)
    assert np.array_equal(box[:, :3], np.array(cells)[:, :3] * 0.1)
    assert np.array_equal(box[:, 3:], np.array(cells)[:, 3:])


def test_dcd_errors(tmp_path):
    with pytest.raises(OSError, match="Failed to open DCD file"):
        rst.DcdReader(tmp_path / "none.dcd")
    p = tmp_path / "bad.dcd"
    p.write_bytes(struct.pack("<i", 12) + b"x" * 100)
    with pytest.raises(OSError, match="first block size is 12"):
        rst.DcdReader(p)


def _write_mdcrd(path, frames_A, boxes=None):
    lines = ["title line"]
    for f in range(frames_A.shape[0]):
        v = list(frames_A[f].ravel())
        for k in range(0, len(v), 10):
            lines.append("".join(f"{x:8.3f}" for x in v[k:k + 10]))
        if boxes is not None:
            lines.append("".join(f"{x:8.3f}" for x in boxes[f]))
    open(path, "w").write("\n".join(lines) + "\n")


def test_mdcrd_round_trip_and_errors(tmp_path):
    rng = np.random.default_rng(7)
    fr = np.round(rng.uniform(-99, 99, (2, 13, 3)), 3)
    p = tmp_path / "t.mdcrd"
    _write_mdcrd(p, fr, [[50.0, 51.0, 52.0]] * 2)
    r = rst.MdcrdReader(p, 13, True)
    for f in range(2):
        assert np.array_equal(r.read_frame(), fr[f])
    assert r.read_frame() is None and r.current_frame == 2
    q = tmp_path / "short.mdcrd"
    _write_mdcrd(q, fr)
    # 2 x 39 values on lines of 10, 10, 10, 9: the first 60-value frame ends
    # one value into the seventh line, whose other 9 values are dropped (the
    # reference reads whole lines); the eighth line leaves 9 of 60
    short = rst.MdcrdReader(q, 20, False)
    assert short.read_frame().shape == (20, 3)
    with pytest.raises(OSError, match=r"Unexpected EOF at frame 1 \(got 9/60 values\)"):
        short.read_frame()
    with pytest.raises(OSError, match="Failed to open mdcrd"):
        rst.MdcrdReader(tmp_path / "none.mdcrd", 3, False)


def test_reimage_closed_forms():
    box = np.array([20.0, 20.0, 20.0])
    # a 3-atom chain wrapped across x = 20: unwrapped to be contiguous with atom 0
    x = np.array([[19.5, 5.0, 5.0], [0.3, 5.0, 5.0], [1.1, 5.0, 5.0], [15.0, 10.0, 10.0]])
    bonds = np.array([[0, 1], [1, 2]])
    y = rst.reimage(x, bonds, box)
    assert np.allclose(y[:3, 0], [19.5, 20.3, 21.1]) and np.array_equal(y[3], x[3])
    # a lone atom more than half a box from that chain's centre moves by a box
    x[3, 0] = 10.0
    assert rst.reimage(x, bonds, box)[3, 0] == 30.0
    # a second component far across the box moves by one box vector next to the first
    x2 = np.array([[1.0, 1.0, 1.0], [2.0, 1.0, 1.0], [18.0, 1.5, 1.0], [19.0, 1.5, 1.0]])
    y2 = rst.reimage(x2, np.array([[0, 1], [2, 3]]), box)
    assert np.allclose(y2[2:, 0], [-2.0, -1.0]) and np.array_equal(y2[:2], x2[:2])
    # single component, nothing to move; no bonds: each atom its own component
    assert np.array_equal(rst.reimage(x2[:2], np.array([[0, 1]]), box), x2[:2])


def test_reimage_matches_oracle_on_wrapped_complex(oracle):
    """The library's reimage_complex equals the oracle's restatement bit for
    bit on a 700-atom complex wrapped into a 40 A box, and both return the
    unwrapped complex up to one rigid box translation."""
    from rust_simulation_tools import synthetic
    top, x, _, _ = synthetic.build_complex(700, 64, None, synthetic.seed_for(5, 0))
    L = 40.0
    box = np.array([L, L, L])
    xc = x - 0.5 * (x.min(0) + x.max(0))      # centred on a box corner: chains cross faces
    wrapped = np.mod(xc, L)
    assert np.ptp(wrapped, axis=0).min() > 0.9 * L
    bonds = np.asarray(top.bonds)[:, :2]
    got = rst.reimage(wrapped, bonds, box)
    want = oracle.reimage(wrapped, bonds, box)
    assert np.array_equal(got, want)
    shift = got - xc
    k = np.round(shift / L)
    assert np.all(np.abs(shift - k * L) < 1e-9)
    assert np.all(k == k[0])                     # one rigid translation for every atom
