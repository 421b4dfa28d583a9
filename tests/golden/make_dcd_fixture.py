"""Writes tests/golden/trajectory_2frames.dcd (the first two frames of the
reference's data/trajectory.dcd, with the header's frame count set to 2) and
tests/golden/trajectory_2frames_expected.npz (header values, unit cells and
positions in nm decoded here with numpy, independently of the library's
reader: the first and last 200 atoms and per-frame sums over all atoms).  Run in a container where the reference checkout is present."""
import os
import struct

import numpy as np

SRC = "/root/reference/data/trajectory.dcd"
HERE = os.path.dirname(os.path.abspath(__file__))


def main():
    raw = open(SRC, "rb").read()
    assert struct.unpack_from("<i", raw, 0)[0] == 84 and raw[4:8] == b"CORD"
    has_cell = struct.unpack_from("<i", raw, 4 + 4 + 40)[0] != 0
    # header records: 84-byte control record, title record, atom-count record
    p = 4 + 84 + 4
    tlen = struct.unpack_from("<i", raw, p)[0]
    ntitle = struct.unpack_from("<i", raw, p + 4)[0]
    p += 4 + tlen + 4
    natom = struct.unpack_from("<i", raw, p + 4)[0]
    first = p + 4 + 4 + 4
    fsize = (4 * natom + 8) * 3 + (56 if has_cell else 0)
    n = 2
    out = bytearray(raw[:first + n * fsize])
    struct.pack_into("<i", out, 8, n)
    open(os.path.join(HERE, "trajectory_2frames.dcd"), "wb").write(bytes(out))
    pos, cells = [], []
    for f in range(n):
        q = first + f * fsize
        if has_cell:
            a, g, b, be, al, c = struct.unpack_from("<6d", raw, q + 4)
            cells.append([a * 0.1, b * 0.1, c * 0.1, al, be, g])
            q += 56
        xyz = []
        for d in range(3):
            xyz.append(np.frombuffer(raw, "<f4", natom, q + 4).astype(np.float64) * 0.1)
            q += 4 * natom + 8
        pos.append(np.stack(xyz, 1))
    pos = np.stack(pos)
    # the first and last 200 atoms in full, and per-frame sums over all atoms
    np.savez_compressed(os.path.join(HERE, "trajectory_2frames_expected.npz"),
                        head=pos[:, :200], tail=pos[:, -200:], sums=pos.sum(axis=1),
                        cells=np.array(cells), n_atoms=natom, n_titles=ntitle,
                        first_frame_offset=first, frame_size=fsize, has_cell=has_cell)


if __name__ == "__main__":
    main()
