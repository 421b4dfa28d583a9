"""Trajectory ingest: the PyO3 DcdReader / read_dcd_header / MdcrdReader
(crates/bindings/src/lib.rs:1015-1155, :2014-2040) over the library's
readers (crates/core/src/trajectory/dcd.rs, crates/mmpbsa/src/mdcrd.rs).
File errors raise OSError with the reference's messages."""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import _abi


def _fail(lib):
    raise OSError(lib.mmpbsa_traj_last_error().decode())


class DcdReader:
    """DCD frames in nm, the unit cell as [a, b, c] nm + [alpha, beta, gamma]."""

    def __init__(self, path):
        self._lib = _abi.load_library()
        self._h = C.c_void_p()
        if self._lib.mmpbsa_dcd_open(os.fsencode(str(path)), C.byref(self._h)) != 0:
            _fail(self._lib)
        self._hd = _abi.DcdHeaderC()
        nt = C.c_int32()
        self._lib.mmpbsa_dcd_get_header(self._h, C.byref(self._hd), C.byref(nt))
        self._titles = []
        buf = C.create_string_buffer(96)
        for k in range(nt.value):
            self._lib.mmpbsa_dcd_title(self._h, k, buf, 96)
            self._titles.append(buf.value.decode("utf-8", "replace"))

    @property
    def n_frames(self):
        return int(self._hd.n_frames)

    @property
    def n_atoms(self):
        return int(self._hd.n_atoms)

    @property
    def has_unit_cell(self):
        return bool(self._hd.has_unit_cell)

    @property
    def current_frame(self):
        return int(self._lib.mmpbsa_dcd_current(self._h))

    def seek(self, frame):
        if frame < 0:
            raise OverflowError("can't convert negative int to unsigned")
        if self._lib.mmpbsa_dcd_seek(self._h, int(frame)) != 0:
            _fail(self._lib)

    def read_frame(self):
        xyz = np.empty((self.n_atoms, 3))
        box = np.zeros(6)
        got = C.c_int32()
        if self._lib.mmpbsa_dcd_read_frame(self._h, xyz.ctypes.data_as(_abi.c_double_p),
                                           box.ctypes.data_as(_abi.c_double_p), C.byref(got)) != 0:
            _fail(self._lib)
        if not got.value:
            return None
        return xyz, (list(box) if self.has_unit_cell else None)

    def read_frame_at(self, frame_index):
        self.seek(frame_index)
        return self.read_frame()

    def read_all(self):
        self.seek(0)
        frames, boxes = [], []
        while True:
            r = self.read_frame()
            if r is None:
                break
            frames.append(r[0])
            boxes.append(r[1])
        if not frames:
            raise ValueError("No frames to read")
        traj = np.stack(frames)
        has_box = any(b is not None for b in boxes)
        box = np.array([b if b is not None else [0.0] * 6 for b in boxes]) if has_box else None
        return traj, box

    def __del__(self):
        try:
            if self._h:
                self._lib.mmpbsa_dcd_close(self._h)
                self._h = C.c_void_p()
        except Exception:
            pass


def read_dcd_header(path) -> dict:
    r = DcdReader(path)
    hd = r._hd
    return {"n_frames": int(hd.n_frames), "n_atoms": int(hd.n_atoms),
            "start_timestep": int(hd.start_timestep), "timestep_interval": int(hd.timestep_interval),
            "timestep": float(hd.timestep), "has_unit_cell": bool(hd.has_unit_cell),
            "is_charmm": bool(hd.is_charmm), "is_64bit": bool(hd.is_64bit),
            "is_big_endian": bool(hd.big_endian), "titles": list(r._titles),
            "first_frame_offset": int(hd.first_frame_offset), "frame_size": int(hd.frame_size)}


class MdcrdReader:
    """AMBER ASCII trajectory frames in A (the box, if any, is read and
    dropped, as the reference's Python reader does)."""

    def __init__(self, path, n_atoms, has_box):
        self._lib = _abi.load_library()
        self._n = int(n_atoms)
        self._h = C.c_void_p()
        if self._lib.mmpbsa_mdcrd_open(os.fsencode(str(path)), self._n, 1 if has_box else 0,
                                       C.byref(self._h)) != 0:
            _fail(self._lib)

    def read_frame(self):
        xyz = np.empty((self._n, 3))
        box = np.zeros(3)
        got = C.c_int32()
        if self._lib.mmpbsa_mdcrd_read_frame(self._h, xyz.ctypes.data_as(_abi.c_double_p),
                                             box.ctypes.data_as(_abi.c_double_p), C.byref(got)) != 0:
            _fail(self._lib)
        return xyz if got.value else None

    @property
    def current_frame(self):
        return int(self._lib.mmpbsa_mdcrd_current(self._h))

    def __del__(self):
        try:
            if self._h:
                self._lib.mmpbsa_mdcrd_close(self._h)
                self._h = C.c_void_p()
        except Exception:
            pass


def reimage(coords_angstrom, bonds, box_angstrom):
    """reimage_complex (binding.rs:234-308) on a copy of (N, 3) A coordinates;
    bonds (i, j) pairs (extra columns ignored), box [X, Y, Z] A."""
    x = np.ascontiguousarray(np.asarray(coords_angstrom, dtype=np.float64)).copy()
    b = np.ascontiguousarray(np.asarray(bonds, dtype=np.int32).reshape(-1, np.shape(bonds)[-1] if
                                                                         np.ndim(bonds) == 2 else 2)[:, :2])
    bx = np.ascontiguousarray(np.asarray(box_angstrom, dtype=np.float64))
    lib = _abi.load_library()
    lib.mmpbsa_reimage(x.ctypes.data_as(_abi.c_double_p), x.shape[0],
                       b.ctypes.data_as(_abi.c_int32_p), b.shape[0],
                       bx.ctypes.data_as(_abi.c_double_p))
    return x
