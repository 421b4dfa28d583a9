"""Host-side topology: the fields of AmberTopology that the MM-PBSA path reads.

Mirrors crates/core/src/amber/prmtop.rs:34-105 (struct), :120-132
(atom_residue_indices) and :165-204 (build_selection), and
crates/mmpbsa/src/subsystem.rs:22-181 (extract_subtopology, extract_coords).
Arrays are numpy, indices 0-based; AMBER's exclusion placeholder is -1.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from ._abi import TopologyView, as_ptr


class _Getter(np.ndarray):
    """An array field that can also be called, so both the dataclass field
    (topo.charges) and the reference's PyO3 getter method (topo.charges(),
    lib.rs getters returning a fresh numpy array) read the same data."""

    def __call__(self):
        return np.array(self)


def _getter(a, dtype):
    return np.ascontiguousarray(np.asarray(a, dtype=dtype)).view(_Getter)


def _f64(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.float64))


def _i32(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.int32))


@dataclass
class AmberTopology:
    charges: np.ndarray                      # e
    charges_amber: np.ndarray                # e * 18.2223
    radii: np.ndarray
    screen: np.ndarray
    atom_type_indices: np.ndarray            # 0-based
    n_types: int
    nb_parm_index: np.ndarray                # int64 [n_types^2], 1-based
    lj_acoef: np.ndarray
    lj_bcoef: np.ndarray
    residue_pointers: np.ndarray             # int32 [n_residues]
    residue_labels: list = field(default_factory=list)
    bonds: np.ndarray = None                 # int32 [nb, 3] i, j, type
    bond_force_constants: np.ndarray = None
    bond_equil_values: np.ndarray = None
    angles: np.ndarray = None                # int32 [na, 4]
    angle_force_constants: np.ndarray = None
    angle_equil_values: np.ndarray = None
    dihedrals: np.ndarray = None             # int32 [nd, 6] i, j, k, l, type, ignore_14
    dihedral_force_constants: np.ndarray = None
    dihedral_periodicities: np.ndarray = None
    dihedral_phases: np.ndarray = None
    num_excluded_atoms: np.ndarray = None    # int32 [n]
    excluded_atoms_list: np.ndarray = None   # int32, -1 = placeholder
    scee_scale_factor: float = 1.2
    scnb_scale_factor: float = 2.0
    atom_names: list = None
    masses: np.ndarray = None                # amu (prmtop MASS)

    def __post_init__(self):
        self.charges = _getter(self.charges, np.float64)
        self.charges_amber = _f64(self.charges_amber)
        self.radii = _f64(self.radii)
        self.screen = _f64(self.screen)
        self.atom_type_indices = _i32(self.atom_type_indices)
        self.nb_parm_index = np.ascontiguousarray(np.asarray(self.nb_parm_index, dtype=np.int64))
        self.lj_acoef = _f64(self.lj_acoef)
        self.lj_bcoef = _f64(self.lj_bcoef)
        self.residue_pointers = _getter(self.residue_pointers, np.int32)
        self.bonds = _i32(np.zeros((0, 3)) if self.bonds is None else self.bonds).reshape(-1, 3)
        self.angles = _i32(np.zeros((0, 4)) if self.angles is None else self.angles).reshape(-1, 4)
        self.dihedrals = _i32(np.zeros((0, 6)) if self.dihedrals is None else self.dihedrals).reshape(-1, 6)
        for k in ("bond_force_constants", "bond_equil_values", "angle_force_constants",
                  "angle_equil_values", "dihedral_force_constants", "dihedral_periodicities",
                  "dihedral_phases"):
            v = getattr(self, k)
            setattr(self, k, _f64(np.zeros(1) if v is None or len(v) == 0 else v))
        n = self.n_atoms
        if self.num_excluded_atoms is None:
            self.num_excluded_atoms = np.zeros(n, np.int32)
        self.num_excluded_atoms = _i32(self.num_excluded_atoms)
        if self.excluded_atoms_list is None:
            self.excluded_atoms_list = np.zeros(0, np.int32)
        self.excluded_atoms_list = _i32(self.excluded_atoms_list)
        if not self.residue_labels:
            self.residue_labels = ["RES"] * len(self.residue_pointers)
        if self.atom_names is None:
            self.atom_names = ["X"] * n
        self.masses = _f64(np.zeros(n) if self.masses is None else self.masses)

    # ---- reference accessors -------------------------------------------
    @property
    def n_atoms(self) -> int:
        return int(self.charges.shape[0])

    @property
    def n_residues(self) -> int:
        return int(self.residue_pointers.shape[0])

    def atom_residue_indices(self) -> np.ndarray:
        """prmtop.rs:120-132"""
        out = np.zeros(self.n_atoms, np.int32)
        ptr = list(self.residue_pointers) + [self.n_atoms]
        for r in range(self.n_residues):
            out[ptr[r]:ptr[r + 1]] = r
        return out

    def _self_lj(self):
        """Per-type sigma (nm) and epsilon (kJ/mol) from the self terms
        (prmtop.rs:530-556)."""
        nt = int(self.n_types)
        sig, eps = np.zeros(nt), np.zeros(nt)
        for i in range(nt):
            k = nt * i + i
            if k >= self.nb_parm_index.shape[0] or self.nb_parm_index[k] < 1:
                continue
            a, b = self.lj_acoef[self.nb_parm_index[k] - 1], self.lj_bcoef[self.nb_parm_index[k] - 1]
            if abs(b) > 1e-20 and abs(a) > 1e-20:
                sig[i] = (a / b) ** (1.0 / 6.0) * 0.1
                eps[i] = b * b / (4.0 * a) * 4.184
        return sig, eps

    def sigmas(self) -> np.ndarray:
        return self._self_lj()[0][self.atom_type_indices]

    def epsilons(self) -> np.ndarray:
        return self._self_lj()[1][self.atom_type_indices]

    def bonds_list(self):
        """The reference's bonds() getter: (i, j) pairs."""
        return [(int(b[0]), int(b[1])) for b in self.bonds]

    def build_selection(self, residue_indices) -> np.ndarray:
        """prmtop.rs:165-204: atom indices of the residues, in the given order."""
        ptr = list(self.residue_pointers) + [self.n_atoms]
        out = []
        for r in residue_indices:
            r = int(r)
            if r < 0 or r >= self.n_residues:
                raise ValueError(
                    f"Residue index {r} out of range (0-{self.n_residues - 1})")
        for r in residue_indices:
            r = int(r)
            out.extend(range(ptr[r], ptr[r + 1]))
        return np.asarray(out, dtype=np.int32)

    # ---- C view ---------------------------------------------------------
    def view(self) -> TopologyView:
        """A TopologyView whose pointers stay valid while self is alive."""
        v = TopologyView()
        v.n_atoms = self.n_atoms
        v.n_residues = self.n_residues
        v.n_types = int(self.n_types)
        v.charges = as_ptr(self.charges, C.c_double)
        v.charges_amber = as_ptr(self.charges_amber, C.c_double)
        v.radii = as_ptr(self.radii, C.c_double)
        v.screen = as_ptr(self.screen, C.c_double)
        v.atom_type = as_ptr(self.atom_type_indices, C.c_int32)
        v.nb_parm_index = as_ptr(self.nb_parm_index, C.c_int64)
        v.lj_acoef = as_ptr(self.lj_acoef, C.c_double)
        v.lj_bcoef = as_ptr(self.lj_bcoef, C.c_double)
        v.n_lj = int(self.lj_acoef.shape[0])
        v.residue_pointers = as_ptr(self.residue_pointers, C.c_int32)
        v.n_bonds = int(self.bonds.shape[0])
        v.bonds = as_ptr(self.bonds, C.c_int32)
        v.bond_k = as_ptr(self.bond_force_constants, C.c_double)
        v.bond_r0 = as_ptr(self.bond_equil_values, C.c_double)
        v.n_bond_types = int(self.bond_force_constants.shape[0])
        v.n_angles = int(self.angles.shape[0])
        v.angles = as_ptr(self.angles, C.c_int32)
        v.angle_k = as_ptr(self.angle_force_constants, C.c_double)
        v.angle_theta0 = as_ptr(self.angle_equil_values, C.c_double)
        v.n_angle_types = int(self.angle_force_constants.shape[0])
        v.n_dihedrals = int(self.dihedrals.shape[0])
        v.dihedrals = as_ptr(self.dihedrals, C.c_int32)
        v.dihedral_pk = as_ptr(self.dihedral_force_constants, C.c_double)
        v.dihedral_pn = as_ptr(self.dihedral_periodicities, C.c_double)
        v.dihedral_phase = as_ptr(self.dihedral_phases, C.c_double)
        v.n_dihedral_types = int(self.dihedral_force_constants.shape[0])
        v.num_excluded = as_ptr(self.num_excluded_atoms, C.c_int32)
        v.excluded_atoms = as_ptr(self.excluded_atoms_list, C.c_int32)
        v.n_excluded = int(self.excluded_atoms_list.shape[0])
        v.scee = float(self.scee_scale_factor)
        v.scnb = float(self.scnb_scale_factor)
        return v


def read_prmtop(path) -> AmberTopology:
    """read_prmtop (lib.rs:979-984 over prmtop.rs:436-733), parsed by the
    library; failures raise OSError with the reference's messages."""
    import os

    from ._abi import PrmtopInfoC, load_library
    lib = load_library()
    h = C.c_void_p()
    rc = lib.mmpbsa_prmtop_read(os.fsencode(str(path)), C.byref(h))
    if rc != 0:
        raise OSError(lib.mmpbsa_prmtop_last_error().decode())
    try:
        v = TopologyView()
        info = PrmtopInfoC()
        lib.mmpbsa_prmtop_view(h, C.byref(v))
        lib.mmpbsa_prmtop_info_get(h, C.byref(info))

        def arr(ptr, count, dtype):
            if count <= 0 or not ptr:
                return np.zeros(0, dtype)
            return np.ctypeslib.as_array(ptr, shape=(count,)).astype(dtype, copy=True)

        def names(base, count):
            return [C.string_at(base + 5 * k).decode("ascii", "replace") for k in range(count)]

        n, nr, nt = v.n_atoms, v.n_residues, v.n_types
        return AmberTopology(
            charges=arr(v.charges, n, np.float64), charges_amber=arr(v.charges_amber, n, np.float64),
            radii=arr(v.radii, n, np.float64), screen=arr(v.screen, n, np.float64),
            atom_type_indices=arr(v.atom_type, n, np.int32), n_types=nt,
            nb_parm_index=arr(v.nb_parm_index, nt * nt, np.int64),
            lj_acoef=arr(v.lj_acoef, v.n_lj, np.float64), lj_bcoef=arr(v.lj_bcoef, v.n_lj, np.float64),
            residue_pointers=arr(v.residue_pointers, nr, np.int32),
            residue_labels=names(info.residue_labels, nr),
            bonds=arr(v.bonds, 3 * v.n_bonds, np.int32).reshape(-1, 3),
            bond_force_constants=arr(v.bond_k, v.n_bond_types, np.float64),
            bond_equil_values=arr(v.bond_r0, v.n_bond_types, np.float64),
            angles=arr(v.angles, 4 * v.n_angles, np.int32).reshape(-1, 4),
            angle_force_constants=arr(v.angle_k, v.n_angle_types, np.float64),
            angle_equil_values=arr(v.angle_theta0, v.n_angle_types, np.float64),
            dihedrals=arr(v.dihedrals, 6 * v.n_dihedrals, np.int32).reshape(-1, 6),
            dihedral_force_constants=arr(v.dihedral_pk, v.n_dihedral_types, np.float64),
            dihedral_periodicities=arr(v.dihedral_pn, v.n_dihedral_types, np.float64),
            dihedral_phases=arr(v.dihedral_phase, v.n_dihedral_types, np.float64),
            num_excluded_atoms=arr(v.num_excluded, n, np.int32),
            excluded_atoms_list=arr(v.excluded_atoms, v.n_excluded, np.int32),
            scee_scale_factor=v.scee, scnb_scale_factor=v.scnb,
            atom_names=names(info.atom_names, n), masses=arr(info.masses, n, np.float64))
    finally:
        lib.mmpbsa_prmtop_destroy(h)


def extract_subtopology(top: AmberTopology, atom_indices) -> AmberTopology:
    """subsystem.rs:22-169: per-atom arrays subset, bonded terms kept when all
    their atoms are selected, exclusion lists rebuilt, residues regrouped."""
    idx = np.asarray(atom_indices, dtype=np.int64)
    n_sub = idx.shape[0]
    old_to_new = np.full(top.n_atoms, -1, np.int64)
    old_to_new[idx] = np.arange(n_sub)
    ares = top.atom_residue_indices()
    res_ptr, labels = [], []
    last = None
    for new_i, old_i in enumerate(idx):
        r = int(ares[old_i])
        if r != last:
            labels.append(top.residue_labels[r])
            res_ptr.append(new_i)
            last = r

    def remap(rows, ncol):
        if rows.shape[0] == 0:
            return rows
        m = old_to_new[rows[:, :ncol]]
        keep = np.all(m >= 0, axis=1)
        out = rows[keep].copy()
        out[:, :ncol] = m[keep]
        return out

    bonds = remap(top.bonds, 2)
    angles = remap(top.angles, 3)
    dihedrals = remap(top.dihedrals, 4)
    # exclusion lists (subsystem.rs:113-138)
    nex = np.zeros(n_sub, np.int32)
    exl = []
    off = 0
    lst = top.excluded_atoms_list
    for i in range(top.n_atoms):
        cnt = int(top.num_excluded_atoms[i]) if i < top.num_excluded_atoms.shape[0] else 0
        ni = old_to_new[i]
        if ni >= 0:
            c = 0
            for k in range(cnt):
                if off + k < lst.shape[0]:
                    j = int(lst[off + k])
                    if 0 <= j < top.n_atoms and old_to_new[j] >= 0:
                        exl.append(int(old_to_new[j]))
                        c += 1
            nex[ni] = c
        off += cnt
    return AmberTopology(
        charges=top.charges[idx], charges_amber=top.charges_amber[idx], radii=top.radii[idx],
        screen=top.screen[idx], atom_type_indices=top.atom_type_indices[idx],
        n_types=top.n_types, nb_parm_index=top.nb_parm_index, lj_acoef=top.lj_acoef,
        lj_bcoef=top.lj_bcoef, residue_pointers=np.asarray(res_ptr, np.int32),
        residue_labels=labels, bonds=bonds, bond_force_constants=top.bond_force_constants,
        bond_equil_values=top.bond_equil_values, angles=angles,
        angle_force_constants=top.angle_force_constants,
        angle_equil_values=top.angle_equil_values, dihedrals=dihedrals,
        dihedral_force_constants=top.dihedral_force_constants,
        dihedral_periodicities=top.dihedral_periodicities,
        dihedral_phases=top.dihedral_phases, num_excluded_atoms=nex,
        excluded_atoms_list=np.asarray(exl, np.int32),
        scee_scale_factor=top.scee_scale_factor, scnb_scale_factor=top.scnb_scale_factor,
        atom_names=[top.atom_names[i] for i in idx])


def extract_coords(coords, atom_indices):
    """subsystem.rs:172-175"""
    return np.ascontiguousarray(np.asarray(coords, np.float64)[np.asarray(atom_indices)])
