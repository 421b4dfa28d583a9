"""Seeded synthetic receptor-ligand complexes for the MM-PBSA benchmarks.

Layout: residues of 16 consecutive atoms, chains of up to 64 residues
(1,024 atoms), receptor chains first, then the ligand.  Coordinates are a
self-avoiding random walk with 1.5 A steps, a 1.5 A minimum separation from
every earlier atom, confined to a sphere of radius R; the ligand grows in a
sub-sphere of radius R_L centred at (0, 0, R - R_L/2) from which the receptor
is excluded, which forms an interface.  When a walk cannot extend it starts a
new chain segment at a free random point.

Elements H/C/N/O/S (p = .5/.3/.08/.1/.02) give radii 1.2/1.7/1.55/1.5/1.8 A and
HCT screening factors .85/.72/.79/.85/.96.  Charges ~ N(0, 0.35) e shifted so
each partner is neutral; charges_amber = 18.2223 q.  LJ well depths
~ U(0.02, 0.2) kcal/mol quantised to 16 bins per element (80 types),
R_min/2 = element radius, Lorentz-Berthelot A = eps R^12, B = 2 eps R^6.
Bonded terms use the constants of crates/mmpbsa/benches/mmpbsa_bench.rs:163-169
(k_b 300, r0 1.52; k_theta 50, theta0 1.911; pk 2, n 2, gamma pi); exclusions
are 1-2/1-3/1-4 (listed for j > i, AMBER placeholder when none), 1-4 pairs come
from every dihedral, scee 1.2, scnb 2.0.
Frames are base + N(0, sigma^2) per coordinate (seed + 1 + frame).
"""
from __future__ import annotations

import math

import numpy as np

from .topology import AmberTopology

ELEM_P = np.array([0.5, 0.3, 0.08, 0.1, 0.02])
ELEM_RADII = np.array([1.2, 1.7, 1.55, 1.5, 1.8])
ELEM_SCREEN = np.array([0.85, 0.72, 0.79, 0.85, 0.96])
ELEM_NAMES = ["H", "C", "N", "O", "S"]
N_BINS = 16
RES_ATOMS = 16
CHAIN_ATOMS = 1024
STEP = 1.5
MIN_SEP = 1.5


def seed_for(config: int, complex_index: int = 0) -> int:
    return 0x5EED0000 + 1000 * config + complex_index


def sphere_radius_for(n_atoms: int) -> float:
    """~16 A^3 per atom."""
    return (3.0 * n_atoms * 16.0 / (4.0 * math.pi)) ** (1.0 / 3.0)


class _Cells:
    """Occupancy hash with 1.5 A cells for the minimum-separation test."""

    def __init__(self, lo, hi):
        self.lo = np.asarray(lo, float) - 2 * MIN_SEP
        ext = np.asarray(hi, float) + 2 * MIN_SEP - self.lo
        self.dims = np.maximum(np.ceil(ext / MIN_SEP).astype(int), 1)
        self.cells: dict = {}
        self.pos = []

    def key(self, p):
        return tuple(np.floor((p - self.lo) / MIN_SEP).astype(int))

    def ok(self, p, skip=-1):
        k = self.key(p)
        for dz in (-1, 0, 1):
            for dy in (-1, 0, 1):
                for dx in (-1, 0, 1):
                    lst = self.cells.get((k[0] + dx, k[1] + dy, k[2] + dz))
                    if not lst:
                        continue
                    for a in lst:
                        if a == skip:
                            continue
                        d = self.pos[a] - p
                        if d[0] * d[0] + d[1] * d[1] + d[2] * d[2] < MIN_SEP * MIN_SEP:
                            return False
        return True

    def add(self, p):
        i = len(self.pos)
        self.pos.append(p)
        self.cells.setdefault(self.key(p), []).append(i)
        return i


def _unit(rng, m):
    v = rng.normal(size=(m, 3))
    return v / np.linalg.norm(v, axis=1, keepdims=True)


def _walk(rng, n_rec, n_lig, R, R_L):
    cL = np.array([0.0, 0.0, R - 0.5 * R_L])
    lo = np.minimum(-R, cL - R_L)
    hi = np.maximum(R, cL + R_L)
    cells = _Cells(lo, hi)
    xyz = np.zeros((n_rec + n_lig, 3))
    seg_start = np.zeros(n_rec + n_lig, bool)   # atom begins a new chain segment

    def allowed(p, lig):
        if lig:
            return np.dot(p - cL, p - cL) <= R_L * R_L
        return np.dot(p, p) <= R * R and np.dot(p - cL, p - cL) >= R_L * R_L

    def random_free(lig):
        for _ in range(100000):
            if lig:
                p = cL + R_L * (rng.uniform(-1, 1, 3))
            else:
                p = R * rng.uniform(-1, 1, 3)
            if allowed(p, lig) and cells.ok(p):
                return p
        raise RuntimeError("synthetic packing failed: sphere too dense")

    for part, (a0, a1) in enumerate(((0, n_rec), (n_rec, n_rec + n_lig))):
        lig = part == 1
        chain_pos = 0
        prev = -1
        for a in range(a0, a1):
            new_chain = (chain_pos % CHAIN_ATOMS == 0) or prev < 0
            placed = None
            if not new_chain:
                cand = xyz[prev] + STEP * _unit(rng, 24)
                for c in cand:
                    if allowed(c, lig) and cells.ok(c, skip=prev):
                        placed = c
                        break
            if placed is None:
                placed = random_free(lig)
                seg_start[a] = True
                chain_pos = 0
            xyz[a] = placed
            cells.add(placed)
            prev = a
            chain_pos += 1
    return xyz, seg_start


def build_complex(n_atoms: int, n_ligand: int, radius: float | None = None,
                  seed: int = 0x5EED0000):
    """Returns (AmberTopology, base coords [n,3] A, receptor residue ids,
    ligand residue ids)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    n_rec = n_atoms - n_ligand
    R = float(radius) if radius is not None else sphere_radius_for(n_atoms)
    R_L = sphere_radius_for(max(n_ligand, 1))
    elem = rng.choice(5, size=n_atoms, p=ELEM_P)
    radii = ELEM_RADII[elem]
    screen = ELEM_SCREEN[elem]
    q = rng.normal(0.0, 0.35, n_atoms)
    if n_rec > 0:
        q[:n_rec] -= q[:n_rec].mean()
    if n_ligand > 0:
        q[n_rec:] -= q[n_rec:].mean()
    eps = rng.uniform(0.02, 0.2, n_atoms)
    width = 0.18 / N_BINS
    b = np.clip(np.floor((eps - 0.02) / width).astype(int), 0, N_BINS - 1)
    atype = (elem * N_BINS + b).astype(np.int32)
    n_types = 5 * N_BINS
    t_eps = 0.02 + (np.arange(n_types) % N_BINS + 0.5) * width
    t_rh = ELEM_RADII[np.arange(n_types) // N_BINS]
    # AMBER-style packed upper triangle, 1-based index table
    nb = np.zeros((n_types, n_types), np.int64)
    A, B = [], []
    k = 0
    for i in range(n_types):
        for j in range(i, n_types):
            e = math.sqrt(t_eps[i] * t_eps[j])
            rr = t_rh[i] + t_rh[j]
            A.append(e * rr ** 12)
            B.append(2.0 * e * rr ** 6)
            k += 1
            nb[i, j] = nb[j, i] = k
    xyz, seg = _walk(rng, n_rec, n_ligand, R, R_L)
    # bonded terms within chain segments
    seg_id = np.cumsum(seg) - 1
    bonds, angles, dihs = [], [], []
    for i in range(n_atoms - 1):
        if seg_id[i + 1] == seg_id[i]:
            bonds.append((i, i + 1, 0))
    for i in range(n_atoms - 2):
        if seg_id[i + 2] == seg_id[i]:
            angles.append((i, i + 1, i + 2, 0))
    for i in range(n_atoms - 3):
        if seg_id[i + 3] == seg_id[i]:
            dihs.append((i, i + 1, i + 2, i + 3, 0, 0))
    nex = np.zeros(n_atoms, np.int32)
    exl = []
    for i in range(n_atoms):
        cnt = 0
        for d in (1, 2, 3):
            j = i + d
            if j < n_atoms and seg_id[j] == seg_id[i]:
                exl.append(j)
                cnt += 1
        if cnt == 0:
            exl.append(-1)       # AMBER placeholder entry
            cnt = 1
        nex[i] = cnt
    # residues: 16-atom blocks within each partner
    rp = list(range(0, n_rec, RES_ATOMS)) + list(range(n_rec, n_atoms, RES_ATOMS))
    n_rres = len(range(0, n_rec, RES_ATOMS))
    labels = ["REC"] * n_rres + ["LIG"] * (len(rp) - n_rres)
    top = AmberTopology(
        charges=q, charges_amber=q * 18.2223, radii=radii, screen=screen,
        atom_type_indices=atype, n_types=n_types, nb_parm_index=nb.reshape(-1),
        lj_acoef=np.array(A), lj_bcoef=np.array(B), residue_pointers=np.array(rp, np.int32),
        residue_labels=labels, bonds=np.array(bonds, np.int32).reshape(-1, 3),
        bond_force_constants=[300.0], bond_equil_values=[1.52],
        angles=np.array(angles, np.int32).reshape(-1, 4), angle_force_constants=[50.0],
        angle_equil_values=[1.911], dihedrals=np.array(dihs, np.int32).reshape(-1, 6),
        dihedral_force_constants=[2.0], dihedral_periodicities=[2.0],
        dihedral_phases=[math.pi], num_excluded_atoms=nex,
        excluded_atoms_list=np.array(exl, np.int32), scee_scale_factor=1.2,
        scnb_scale_factor=2.0, atom_names=[ELEM_NAMES[e] for e in elem])
    rec_res = np.arange(n_rres, dtype=np.int32)
    lig_res = np.arange(n_rres, len(rp), dtype=np.int32)
    return top, xyz, rec_res, lig_res


def noisy_frame(base: np.ndarray, seed: int, frame: int, sigma: float = 0.3) -> np.ndarray:
    rng = np.random.Generator(np.random.PCG64(seed + 1 + frame))
    return base + rng.normal(0.0, sigma, base.shape)


def config_system(config: int, complex_index: int = 0):
    """The receptor-ligand complex of a benchmark configuration.
    config 0: 2,500 atoms (2,300 + 200) in a 22 A sphere.
    config 1: 25,000 atoms (23,500 + 1,500) in a 45 A sphere.
    config 4: 150,000 atoms (142,500 + 7,500), ~9,400 residues of 16 atoms,
    in the 16 A^3-per-atom sphere (R ~ 82 A)."""
    s = seed_for(config, complex_index)
    if config == 0:
        return build_complex(2500, 200, 22.0, s) + (s,)
    if config == 1:
        return build_complex(25000, 1500, 45.0, s) + (s,)
    if config == 4:
        return build_complex(150000, 7500, None, s) + (s,)
    raise ValueError(f"no fixed system for configuration {config}")


def interface_center(xyz, rec_atoms, lig_atoms, cutoff: float = 6.0) -> np.ndarray:
    """Centroid of the interface: the receptor atoms within ``cutoff`` of a
    ligand atom and the ligand atoms within ``cutoff`` of a receptor atom.
    Falls back to the ligand centroid when the partners do not touch."""
    from scipy.spatial import cKDTree

    x = np.asarray(xyz, np.float64)
    r, l = x[np.asarray(rec_atoms)], x[np.asarray(lig_atoms)]
    if r.shape[0] == 0 or l.shape[0] == 0:
        return (l if l.shape[0] else r).mean(axis=0)
    near_r = np.isfinite(cKDTree(l).query(r, k=1, distance_upper_bound=cutoff)[0])
    near_l = np.isfinite(cKDTree(r).query(l, k=1, distance_upper_bound=cutoff)[0])
    pts = np.concatenate([r[near_r], l[near_l]])
    return (pts if pts.shape[0] else l).mean(axis=0)


def focusing_grids(xyz, rec_atoms, lig_atoms, coarse_dims=193, coarse_spacing=1.0,
                   fine_dims=321, fine_spacing=0.4, cutoff: float = 6.0) -> dict:
    """Explicit coarse-to-fine PB grids shared by complex, receptor and
    ligand: a coarse grid centred on the complex's bounding box and a fine
    grid centred on the interface centroid (PbParams keyword arguments; the
    large-decomposition workload uses the defaults, 193^3 at 1.0 A then
    321^3 at 0.4 A)."""
    x = np.asarray(xyz, np.float64)
    sel = np.concatenate([np.asarray(rec_atoms), np.asarray(lig_atoms)])
    mn, mx = x[sel].min(axis=0), x[sel].max(axis=0)
    cd = (coarse_dims,) * 3 if np.isscalar(coarse_dims) else tuple(coarse_dims)
    fd = (fine_dims,) * 3 if np.isscalar(fine_dims) else tuple(fine_dims)
    return dict(coarse_dims=cd, coarse_spacing=float(coarse_spacing),
                coarse_center=tuple(0.5 * (mn + mx)), fine_dims=fd,
                grid_spacing=float(fine_spacing),
                fine_center=tuple(interface_center(x, rec_atoms, lig_atoms, cutoff)))
