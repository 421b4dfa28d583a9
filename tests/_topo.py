"""Small hand-built topologies for unit tests (the reference's tests build
AmberTopology literals the same way: gb_energy.rs:410-448, mm_energy.rs:277-314,
pb_energy.rs:489-526, sa_energy.rs:85-120)."""
import numpy as np

from rust_simulation_tools.topology import AmberTopology


def make_topology(charges_e, radii, screen=None, types=None, n_types=1, lj_a=(0.0,),
                  lj_b=(0.0,), nb_parm_index=None, residue_pointers=(0,), bonds=None,
                  bond_k=(300.0,), bond_r0=(1.5,), angles=None, angle_k=(50.0,),
                  angle_t0=(np.pi,), dihedrals=None, dih_pk=(2.0,), dih_pn=(2.0,),
                  dih_phase=(np.pi,), num_excluded=None, excluded=None, scee=1.2, scnb=2.0):
    q = np.asarray(charges_e, float)
    n = q.shape[0]
    if screen is None:
        screen = np.full(n, 0.8)
    if types is None:
        types = np.zeros(n, np.int32)
    if nb_parm_index is None:
        nb_parm_index = np.arange(1, n_types * n_types + 1)
    if num_excluded is None:
        num_excluded = np.zeros(n, np.int32)
    return AmberTopology(
        charges=q, charges_amber=q * 18.2223, radii=np.asarray(radii, float),
        screen=np.asarray(screen, float), atom_type_indices=types, n_types=n_types,
        nb_parm_index=nb_parm_index, lj_acoef=lj_a, lj_bcoef=lj_b,
        residue_pointers=np.asarray(residue_pointers, np.int32),
        bonds=None if bonds is None else np.asarray(bonds, np.int32),
        bond_force_constants=bond_k, bond_equil_values=bond_r0,
        angles=None if angles is None else np.asarray(angles, np.int32),
        angle_force_constants=angle_k, angle_equil_values=angle_t0,
        dihedrals=None if dihedrals is None else np.asarray(dihedrals, np.int32),
        dihedral_force_constants=dih_pk, dihedral_periodicities=dih_pn, dihedral_phases=dih_phase,
        num_excluded_atoms=num_excluded,
        excluded_atoms_list=None if excluded is None else np.asarray(excluded, np.int32),
        scee_scale_factor=scee, scnb_scale_factor=scnb)


def random_cluster(n, seed, box=12.0, min_sep=1.3, n_types=3):
    """n atoms uniformly in a box with a minimum separation, random radii,
    charges and LJ types (Lorentz-Berthelot A = eps R^12, B = 2 eps R^6)."""
    rng = np.random.default_rng(seed)
    pts = []
    while len(pts) < n:
        p = rng.uniform(-box / 2, box / 2, 3)
        if all(np.linalg.norm(p - o) >= min_sep for o in pts):
            pts.append(p)
    x = np.array(pts)
    radii = rng.choice([1.2, 1.5, 1.7, 1.8], n)
    q = rng.normal(0, 0.4, n)
    types = rng.integers(0, n_types, n).astype(np.int32)
    eps = rng.uniform(0.02, 0.2, n_types)
    rh = rng.uniform(1.0, 2.0, n_types)
    A, B, idx = [], [], np.zeros((n_types, n_types), np.int64)
    k = 0
    for i in range(n_types):
        for j in range(i, n_types):
            e = np.sqrt(eps[i] * eps[j])
            R = rh[i] + rh[j]
            A.append(e * R ** 12)
            B.append(2 * e * R ** 6)
            k += 1
            idx[i, j] = idx[j, i] = k
    top = make_topology(q, radii, screen=rng.uniform(0.7, 0.95, n), types=types,
                        n_types=n_types, lj_a=A, lj_b=B, nb_parm_index=idx.reshape(-1))
    return top, x


def hollow_shell(open_cap):
    """150 atoms on a 7 A sphere (Fibonacci points, spacing ~2.3 A, radius
    1.7 A: a 1.4 A probe cannot pass between them) around a +1 e atom of
    radius 1.2 A at the origin; open_cap removes the atoms within 45 degrees
    of +z, a ~5 A hole the probe passes.  Returns (topology, xyz in A)."""
    n, R = 150, 7.0
    k = np.arange(n) + 0.5
    z = 1.0 - 2.0 * k / n
    r = np.sqrt(1.0 - z * z)
    phi = np.pi * (3.0 - np.sqrt(5.0)) * k
    pts = R * np.stack([r * np.cos(phi), r * np.sin(phi), z], 1)
    if open_cap:
        pts = pts[pts[:, 2] / R < np.cos(np.pi / 4)]
    q = np.where(np.arange(len(pts)) % 2 == 0, 0.3, -0.3)
    q -= q.mean()  # a neutral shell
    x = np.vstack([[0.0, 0.0, 0.0], pts])
    charges = np.concatenate([[1.0], q])
    radii = np.concatenate([[1.2], np.full(len(pts), 1.7)])
    return make_topology(charges, radii), x
