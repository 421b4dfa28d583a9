"""Pins of the FP64 oracle (CPU): the reference's own known-answer tests,
closed forms, brute force on tiny inputs, and a scipy direct solve of the
discrete LPBE.  A plausible slip (a dropped term, a sign, a transposed index)
fails at least one of these."""
import ctypes as C
import math

import numpy as np
import pytest
from scipy import integrate, sparse
from scipy.sparse.linalg import spsolve

from _topo import make_topology, random_cluster

dp = C.POINTER(C.c_double)


def P(a):
    return a.ctypes.data_as(dp)


# ---- reference unit tests ---------------------------------------------------
def test_kappa_known_answer(oracle):
    # gb_energy.rs:394-408
    assert abs(oracle.compute_kappa(0.15, 78.3, 300.0) - 0.127315) < 1e-3
    assert oracle.compute_kappa(0.0, 78.3, 300.0) == 0.0
    assert oracle.compute_kappa(-1.0, 78.3, 300.0) == 0.0
    # closed form: kappa^2 = 8 pi l_B N_A c 1e-27 with l_B = 332.0522 / (eps kT)
    k = oracle.compute_kappa(0.15, 80.0, 300.0)
    lb = 332.0522 / (80.0 * 0.00198688 * 300.0)
    assert abs(k * k - 8 * math.pi * lb * 6.022e-4 * 0.15) < 1e-12


def test_single_ion_gb(oracle):
    # gb_energy.rs:451-494
    top = make_topology([1.0], [1.5], screen=[0.8])
    e, R = oracle.gb_energy(top, np.zeros((1, 3)), oracle.gb_params(model=1, salt_concentration=0.0))
    rho = 1.5 - 0.09
    assert abs(R[0] - rho) < 1e-10
    assert abs(e - (-0.5 * (1 - 1 / 78.3) * 18.2223 ** 2 / rho)) < 0.01


def test_bonded_known_answers(oracle):
    # mm_energy.rs:317-399
    top = make_topology([0, 0], [1.5, 1.5], bonds=[[0, 1, 0]], bond_k=[300.0], bond_r0=[1.5],
                        num_excluded=[1, 0], excluded=[1])
    assert abs(oracle.mm_energy(top, np.array([[0, 0, 0], [2.0, 0, 0]]))["bond"] - 75.0) < 1e-10
    top = make_topology([0] * 3, [1.5] * 3, angles=[[0, 1, 2, 0]], angle_k=[50.0],
                        angle_t0=[math.pi], num_excluded=[2, 1, 0], excluded=[1, 2, 2])
    x = np.array([[1.0, 0, 0], [0, 0, 0], [0, 1.0, 0]])
    assert abs(oracle.mm_energy(top, x)["angle"] - 50 * (math.pi / 2) ** 2) < 1e-10
    top = make_topology([0] * 4, [1.5] * 4, dihedrals=[[0, 1, 2, 3, 0, 0]], dih_pk=[2.0],
                        dih_pn=[2.0], dih_phase=[math.pi], num_excluded=[3, 2, 1, 0],
                        excluded=[1, 2, 3, 2, 3, 3])
    x = np.array([[1.0, 1.0, 0], [0, 0, 0], [1.0, 0, 0], [2.0, 1.0, 0]])
    d = oracle.mm_energy(top, x)["dihedral"]
    assert 0.0 <= d < 10.0
    # closed forms of 2 (1 + cos(2 phi - pi)): cis phi = 0 -> 0, phi = 90 deg -> 4
    x_cis = np.array([[0, 1.0, 0], [0, 0, 0], [1.0, 0, 0], [1.0, 1.0, 0]])
    x_perp = np.array([[0, 1.0, 0], [0, 0, 0], [1.0, 0, 0], [1.0, 0, 1.0]])
    assert abs(oracle.mm_energy(top, x_cis)["dihedral"] - 0.0) < 1e-12
    assert abs(oracle.mm_energy(top, x_perp)["dihedral"] - 4.0) < 1e-12


def test_single_atom_sasa(oracle):
    # sa_energy.rs:82-141; an isolated atom keeps all points: exactly 4 pi (r+p)^2
    top = make_topology([0.0], [1.7])
    t, ts, per = oracle.sa_energy(top, np.zeros((1, 3)), oracle.sa_params())
    exact = 4 * math.pi * (1.7 + 1.4) ** 2
    assert abs(ts - exact) < 1e-9
    assert abs(t - 0.0072 * exact) < 1e-12


def test_sasa_two_atoms_burial(oracle):
    # a fully buried point set: atom 0 inside atom 1's extended sphere -> SASA 0
    top = make_topology([0, 0], [1.0, 3.0])
    _, _, per = oracle.sa_energy(top, np.array([[0, 0, 0], [0.1, 0, 0]]), oracle.sa_params())
    assert per[0] == 0.0 and per[1] > 0.0
    # brute-force count on a random cluster against a direct numpy evaluation
    top, x = random_cluster(30, 3, box=9.0)
    _, _, per = oracle.sa_energy(top, x, oracle.sa_params(n_sphere_points=96))
    n = 96
    i = np.arange(n)
    y = 1 - (i / (n - 1)) * 2
    th = math.pi * (3 - math.sqrt(5)) * i
    pts = np.stack([np.cos(th) * np.sqrt(1 - y * y), y, np.sin(th) * np.sqrt(1 - y * y)], 1)
    ext = top.radii + 1.4
    for a in range(x.shape[0]):
        tp = x[a] + pts * ext[a]
        d2 = ((tp[:, None, :] - x[None, :, :]) ** 2).sum(-1)
        buried = (d2 <= ext[None, :] ** 2)
        buried[:, a] = False
        cnt = np.sum(~buried.any(1))
        assert per[a] == pytest.approx(cnt / n * 4 * math.pi * ext[a] ** 2, rel=0, abs=1e-12)


def test_grid_known_answers(oracle):
    L = oracle.lib()
    g = oracle.OrcGrid()
    x = np.array([[0.0, 0, 0], [10.0, 10, 10]])
    L.orc_auto_grid(2, P(x), 0.5, 5.0, C.byref(g))  # pb_grid.rs:1004-1010
    assert all(g.n[k] % 2 == 1 for k in range(3))
    x = np.array([[5.0, 5.0, 5.0]])
    L.orc_auto_grid(1, P(x), 1.0, 5.0, C.byref(g))  # pb_grid.rs:1013-1025
    rho = np.zeros(g.n[0] * g.n[1] * g.n[2])
    q = np.array([1.0])
    L.orc_map_charges(C.byref(g), 1, P(x), P(q), P(rho))
    assert abs(rho.sum() * g.h[0] ** 3 - 1.0) < 1e-10
    # off-node charge: conservation and first moment (trilinear weights)
    x = np.array([[5.3, 4.71, 5.05]])
    L.orc_map_charges(C.byref(g), 1, P(x), P(q), P(rho))
    nodes = np.stack(np.meshgrid(*[g.o[k] + np.arange(g.n[k]) * g.h[k] for k in range(3)],
                                 indexing="ij"), -1).transpose(2, 1, 0, 3).reshape(-1, 3)
    assert abs(rho.sum() - 1.0) < 1e-12
    assert np.allclose((rho[:, None] * nodes).sum(0), x[0], atol=1e-12)
    # interpolation at a node (pb_grid.rs:1028-1041) through the energy gather
    g3 = oracle.OrcGrid()
    for k in range(3):
        g3.n[k] = 3
        g3.h[k] = 1.0
        g3.o[k] = 0.0
    phi = np.zeros(27)
    phi[1 + 3 * (1 + 3 * 1)] = 1.0
    e = L.orc_elec_energy(C.byref(g3), P(phi), 1, P(np.array([1.0, 1.0, 1.0])), P(np.array([2.0])))
    assert abs(e - 1.0) < 1e-10


def test_auto_grid_invariants(oracle):
    """auto_grid (pb_grid.rs:188-215) on point clouds of real extent, checked
    through what any correct placement must satisfy on each axis: the grid is
    centred on the bounding box, covers it with the buffer on both sides, and
    has the fewest odd node count that does ((n - 3) h < extent + 2 buffer
    <= (n - 1) h)."""
    L = oracle.lib()
    rng = np.random.default_rng(11)
    for trial in range(40):
        n = int(rng.integers(1, 60))
        x = rng.normal(0, rng.uniform(1, 15), (n, 3)) + rng.uniform(-50, 50, 3)
        h = float(rng.choice([0.37, 0.5, 0.6, 0.7, 1.0]))
        buf = float(rng.choice([0.0, 2.5, 5.0, 10.0]))
        g = oracle.OrcGrid()
        L.orc_auto_grid(n, P(np.ascontiguousarray(x)), h, buf, C.byref(g))
        mn, mx = x.min(0), x.max(0)
        for k in range(3):
            nk, o = g.n[k], g.o[k]
            assert g.h[k] == h and nk % 2 == 1
            span = (nk - 1) * h
            ext = mx[k] - mn[k] + 2 * buf
            assert abs((o + 0.5 * span) - 0.5 * (mn[k] + mx[k])) <= 1e-12 * (1 + abs(mn[k]) + abs(mx[k]))
            assert o <= mn[k] - buf + 1e-9 and o + span >= mx[k] + buf - 1e-9
            assert (nk - 3) * h < ext <= span + 1e-9


def test_restrict_prolong_constants(oracle):
    # pb_solver.rs:1260-1283 (and exactness for odd dimensions)
    L = oracle.lib()
    fd = np.array([9, 9, 9], np.int64)
    cd = np.array([5, 5, 5], np.int64)
    fine = np.full(729, 3.0)
    coarse = np.zeros(125)
    i64 = C.POINTER(C.c_int64)
    L.orc_restrict(P(fine), fd.ctypes.data_as(i64), P(coarse), cd.ctypes.data_as(i64), 1.0)
    assert np.all(np.abs(coarse - 3.0) < 1e-12)
    out = np.zeros(729)
    L.orc_prolongate_add(P(coarse), cd.ctypes.data_as(i64), P(out), fd.ctypes.data_as(i64))
    assert np.all(np.abs(out - 3.0) < 1e-12)
    # even fine dims extrapolate on the last plane (the clamped coarse cell)
    fd = np.array([6, 6, 6], np.int64)
    cd = np.array([3, 3, 3], np.int64)
    coarse = np.zeros(27)
    coarse[1 + 3 * (1 + 3 * 1)] = 1.0
    out = np.zeros(216)
    L.orc_prolongate_add(P(coarse), cd.ctypes.data_as(i64), P(out), fd.ctypes.data_as(i64))
    # node (5,2,2): x = 2.5 -> cell 1, weight 1.5 on coarse x=2 (0) and -0.5 on x=1
    assert abs(out[5 + 6 * (2 + 6 * 2)] - (-0.5)) < 1e-15


def _maps(oracle, g, x, radii, q, eps_in, eps_out, kappa, probe=1.4, ion=2.0):
    L = oracle.lib()
    N = g.n[0] * g.n[1] * g.n[2]
    n = x.shape[0]
    rho = np.zeros(N)
    ex, ey, ez, k2 = np.zeros(N), np.zeros(N), np.zeros(N), np.zeros(N)
    L.orc_map_charges(C.byref(g), n, P(x), P(q), P(rho))
    L.orc_assign_dielectrics(C.byref(g), n, P(x), P(radii), probe, eps_in, eps_out, P(ex), P(ey),
                             P(ez))
    if kappa > 0:
        L.orc_assign_kappa(C.byref(g), n, P(x), P(radii), ion, kappa, eps_out, P(k2))
    return rho, ex, ey, ez, k2


def test_sor_and_multigrid_converge(oracle):
    # pb_solver.rs:1194-1257
    L = oracle.lib()
    x = np.array([[0.0, 0.0, 0.0]])
    g = oracle.OrcGrid()
    L.orc_auto_grid(1, P(x), 1.0, 10.0, C.byref(g))
    rho, ex, ey, ez, k2 = _maps(oracle, g, x, np.array([1.5]), np.array([1.0]), 1.0, 1.0, 0.0)
    for fn in (L.orc_solve_sor, L.orc_solve_multigrid):
        phi = np.zeros_like(rho)
        r = oracle.OrcSolve()
        fn(C.byref(g), P(rho), P(ex), P(ey), P(ez), P(k2), 0, 1, P(x), P(np.array([1.0])), 0.0,
           1.0, 1e-6, 5000, P(phi), C.byref(r))
        assert r.converged == 1, r.final_residual


def _direct_solve(g, rho, ex, ey, ez, k2, phi_bc):
    """scipy direct solve of the discrete LPBE (stencil_at, pb_solver.rs:159-194)
    with Dirichlet values phi_bc on the boundary."""
    nx, ny, nz = g.n[0], g.n[1], g.n[2]
    h2 = 1.0 / (g.h[0] * g.h[0])
    idx = np.arange(nx * ny * nz).reshape(nz, ny, nx)
    inner = idx[1:-1, 1:-1, 1:-1].ravel()
    pos = -np.ones(nx * ny * nz, np.int64)
    pos[inner] = np.arange(inner.size)
    rows, cols, vals = [], [], []
    b = 4 * math.pi * 332.0522 * rho[inner].copy()
    diag = np.zeros(inner.size)
    for off, ep, em in ((1, ex[inner], ex[inner - 1]), (nx, ey[inner], ey[inner - nx]),
                        (nx * ny, ez[inner], ez[inner - nx * ny])):
        diag += (ep + em) * h2
        for nb, e in ((inner + off, ep), (inner - off, em)):
            m = pos[nb] >= 0
            rows.append(np.arange(inner.size)[m])
            cols.append(pos[nb][m])
            vals.append(-e[m] * h2)
            b[~m] += e[~m] * h2 * phi_bc[nb[~m]]
    diag += k2[inner]
    A = sparse.csr_matrix((np.concatenate(vals + [diag]),
                           (np.concatenate(rows + [np.arange(inner.size)]),
                            np.concatenate(cols + [np.arange(inner.size)]))),
                          shape=(inner.size, inner.size))
    out = phi_bc.copy()
    out[inner] = spsolve(A, b)
    return out


@pytest.mark.parametrize("eps_out,kappa", [(80.0, 0.125733), (4.0, 0.0)])
def test_multigrid_matches_direct_solve(oracle, eps_out, kappa):
    """On 17^3 (levels 17-9-5, all odd: no boundary drift) the multigrid
    iterate at a tight tolerance equals scipy's direct solution."""
    L = oracle.lib()
    top, x = random_cluster(12, 7, box=5.0)
    g = oracle.OrcGrid()
    L.orc_auto_grid(12, P(x), 0.8, 0.5, C.byref(g))
    for k in range(3):
        g.n[k] = 17
        g.o[k] = x[:, k].mean() - 8 * 0.8
    q = top.charges.copy()
    rho, ex, ey, ez, k2 = _maps(oracle, g, x, top.radii, q, 4.0, eps_out, kappa)
    phi = np.zeros_like(rho)
    L.orc_dh_boundary(C.byref(g), 12, P(x), P(q), kappa, eps_out, P(phi))
    bc = phi.copy()
    r = oracle.OrcSolve()
    L.orc_solve_multigrid(C.byref(g), P(rho), P(ex), P(ey), P(ez), P(k2), 2, 12, P(x), P(q), kappa,
                          eps_out, 1e-10, 100000, P(phi), C.byref(r))
    assert r.converged == 1
    ref = _direct_solve(g, rho, ex, ey, ez, k2, bc)
    assert np.max(np.abs(phi - ref)) < 1e-8 * max(1.0, np.max(np.abs(ref)))


def test_dh_boundary_closed_form(oracle):
    L = oracle.lib()
    g = oracle.OrcGrid()
    for k in range(3):
        g.n[k] = 5
        g.h[k] = 1.0
        g.o[k] = -2.0
    x = np.array([[0.3, -0.2, 0.1]])
    q = np.array([0.7])
    phi = np.zeros(125)
    L.orc_dh_boundary(C.byref(g), 1, P(x), P(q), 0.1, 80.0, P(phi))
    node = np.array([-2.0, 1.0, 2.0])  # (0, 3, 4)
    r = np.linalg.norm(node - x[0])
    assert abs(phi[0 + 5 * (3 + 5 * 4)] - 332.0522 * 0.7 * math.exp(-0.1 * r) / (80 * r)) < 1e-13
    assert phi[2 + 5 * (2 + 5 * 2)] == 0.0  # interior untouched


def test_ses_brute_force(oracle):
    """SAS, connected exterior and the SES test against naive BFS and an
    O(G^2) distance search on a small grid."""
    L = oracle.lib()
    top, x = random_cluster(25, 11, box=7.0, min_sep=1.1)
    g = oracle.OrcGrid()
    L.orc_auto_grid(25, P(x), 0.7, 2.5, C.byref(g))
    nx, ny, nz = g.n[0], g.n[1], g.n[2]
    N = nx * ny * nz
    inside, sas, ext = (np.zeros(N, np.uint8) for _ in range(3))
    u8 = C.POINTER(C.c_uint8)
    L.orc_inside_ses(C.byref(g), 25, P(x), P(top.radii), 1.4, inside.ctypes.data_as(u8),
                     sas.ctypes.data_as(u8), ext.ctypes.data_as(u8))
    zz, yy, xx = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    pts = np.stack([g.o[0] + xx * g.h[0], g.o[1] + yy * g.h[1], g.o[2] + zz * g.h[2]], -1).reshape(-1, 3)
    d2 = ((pts[:, None, :] - x[None]) ** 2).sum(-1)
    sas_bf = (d2 < (top.radii + 1.4)[None] ** 2).any(1)
    assert np.array_equal(sas.astype(bool), sas_bf)
    # naive BFS
    from collections import deque
    occ = sas_bf.reshape(nz, ny, nx)
    seen = np.zeros_like(occ)
    dq = deque()
    for k, j, i in zip(*np.nonzero(~occ)):
        if i in (0, nx - 1) or j in (0, ny - 1) or k in (0, nz - 1):
            seen[k, j, i] = True
            dq.append((k, j, i))
    while dq:
        k, j, i = dq.popleft()
        for dk, dj, di in ((1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1)):
            a, b_, c = k + dk, j + dj, i + di
            if 0 <= a < nz and 0 <= b_ < ny and 0 <= c < nx and not occ[a, b_, c] and not seen[a, b_, c]:
                seen[a, b_, c] = True
                dq.append((a, b_, c))
    assert np.array_equal(ext.astype(bool), seen.ravel())
    # SES: inside iff no exterior node within probe/h (index distance)
    ei = np.argwhere(seen)
    ii = np.stack([zz, yy, xx], -1).reshape(-1, 3)
    dmin = np.full(N, 2 * N, float)
    for blk in range(0, N, 4096):
        dd = ((ii[blk:blk + 4096, None, :] - ei[None]) ** 2).sum(-1)
        dmin[blk:blk + 4096] = dd.min(1) if ei.size else 2 * N
    th = (1.4 / g.h[0]) ** 2
    assert np.array_equal(inside.astype(bool), dmin > th)


def test_ses_shell_cavity(oracle):
    """The connected exterior around an enclosed cavity: inside a closed
    hollow shell the annulus between the centre atom's and the shell's SAS
    is outside the SAS but not reachable (SES interior); with a cap removed
    the same points are exterior."""
    from _topo import hollow_shell
    L = oracle.lib()
    u8 = C.POINTER(C.c_uint8)
    for open_cap in (False, True):
        top, x = hollow_shell(open_cap)
        n = len(x)
        g = oracle.OrcGrid()
        L.orc_auto_grid(n, P(x), 0.5, 10.0, C.byref(g))
        nx, ny, nz = g.n[0], g.n[1], g.n[2]
        N = nx * ny * nz
        inside, sas, ext = (np.zeros(N, np.uint8) for _ in range(3))
        L.orc_inside_ses(C.byref(g), n, P(x), P(top.radii), 1.4, inside.ctypes.data_as(u8),
                         sas.ctypes.data_as(u8), ext.ctypes.data_as(u8))
        for pt in ([0.0, 0.0, 3.3], [3.3, 0.0, 0.0], [0.0, 0.0, -3.3], [0.0, -3.3, 0.0]):
            i = [int(round((pt[d] - g.o[d]) / g.h[d])) for d in range(3)]
            k = i[0] + nx * (i[1] + ny * i[2])
            assert sas[k] == 0
            assert ext[k] == (1 if open_cap else 0)
            assert inside[k] == (0 if open_cap else 1)


def test_face_fraction_sampling(oracle):
    L = oracle.lib()
    top, x = random_cluster(20, 5, box=6.0)
    rng = np.random.default_rng(1)
    for _ in range(40):
        axis = int(rng.integers(0, 3))
        lo = rng.uniform(-2, 2)
        hi = lo + 0.5
        p1, p2 = rng.uniform(-2, 2, 2)
        f = L.orc_fraction_inside(lo, hi, p1, p2, axis, 20, P(x), P(top.radii), 0.0)
        t = np.linspace(lo, hi, 20001)[:-1] + 0.5 * (hi - lo) / 20000
        pts = np.zeros((t.size, 3))
        pts[:, axis] = t
        a1 = 1 if axis == 0 else 0
        a2 = 1 if axis == 2 else 2
        pts[:, a1] = p1
        pts[:, a2] = p2
        ins = (((pts[:, None] - x[None]) ** 2).sum(-1) < top.radii[None] ** 2).any(1)
        assert abs(f - ins.mean()) < 2e-4


def test_gb_hct_integral_quadrature(oracle):
    """Two-atom HCT descreening sum against quadrature of the descreening
    integral 1/2 int (1 - cos theta(t)) / t^2 dt over the part of j's scaled
    sphere outside i's intrinsic sphere (far and partial-overlap cases)."""
    for r, rad_j, scr in ((6.0, 1.7, 0.8), (2.2, 1.5, 0.85), (1.6, 1.8, 0.9), (25.0, 1.2, 0.72)):
        top = make_topology([0.0, 0.0], [1.5, rad_j], screen=[0.8, scr])
        x = np.array([[0.0, 0, 0], [r, 0, 0]])
        _, R = oracle.gb_energy(top, x, oracle.gb_params(model=0))
        psi = 1.0 / (1.5 - 0.09) - 1.0 / R[0]
        rho_i, s = 1.5 - 0.09, scr * (rad_j - 0.09)
        lo, hi = max(rho_i, abs(r - s)), r + s

        def f(t):
            c = np.clip((t * t + r * r - s * s) / (2 * t * r), -1, 1)
            return 0.5 * (1 - c) / (t * t)
        ref, _ = integrate.quad(f, lo, hi, epsabs=1e-15, epsrel=1e-13, limit=200)
        assert abs(psi - ref) <= 1e-11 * max(1.0, abs(ref)), (r, psi, ref)
        # far field: (1/r) sum k/(2k+1) x^(2k+1)
        if rho_i <= r - s:
            xx = s / r
            ser = sum(k / (2 * k + 1) * xx ** (2 * k + 1) for k in range(1, 40)) / r
            assert abs(psi - ser) <= 1e-12 * max(ser, 1e-300) + 1e-16


@pytest.mark.parametrize("model", [0, 1, 2])
@pytest.mark.parametrize("cutoff", [9999.0, 4.5])
def test_gb_radii_multi_atom_quadrature(oracle, model, cutoff):
    """Born radii of a 12-atom cluster (close and overlapping pairs, one atom
    engulfed by a neighbour's scaled sphere) against an independent evaluation:
    each pair's descreening integral 1/2 int (1 - cos theta(t)) / t^2 dt from
    rho_i to r + s_j by quadrature, summed over the pairs within the cutoff,
    then mapped by HCT (1/R = 1/rho - psi) or by the published OBC map
    (Onufriev, Bashford and Case, Proteins 55:383 (2004): 1/R = 1/rho -
    tanh(a P - b P^2 + c P^3) / radius with P = psi rho; OBC I (0.8, 0,
    2.909125), OBC II (1.0, 0.8, 4.85)), capped at rgbmax.  For an engulfed
    atom (rho_i < s_j - r) the reference adds the inner shell's integral
    1/rho_i - 1/(s_j - r) twice, once more than the integral itself
    (gb_energy.rs:207-211), and so does this evaluation."""
    rng = np.random.default_rng(7)
    n = 12
    x = rng.uniform(-3.5, 3.5, (n, 3))
    radii = rng.choice([1.2, 1.5, 1.7, 1.8, 2.0], n)
    screen = rng.uniform(0.7, 0.9, n)
    radii[0], radii[1], screen[1] = 1.0, 3.0, 0.9
    x[0] = x[1] + np.array([0.4, 0.1, 0.0])
    top = make_topology(np.zeros(n), radii, screen=screen)
    gp = oracle.gb_params(model=model, cutoff=cutoff)
    _, R = oracle.gb_energy(top, x, gp)
    rho = radii - gp.offset
    psi = np.zeros(n)
    engulfed = 0
    for i in range(n):
        for j in range(n):
            r = float(np.linalg.norm(x[i] - x[j]))
            if i == j or r > cutoff:
                continue
            s = screen[j] * rho[j]
            if rho[i] >= r + s:
                continue

            def f(t):
                c = np.clip((t * t + r * r - s * s) / (2 * t * r), -1, 1)
                return 0.5 * (1 - c) / (t * t)
            kink = abs(r - s)
            pts = [kink] if rho[i] < kink < r + s else None
            val, _ = integrate.quad(f, rho[i], r + s, points=pts, epsabs=1e-15, epsrel=1e-13,
                                    limit=200)
            if rho[i] < s - r:
                val += 1.0 / rho[i] - 1.0 / (s - r)
                engulfed += 1
            psi[i] += val
    assert engulfed == 1
    if model == 0:
        inv = 1.0 / rho - psi
    else:
        a, b, c = (0.8, 0.0, 2.909125) if model == 1 else (1.0, 0.8, 4.85)
        P = psi * rho
        inv = 1.0 / rho - np.tanh(a * P - b * P * P + c * P ** 3) / radii
    ref = np.where(inv > 0, 1.0 / np.where(inv > 0, inv, 1.0), gp.rgbmax)
    ref = np.minimum(ref, gp.rgbmax)
    assert np.all(np.abs(R - ref) <= 1e-10 * ref), (model, cutoff, np.max(np.abs(R - ref) / ref))


def test_gb_energy_limits(oracle):
    top, x = random_cluster(30, 2, box=10.0)
    # eps_in == eps_out and no salt: zero exactly
    e, _ = oracle.gb_energy(top, x, oracle.gb_params(solvent_dielectric=1.0, salt_concentration=0.0))
    assert e == 0.0
    # salt -> 0 converges to the no-salt branch
    e0, _ = oracle.gb_energy(top, x, oracle.gb_params(salt_concentration=0.0))
    e1, _ = oracle.gb_energy(top, x, oracle.gb_params(salt_concentration=1e-12))
    assert abs(e0 - e1) < 1e-6 * abs(e0)
    # two distant atoms: pair -> -(1/ein - exp(-0.73 kappa r)/eout) qi qj / r (+ self terms)
    t2 = make_topology([0.5, -0.3], [1.5, 1.5])
    xx = np.array([[0.0, 0, 0], [400.0, 0, 0]])
    p = oracle.gb_params(model=1)
    e, R = oracle.gb_energy(t2, xx, p)
    k = oracle.compute_kappa(0.15, 78.3, 300.0)
    qa = t2.charges_amber
    self_ = sum(-0.5 * q * q * (1 - math.exp(-0.73 * k * r) / 78.3) / r for q, r in zip(qa, R))
    pair = -qa[0] * qa[1] * (1 - math.exp(-0.73 * k * 400.0) / 78.3) / 400.0
    assert abs(e - (self_ + pair)) < 1e-9


def test_mm_pair_closed_forms(oracle):
    # LJ minimum: A = eps R^12, B = 2 eps R^6 -> -eps at r = R; Coulomb q q / r
    eps, R = 0.15, 3.4
    top = make_topology([0.2, -0.4], [1.5, 1.5], lj_a=[eps * R ** 12], lj_b=[2 * eps * R ** 6])
    m = oracle.mm_energy(top, np.array([[0.0, 0, 0], [R, 0, 0]]))
    assert abs(m["vdw"] + eps) < 1e-12
    assert abs(m["elec"] - 0.2 * -0.4 * 18.2223 ** 2 / R) < 1e-12
    # 4-atom chain: 1-2/1-3 excluded, 1-4 scaled, 1-4 listed as excluded too
    # (AMBER lists it in both; 1-4 wins, mm_energy.rs:230-252)
    top = make_topology([0.3, -0.1, 0.2, -0.4], [1.5] * 4, lj_a=[eps * R ** 12],
                        lj_b=[2 * eps * R ** 6], dihedrals=[[0, 1, 2, 3, 0, 0]],
                        num_excluded=[3, 2, 1, 0], excluded=[1, 2, 3, 2, 3, 3], scee=1.2, scnb=2.0)
    x = np.array([[0.0, 0, 0], [1.5, 0, 0], [1.5, 1.5, 0], [3.0, 1.5, 0.4]])
    m = oracle.mm_energy(top, x)
    r = np.linalg.norm(x[3] - x[0])
    assert m["vdw"] == 0.0 and m["elec"] == 0.0
    assert abs(m["vdw_14"] - (eps * R ** 12 / r ** 12 - 2 * eps * R ** 6 / r ** 6) / 2.0) < 1e-12
    assert abs(m["elec_14"] - 0.3 * -0.4 * 18.2223 ** 2 / r / 1.2) < 1e-12


def test_pb_zero_when_media_equal(oracle, cfg0):
    top, x, _, _, _ = cfg0
    p = oracle.pb_params(grid_spacing=1.0, solute_dielectric=4.0, solvent_dielectric=4.0,
                         salt_concentration=0.0, fine_dims=(33, 33, 33))
    r = oracle.pb_energy(top, x, p)
    assert r.total == 0.0


@pytest.mark.slow
def test_pb_born_ion(oracle):
    # pb_energy.rs:529-570: within 15% of the Born energy
    top = make_topology([1.0], [2.0])
    p = oracle.pb_params(grid_spacing=0.25, grid_buffer=15.0, solute_dielectric=1.0,
                         solvent_dielectric=80.0, salt_concentration=0.0, probe_radius=0.0,
                         ion_radius=0.0, tolerance=1e-8, max_iterations=20000)
    r = oracle.pb_energy(top, np.zeros((1, 3)), p)
    born = -0.5 * (1 - 1 / 80.0) * 332.0522 / 2.0
    assert abs((r.total - born) / born) < 0.15


def test_decomposition_sums(oracle):
    """Per-residue sums reproduce the cross-pair, delta-GB and delta-SA totals."""
    from rust_simulation_tools import synthetic
    top, x, rec, lig = synthetic.build_complex(600, 80, seed=77)
    rr, ll = oracle.decompose(top, x, rec, lig)
    fe = oracle.frame_energy(top, x, rec, lig, method=0)
    allr = rr + ll
    vdw = sum(d["vdw"] for d in allr)
    elec = sum(d["elec"] for d in allr)
    m_c = oracle.mm_energy(top, x)
    from rust_simulation_tools.topology import extract_subtopology
    ra, la = top.build_selection(rec), top.build_selection(lig)
    m_r = oracle.mm_energy(extract_subtopology(top, ra), x[ra])
    m_l = oracle.mm_energy(extract_subtopology(top, la), x[la])
    dv = m_c["vdw"] - m_r["vdw"] - m_l["vdw"]
    de = m_c["elec"] - m_r["elec"] - m_l["elec"]
    assert abs(vdw - dv) < 1e-6 * max(1.0, abs(dv))
    assert abs(elec - de) < 1e-8 * max(1.0, abs(de))
    pol = sum(d["polar"] for d in allr)
    gbp = oracle.gb_params()
    dgb = oracle.gb_energy(top, x, gbp)[0] - \
        oracle.gb_energy(extract_subtopology(top, ra), x[ra], gbp)[0] - \
        oracle.gb_energy(extract_subtopology(top, la), x[la], gbp)[0]
    assert abs(pol - dgb) < 1e-8
    sa = sum(d["sa"] for d in allr)
    assert abs(sa - fe["delta_sa"]) < 1e-9


def test_decomposition_gb_zero_shares(oracle):
    """The reference shares dGB out among residues in proportion to
    |q^2 (1/R_complex - 1/R_alone)| summed over their atoms
    (decomposition.rs:309-351).  Independently of the proportions, a residue
    of neutral atoms and a residue farther than the GB cutoff from every
    partner atom (its atoms' Born radii cannot change) must get exactly zero,
    and the two interface residues then carry all of dGB.  Sharing pair terms,
    or spreading dGB by charge or atom count, would give them a share."""
    rng = np.random.default_rng(5)
    centres = np.array([[0.0, 0, 0], [0, 4.5, 0], [22.0, 0, 0], [26.0, 0, 0], [0, 0, 4.5]])
    x = np.concatenate([c + rng.normal(0, 0.9, (4, 3)) for c in centres])
    q = rng.normal(0, 0.5, 20)
    q[4:8] = 0.0  # residue 1: neutral
    radii = rng.choice([1.2, 1.5, 1.7], 20)
    top = make_topology(q, radii, residue_pointers=(0, 4, 8, 12, 16))
    gb = oracle.gb_params(cutoff=8.0)
    lig = x[16:20]
    far = np.linalg.norm(x[8:16, None, :] - lig[None, :, :], axis=-1).min()
    assert far > 8.0 + 1.0
    rr, ll = oracle.decompose(top, x, [0, 1, 2, 3], [4], gb=gb)
    pol = {d["residue_index"]: d["polar"] for d in rr + ll}
    from rust_simulation_tools.topology import extract_subtopology
    ra, la = top.build_selection([0, 1, 2, 3]), top.build_selection([4])
    dgb = oracle.gb_energy(top, x, gb)[0] - \
        oracle.gb_energy(extract_subtopology(top, ra), x[ra], gb)[0] - \
        oracle.gb_energy(extract_subtopology(top, la), x[la], gb)[0]
    assert pol[1] == 0.0 and pol[2] == 0.0 and pol[3] == 0.0
    assert pol[0] != 0.0 and pol[4] != 0.0 and abs(dgb) > 1e-3
    assert abs(pol[0] + pol[4] - dgb) <= 1e-12 * max(1.0, abs(dgb))


def _pb_small(**kw):
    base = dict(grid_spacing=1.0, grid_buffer=4.0, solute_dielectric=4.0,
                solvent_dielectric=80.0, salt_concentration=0.15, tolerance=1e-6)
    base.update(kw)
    return base


def test_pb_explicit_grids_match_auto(oracle):
    """Explicit focusing grids set to the geometry auto_grid and
    build_coarse_grid (pb_grid.rs:188-215, pb_energy.rs:134-148) would pick
    reproduce the reference path bit for bit; a shifted centre does not."""
    from rust_simulation_tools import synthetic
    top, x, _, _ = synthetic.build_complex(300, 40, seed=91)
    p0 = oracle.pb_params(**_pb_small())
    r0 = oracle.pb_energy(top, x, p0)
    mn, mx = x.min(axis=0), x.max(axis=0)
    c = 0.5 * (mn + mx)
    cs = 1.0 * 8
    cd = [int(math.ceil((mx[k] - mn[k]) * 4.0 / cs)) + 1 for k in range(3)]
    cd = [m + 1 if m % 2 == 0 else m for m in cd]
    assert list(r0.coarse_dims) == cd
    p1 = oracle.pb_params(**_pb_small(fine_dims=tuple(r0.fine_dims), fine_center=tuple(c),
                                      coarse_dims=tuple(cd), coarse_spacing=cs,
                                      coarse_center=tuple(c)))
    r1 = oracle.pb_energy(top, x, p1)
    assert r1.total == r0.total and r1.solv.energy == r0.solv.energy
    assert list(r1.coarse_dims) == cd and list(r1.fine_dims) == list(r0.fine_dims)
    p2 = oracle.pb_params(**_pb_small(fine_dims=tuple(r0.fine_dims),
                                      fine_center=tuple(c + 0.37), coarse_dims=tuple(cd),
                                      coarse_spacing=cs, coarse_center=tuple(c)))
    assert oracle.pb_energy(top, x, p2).total != r0.total


def test_pb_residue_terms_sum_to_delta_pb(oracle):
    """The per-atom terms 0.5 q (phi_solv - phi_ref)(r) sum, over all residues
    of both partners, to dG_PB = G_complex - G_receptor - G_ligand, each G from
    compute_pb_energy's own gather (pb_solver.rs:1174-1186); the order of the
    requested residues only permutes the output (to rounding)."""
    from rust_simulation_tools import synthetic
    top, x, rec, lig = synthetic.build_complex(400, 48, seed=92)
    p = oracle.pb_params(**_pb_small())
    tr, tl, d = oracle.pb_residue_terms(top, x, rec, lig, p)
    assert abs(tr.sum() + tl.sum() - d) < 1e-9 * max(1.0, abs(d))
    assert np.abs(tr).max() > 1e-3 and d != 0.0
    tr2, tl2, d2 = oracle.pb_residue_terms(top, x, list(rec)[::-1], lig, p)
    # (the receptor's atoms are taken in the requested residue order, so the
    # sums round differently)
    assert abs(d2 - d) < 1e-10 * abs(d)
    assert np.allclose(tr2, tr[::-1], rtol=0, atol=1e-9) and np.allclose(tl2, tl, rtol=0, atol=1e-9)


def test_pb_residue_terms_far_apart_closed_form(oracle):
    """Partners 40 A apart on one shared explicit grid, at 0.15 M salt (Debye
    length ~8 A): in the solvated medium the far partner's field is screened
    away, while the reference medium (eps_out = eps_in, no ions) keeps its
    unscreened Coulomb field.  So each atom's term tends to
    -0.5 q_i sum_j C q_j / (eps_in r_ij) over the other partner's atoms, and
    every residue's PB term to the sum of those (within ~1%, the remainder
    being the boundary and cavity effects that fall off with distance)."""
    from rust_simulation_tools import synthetic
    top, x, rec, lig = synthetic.build_complex(160, 32, seed=93)
    ra, la = top.build_selection(rec), top.build_selection(lig)
    x = x.copy()
    x[la] += np.array([40.0, 0.0, 0.0]) - (x[la].mean(axis=0) - x[ra].mean(axis=0))
    c = 0.5 * (x.min(axis=0) + x.max(axis=0))
    p = oracle.pb_params(**_pb_small(fine_dims=(73, 33, 33), fine_center=tuple(c),
                                     coarse_dims=(37, 17, 17), coarse_spacing=4.0,
                                     coarse_center=tuple(c)))
    tr, tl, d = oracle.pb_residue_terms(top, x, rec, lig, p)
    q = np.asarray(top.charges)
    r = np.linalg.norm(x[ra][:, None, :] - x[la][None, :, :], axis=2)
    pair = 332.0522 * q[ra][:, None] * q[la][None, :] / (4.0 * r)
    ptr = np.append(np.asarray(top.residue_pointers), top.n_atoms)
    ares = np.searchsorted(ptr, np.arange(top.n_atoms), side="right") - 1
    er = np.array([-0.5 * pair.sum(axis=1)[ares[ra] == k].sum() for k in rec])
    el = np.array([-0.5 * pair.sum(axis=0)[ares[la] == k].sum() for k in lig])
    scale = max(np.abs(er).max(), np.abs(el).max())
    assert scale > 0.1
    assert np.abs(tr - er).max() < 0.02 * scale and np.abs(tl - el).max() < 0.02 * scale
    assert abs(tr.sum() + tl.sum() - d) < 1e-9 * max(1.0, abs(d))


def _gb_from_radii(oracle, top, x, p, R):
    L = oracle.lib()
    v = top.view()
    x = np.ascontiguousarray(x, dtype=np.float64)
    R = np.ascontiguousarray(R, dtype=np.float64)
    return L.orc_gb_energy_from_radii(C.byref(v), P(x), C.byref(p), P(R))


def test_gb_pair_distance_still_formula(oracle):
    """The GB pair term at intermediate distances (gb_energy.rs:322-352).  The
    effective distance the oracle uses is read back from its energy at fixed
    Born radii (no salt: pair = -q_i q_j (1/eps_in - 1/eps_out) / f) and
    compared with Still et al., JACS 112:6127 (1990):
    f = sqrt(r^2 + R_i R_j exp(-r^2 / (4 R_i R_j))), which is sqrt(R_i R_j) at
    r = 0 and tends to r for r >> R.  With salt the same f enters
    exp(-0.73 kappa f) (the screening factor of the OBC papers)."""
    Ri, Rj = 1.3, 2.1
    top = make_topology([0.5, -0.3], [1.5, 1.5])
    qa = np.asarray(top.charges_amber)
    R = np.array([Ri, Rj])
    rr = Ri * Rj
    p0 = oracle.gb_params(salt_concentration=0.0)
    ps = oracle.gb_params(salt_concentration=0.15)
    k = oracle.compute_kappa(0.15, 78.3, 300.0)
    far = np.array([[0.0, 0, 0], [5000.0, 0, 0]])  # inside the default 9999 A cutoff
    self0 = _gb_from_radii(oracle, top, far, p0, R) + qa[0] * qa[1] * (1 - 1 / 78.3) / 5000.0
    selfs = sum(-0.5 * q * q * (1 - math.exp(-0.73 * k * r) / 78.3) / r for q, r in zip(qa, R))
    for r in (0.0, 0.6, math.sqrt(rr), 2.5, Ri + Rj, 6.0, 15.0):
        x = np.array([[0.0, 0, 0], [r * 0.6, r * 0.8, 0.0]])
        pair0 = _gb_from_radii(oracle, top, x, p0, R) - self0
        f_seen = -qa[0] * qa[1] * (1 - 1 / 78.3) / pair0
        f_still = math.sqrt(r * r + rr * math.exp(-r * r / (4 * rr)))
        assert abs(f_seen - f_still) < 1e-9 * f_still, (r, f_seen, f_still)
        pairs = _gb_from_radii(oracle, top, x, ps, R) - selfs
        want = -qa[0] * qa[1] * (1 - math.exp(-0.73 * k * f_still) / 78.3) / f_still
        assert abs(pairs - want) < 1e-10 * abs(want), (r, pairs, want)
    # r >> R: f -> r
    x = np.array([[0.0, 0, 0], [40.0, 0, 0]])
    f40 = -qa[0] * qa[1] * (1 - 1 / 78.3) / (_gb_from_radii(oracle, top, x, p0, R) - self0)
    assert abs(f40 - 40.0) < 1e-9


@pytest.mark.parametrize("a", [2.0, 3.0])
def test_pb_salt_born_ion_debye_huckel(oracle, a):
    """The kappa-bar^2 = eps_s kappa^2 map (pb_grid.rs:951-985) in isolation:
    one ion of radius a in a uniform eps = 80 medium (eps_in = eps_out, so the
    solvated and reference problems differ only by the salt), ion radius 0 so
    the salt is excluded from the sphere r < a.  Debye-Hueckel for a charge at
    the centre of a salt-free sphere gives the extra potential at the charge
    -C q kappa / (eps (1 + kappa a)), so dG_PB = -(C q^2 / 2) kappa /
    (eps (1 + kappa a)), C = 332.0522.  Dropping eps_s from the map (kappa_eff
    = kappa / sqrt(80)) or doubling it moves dG by far more than the 2% bar."""
    top = make_topology([1.0], [a])
    p = oracle.pb_params(grid_spacing=0.5, grid_buffer=16.0 if a == 2.0 else 20.0,
                         solute_dielectric=80.0, solvent_dielectric=80.0,
                         salt_concentration=0.15, probe_radius=0.0, ion_radius=0.0,
                         tolerance=1e-9, max_iterations=20000, fillratio=0.0)
    r = oracle.pb_energy(top, np.zeros((1, 3)), p)
    k = oracle.compute_kappa(0.15, 80.0, 300.0)
    want = -0.5 * 332.0522 * k / (80.0 * (1.0 + k * a))
    assert r.solv.converged and r.ref.converged
    assert abs(r.total - want) < 0.02 * abs(want), (r.total, want)


def test_coarse_grid_invariants(oracle):
    """build_coarse_grid (pb_energy.rs:92-125) on point clouds, through what a
    correct coarse grid must satisfy on each axis: spacing h * fscale, an odd
    node count, centred on the bounding box, spanning at least fillratio times
    the extent, and the fewest odd nodes that do ((n - 3) cs < F ext)."""
    L = oracle.lib()
    rng = np.random.default_rng(23)
    for trial in range(60):
        n = int(rng.integers(2, 80))
        x = rng.normal(0, rng.uniform(0.5, 30), (n, 3)) + rng.uniform(-80, 80, 3)
        h = float(rng.choice([0.3, 0.5, 0.6, 1.0]))
        F = float(rng.choice([1.5, 2.0, 4.0, 6.0]))
        s = int(rng.choice([2, 4, 8]))
        g = oracle.OrcGrid()
        L.orc_coarse_grid(n, P(np.ascontiguousarray(x)), h, F, s, C.byref(g))
        mn, mx = x.min(0), x.max(0)
        cs = h * s
        for k in range(3):
            nk, o = g.n[k], g.o[k]
            assert g.h[k] == cs and nk % 2 == 1
            span = (nk - 1) * cs
            ext = F * (mx[k] - mn[k])
            assert abs((o + 0.5 * span) - 0.5 * (mn[k] + mx[k])) <= 1e-12 * (1 + abs(mn[k]) + abs(mx[k]))
            assert span >= ext - 1e-9 * max(1.0, ext)
            assert (nk - 3) * cs < ext


def test_decomposition_gb_weight_charge_power(oracle):
    """The GB share of a residue is proportional to sum |q^2 (1/R_complex -
    1/R_alone)| over its atoms (decomposition.rs:309-351).  Born radii do not
    depend on charges, so scaling one interface residue's charges by c scales
    its share ratio to any other residue by exactly c^2 (whatever dGB does);
    a weight linear in |q|, or one taken from pair terms, scales by c or not
    at all."""
    rng = np.random.default_rng(8)
    centres = np.array([[0.0, 0, 0], [0, 4.2, 0], [4.2, 0, 0], [0, 0, 4.4]])
    x = np.concatenate([c + rng.normal(0, 0.9, (4, 3)) for c in centres])
    q = rng.normal(0, 0.5, 16)
    radii = rng.choice([1.2, 1.5, 1.7], 16)
    rec, lig = [0, 1, 2], [3]
    shares = []
    for c in (1.0, 2.5):
        qq = q.copy()
        qq[0:4] *= c
        top = make_topology(qq, radii, residue_pointers=(0, 4, 8, 12))
        rr, ll = oracle.decompose(top, x, rec, lig)
        shares.append({d["residue_index"]: d["polar"] for d in rr + ll})
    for other in (1, 2, 3):
        r1 = shares[0][0] / shares[0][other]
        r2 = shares[1][0] / shares[1][other]
        assert abs(shares[0][other]) > 1e-6
        assert abs(r2 / r1 - 2.5 ** 2) < 1e-9 * 6.25, (other, r2 / r1)


# ---- quasi-harmonic entropy (entropy.rs:97-275) ---------------------------
def test_jacobi_eigenvalues_known_answers(oracle):
    # the reference's own unit tests: identity, and [[2, 1], [1, 2]] -> 3, 1
    assert np.allclose(oracle.jacobi_eigenvalues(np.eye(4)), 1.0, rtol=0, atol=1e-10)
    e = oracle.jacobi_eigenvalues([[2.0, 1.0], [1.0, 2.0]])
    assert abs(e[0] - 3.0) < 1e-10 and abs(e[1] - 1.0) < 1e-10
    # a dense symmetric matrix against LAPACK's (a wrong rotation fails here)
    rng = np.random.default_rng(5)
    b = rng.normal(size=(12, 12))
    a = b @ b.T + np.diag(rng.uniform(0, 3, 12))
    want = np.sort(np.linalg.eigvalsh(a))[::-1]
    got = oracle.jacobi_eigenvalues(a)
    assert np.allclose(got, want, rtol=1e-10, atol=1e-10)


def _qho_entropy_over_k(x):
    """Entropy of one quantum harmonic oscillator in units of k_B, x = hbar
    omega / kT: S/k = x / (e^x - 1) - ln(1 - e^-x)."""
    return x / math.expm1(x) - math.log(-math.expm1(-x))


def test_quasi_harmonic_single_mode_closed_form(oracle):
    # one 12 amu atom at +-1 A along x over two frames (the reference's test
    # system): the mass-weighted variance is 12 A^2 amu (divided by F, not
    # F - 1) along one axis and 0 along the others, so there is one mode with
    # omega^2 = kT / lambda
    T = 298.15
    got = oracle.quasi_harmonic_entropy(np.array([[[1.0, 0, 0]], [[-1.0, 0, 0]]]), [12.0], T)
    lam = 12.0 * 1e-20 * 1.66054e-27  # m^2 kg
    x = 1.054571817e-34 / math.sqrt(1.380649e-23 * T * lam)
    want = -T * 0.00198688 * _qho_entropy_over_k(x)
    assert got == pytest.approx(want, rel=1e-12)
    # the classical limit of the same expression (hbar omega << kT):
    # S/k -> 1 - ln x; at this x (~0.3) the next term x^2 / 24 decides
    assert _qho_entropy_over_k(x) == pytest.approx(1 - math.log(x) + x * x / 24, rel=2e-4)


def test_quasi_harmonic_limits_and_none(oracle):
    # fewer than two frames or no atoms: the reference's None
    assert oracle.quasi_harmonic_entropy(np.zeros((1, 3, 3)), np.ones(3)) is None
    assert oracle.quasi_harmonic_entropy(np.zeros((4, 0, 3)), np.ones(0)) is None
    # no motion: every mode is zero and skipped
    still = np.tile(np.arange(9.0).reshape(1, 3, 3), (5, 1, 1))
    assert oracle.quasi_harmonic_entropy(still, [12.0, 14.0, 16.0]) == 0.0
    # low temperature: x large, S/k ~ (x + 1) e^-x, so -T dS -> 0 from below
    v = oracle.quasi_harmonic_entropy(np.array([[[0.01, 0, 0]], [[-0.01, 0, 0]]]), [1.0], 5.0)
    assert -1e-30 < v <= 0.0


def test_quasi_harmonic_invariances(oracle):
    rng = np.random.default_rng(11)
    na, nf = 6, 40
    masses = rng.choice([1.008, 12.011, 14.007, 15.999], na)
    base = rng.normal(scale=3.0, size=(na, 3))
    traj = base + rng.normal(scale=0.2, size=(nf, na, 3))
    s0 = oracle.quasi_harmonic_entropy(traj, masses)
    # a rigid rotation and a translation of every frame leave the covariance's
    # spectrum unchanged
    q, _ = np.linalg.qr(rng.normal(size=(3, 3)))
    moved = traj @ q.T + np.array([5.0, -2.0, 1.5])
    assert oracle.quasi_harmonic_entropy(moved, masses) == pytest.approx(s0, rel=1e-9)
    # independent partners: motions whose cross products sum to zero over the
    # frames give a block-diagonal covariance, and additive entropies
    a = np.array([[[0.3, 0.1, 0.0]], [[-0.3, -0.1, 0.0]], [[0.3, 0.1, 0.0]], [[-0.3, -0.1, 0.0]]])
    b = np.array([[[0.0, 0.2, 0.1]], [[0.0, 0.2, 0.1]], [[0.0, -0.2, -0.1]], [[0.0, -0.2, -0.1]]])
    sa = oracle.quasi_harmonic_entropy(a, [12.0])
    sb = oracle.quasi_harmonic_entropy(b, [16.0])
    sab = oracle.quasi_harmonic_entropy(np.concatenate([a, b], axis=1), [12.0, 16.0])
    assert sab == pytest.approx(sa + sb, rel=1e-12)
