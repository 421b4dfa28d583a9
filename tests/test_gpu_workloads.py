"""The benchmark workloads at their full sizes on the GPU.

- 25,000-atom complex, 161^3 PB grid at 0.5 A (the throughput workload):
  one frame against the FP64 oracle, all 13 terms within 0.01 kcal/mol,
  OBC-II Born radii within 1e-5 relative, PB cycle counts equal.
- Uneven complexes through one server round (the batched and served
  workloads' shape): four complexes of 3,000-9,000 atoms on their own
  auto_grid grids (0.5 A, 10 A buffer), two frames each, sharing solver
  launches; every request's first frame against the oracle.
- 150,000-atom complex with focusing 193^3 at 1.0 A then 321^3 at 0.4 A and
  per-residue decomposition: the size-independent identities (residue PB
  terms sum to dG_PB, which equals the frame's dPolar; residue GB, vdW,
  elec and SA terms sum to the frame's deltas) and, for a sample of ligand
  residues, vdW and elec against an FP64 brute-force pair sum.

The oracle runs take tens of seconds each, hence the longer timeouts.
"""
import math

import numpy as np
import pytest

import rust_simulation_tools as rst
from rust_simulation_tools import synthetic

pytestmark = pytest.mark.gpu

TOL = 0.01
TERMS = ("delta_mm", "delta_polar", "delta_sa", "delta_total")
PARTS = ("complex", "receptor", "ligand")


def _frame_errors(f, o):
    err = {}
    for part in PARTS:
        for term in ("mm", "polar", "sa"):
            err[f"{part}_{term}"] = abs(getattr(f, f"{part}_{term}") - o[part][term])
    for term in TERMS:
        err[term] = abs(getattr(f, term) - o[term])
    return err


@pytest.mark.timeout(600)
def test_25k_atom_frame_161_grid_vs_oracle(oracle):
    top, base, rec, lig, seed = synthetic.config_system(1)
    x = (synthetic.noisy_frame(base, seed, 0) / 10.0) * 10.0
    kw = dict(grid_spacing=0.5, solute_dielectric=4.0, solvent_dielectric=80.0,
              salt_concentration=0.15, fine_dims=(161, 161, 161))
    f = rst.compute_binding_energy_single_frame(top, x / 10.0, rec, lig,
                                                pb_params=rst.PbParams(**kw))
    o = oracle.frame_energy(top, x, rec, lig, method=1, pb=oracle.pb_params(**kw))
    err = _frame_errors(f, o)
    print("per-term |GPU - FP64 oracle| (kcal/mol):",
          ", ".join(f"{k} {v:.2e}" for k, v in err.items()))
    assert max(err.values()) < TOL, err
    # PB cycle counts of the complex (coarse and fine, both media)
    r = rst.compute_pb_energy(top, x / 10.0, rst.PbParams(**kw))
    oc = o["complex"]["pb"]
    assert r.total == f.complex_polar
    assert r.iterations == (oc.solv.iterations, oc.ref.iterations)
    assert r.coarse_iterations == (oc.solv.coarse_iterations, oc.ref.coarse_iterations)
    assert r.fine_dims == (161, 161, 161) and r.coarse_dims == tuple(oc.coarse_dims)
    # Born radii of the complex (OBC-II)
    g = rst.compute_gb_energy(top, x / 10.0, rst.GbParams(model=2))
    og, oR = oracle.gb_energy(top, x, oracle.gb_params(model=2))
    rel = np.max(np.abs(g.born_radii() - oR) / oR)
    print(f"Born radii max rel err {rel:.2e}; GB {g.total:.6f} vs {og:.6f}")
    assert rel < 1e-5 and abs(g.total - og) < TOL


@pytest.mark.timeout(900)
def test_uneven_server_round_vs_oracle(oracle):
    rng = np.random.Generator(np.random.PCG64(synthetic.seed_for(2, 500)))
    sizes = (3000, 4500, 6000, 9000)
    kw = dict(grid_spacing=0.5, grid_buffer=10.0, solute_dielectric=4.0,
              solvent_dielectric=80.0, salt_concentration=0.15)
    cx = []
    for i, n in enumerate(sizes):
        seed = synthetic.seed_for(2, 500 + i)
        n_lig = max(16, int(round(n * rng.uniform(0.03, 0.10) / 16.0)) * 16)
        top, base, rec, lig = synthetic.build_complex(n, n_lig, None, seed)
        fr = np.stack([synthetic.noisy_frame(base, seed, k) for k in range(2)]) / 10.0
        cx.append((top, fr, rec, lig))
    srv = rst.Server(hbm_budget_bytes=200 << 30, max_round_frames=16)
    try:
        tickets = [srv.submit(top, fr, rec, lig, pb_params=rst.PbParams(**kw))
                   for top, fr, rec, lig in cx]
        results = [srv.wait(t)[0] for t in tickets]
    finally:
        srv.close()
    dims = set()
    worst = 0.0
    for (top, fr, rec, lig), res in zip(cx, results):
        assert len(res.frames()) == 2
        o = oracle.frame_energy(top, fr[0] * 10.0, rec, lig, method=1,
                                pb=oracle.pb_params(**kw))
        dims.add(tuple(o["complex"]["pb"].fine_dims))
        err = _frame_errors(res.frames()[0], o)
        worst = max(worst, max(err.values()))
        print(f"N={top.n_atoms}: grid {tuple(o['complex']['pb'].fine_dims)}, worst term err "
              f"{max(err.values()):.2e}")
        assert max(err.values()) < TOL, err
    assert len(dims) == len(sizes)


def _lj_coul(top, xa, ia, xb, ib):
    """FP64 brute force: sum over a x b of A/r^12 - B/r^6 and qa qb / r."""
    qa = np.asarray(top.charges_amber)
    ty = np.asarray(top.atom_type_indices)
    nt = top.n_types
    nbi = np.asarray(top.nb_parm_index)
    A, B = np.asarray(top.lj_acoef), np.asarray(top.lj_bcoef)
    d = np.linalg.norm(xa[:, None, :] - xb[None, :, :], axis=2)
    k = nbi[nt * ty[ia][:, None] + ty[ib][None, :]] - 1
    a = np.where(k >= 0, A[np.maximum(k, 0)], 0.0)
    b = np.where(k >= 0, B[np.maximum(k, 0)], 0.0)
    r6 = d ** 6
    return float((a / (r6 * r6) - b / r6).sum()), float((qa[ia][:, None] * qa[ib][None, :] / d).sum())


@pytest.mark.timeout(600)
def test_150k_atom_decomposition_identities():
    top, base, rec, lig, seed = synthetic.config_system(4)
    x = (synthetic.noisy_frame(base, seed, 0) / 10.0) * 10.0
    ra, la = top.build_selection(rec), top.build_selection(lig)
    kw = dict(solute_dielectric=4.0, solvent_dielectric=80.0, salt_concentration=0.15,
              **synthetic.focusing_grids(x, ra, la))
    pb = rst.PbParams(**kw)
    d, dpb = rst.decompose_binding_energy_pb(top, x / 10.0, rec, lig, pb_params=pb)
    f = rst.compute_binding_energy_single_frame(top, x / 10.0, rec, lig, pb_params=pb)
    allr = d.receptor_residues() + d.ligand_residues()
    assert len(allr) == len(rec) + len(lig)
    assert [c.residue_index for c in allr] == list(rec) + list(lig)
    pol = math.fsum(c.polar for c in allr)
    print(f"dG_PB {dpb:.6f}; residue sum - dG_PB {pol - dpb:.2e}; frame dPolar - dG_PB "
          f"{f.delta_polar - dpb:.2e}")
    assert abs(pol - dpb) < 1e-8 * max(1.0, abs(dpb))
    assert abs(f.delta_polar - dpb) < 1e-8 * max(1.0, abs(dpb))
    assert abs(math.fsum(c.sa for c in allr) - f.delta_sa) < 1e-6
    # GB decomposition of the same frame: residue sums equal the frame's deltas
    g = rst.decompose_binding_energy(top, x / 10.0, rec, lig)
    fg = rst.compute_binding_energy_single_frame(top, x / 10.0, rec, lig)
    gr = g.receptor_residues() + g.ligand_residues()
    assert abs(math.fsum(c.polar for c in gr) - fg.delta_polar) < 1e-6 * max(1.0, abs(fg.delta_polar))
    # no receptor-ligand bonds: dMM = the cross-pair sum = sum of residue vdW + elec
    # (the intra-partner terms cancel up to the rounding of ~1e5-atom sums)
    cross = math.fsum(c.vdw + c.elec for c in gr)
    print(f"sum of residue vdW + elec - dMM {cross - fg.delta_mm:.2e}")
    assert abs(cross - fg.delta_mm) < TOL
    # sampled ligand residues against an FP64 brute-force cross sum (1/2 of each pair)
    ptr = np.append(np.asarray(top.residue_pointers), top.n_atoms)
    rng = np.random.default_rng(3)
    worst = 0.0
    for k in rng.choice(len(lig), 12, replace=False):
        res = lig[k]
        ia = np.arange(ptr[res], ptr[res + 1])
        v, e = _lj_coul(top, x[ia], ia, x[ra], ra)
        c = g.ligand_residues()[k]
        worst = max(worst, abs(c.vdw - 0.5 * v), abs(c.elec - 0.5 * e))
        assert abs(c.vdw - 0.5 * v) < TOL and abs(c.elec - 0.5 * e) < TOL
    print(f"sampled ligand residues: worst vdW/elec err vs FP64 brute force {worst:.2e}")
