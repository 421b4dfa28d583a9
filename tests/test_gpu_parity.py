"""GPU parity of every term against the FP64 oracle on the 2,500-atom complex.

Bars: Born radii 1e-5 relative, energy terms 0.01 kcal/mol, SASA per atom and
all index maps bit-exact, PB iteration counts identical to the oracle's
(the GPU runs the same V-cycle algebra in correction form).
"""
import numpy as np
import pytest

import rust_simulation_tools as rst

pytestmark = pytest.mark.gpu

TOL = 0.01
PB_CFG0 = dict(grid_spacing=0.6, solute_dielectric=4.0, solvent_dielectric=80.0,
               fine_dims=(65, 65, 65))


def as_nm(x):
    return np.asarray(x) / 10.0


def as_seen(x):
    """The Angstrom coordinates the library receives from an nm array."""
    return as_nm(x) * 10.0


@pytest.mark.parametrize("model", [0, 1, 2])
def test_born_radii_and_gb_energy(cfg0, oracle, model):
    top, x, _, _, _ = cfg0
    e = rst.compute_gb_energy(top, as_nm(x), rst.GbParams(model=model))
    oe, oR = oracle.gb_energy(top, as_seen(x), oracle.gb_params(model=model))
    rel = np.abs(e.born_radii() - oR) / oR
    print(f"model {model}: max rel Born radius err {rel.max():.3e}, GB {e.total:.6f} vs "
          f"{oe:.6f} (err {abs(e.total - oe):.3e})")
    assert rel.max() < 1e-5
    assert abs(e.total - oe) < TOL


def test_gb_no_salt_and_cutoff(cfg0, oracle):
    top, x, _, _, _ = cfg0
    for kw in (dict(salt_concentration=0.0), dict(cutoff=12.0, model=2)):
        e = rst.compute_gb_energy(top, as_nm(x), rst.GbParams(**kw))
        oe, oR = oracle.gb_energy(top, as_seen(x), oracle.gb_params(**kw))
        assert np.max(np.abs(e.born_radii() - oR) / oR) < 1e-5
        assert abs(e.total - oe) < TOL


def test_mm_terms(cfg0, oracle):
    top, x, _, _, _ = cfg0
    m = rst.compute_mm_energy(top, as_nm(x))
    om = oracle.mm_energy(top, as_seen(x))
    for k in ("bond", "angle", "dihedral", "vdw", "elec", "vdw_14", "elec_14"):
        print(k, getattr(m, k), om[k], abs(getattr(m, k) - om[k]))
        assert abs(getattr(m, k) - om[k]) < TOL, k


def test_sasa_bit_exact(cfg0, oracle):
    top, x, _, _, _ = cfg0
    s = rst.compute_sa_energy(top, as_nm(x))
    ot, ots, oper = oracle.sa_energy(top, as_seen(x), oracle.sa_params())
    assert np.array_equal(s.per_atom_sasa(), oper)
    assert abs(s.total_sasa - ots) < 1e-8
    assert abs(s.total - ot) < 1e-10


def test_pb_energy(cfg0, oracle):
    top, x, _, _, _ = cfg0
    r = rst.compute_pb_energy(top, as_nm(x), rst.PbParams(**PB_CFG0))
    o = oracle.pb_energy(top, as_seen(x), oracle.pb_params(**PB_CFG0))
    print(f"dG_PB {r.total:.6f} vs {o.total:.6f}; E_solv {r.e_solv:.6f} vs {o.solv.energy:.6f}; "
          f"E_ref {r.e_ref:.6f} vs {o.ref.energy:.6f}; cycles {r.iterations} vs "
          f"({o.solv.iterations}, {o.ref.iterations})")
    assert r.fine_dims == tuple(o.fine_dims) and r.coarse_dims == tuple(o.coarse_dims)
    assert r.iterations == (o.solv.iterations, o.ref.iterations)
    assert r.converged == bool(o.solv.converged and o.ref.converged)
    assert abs(r.e_solv - o.solv.energy) < TOL
    assert abs(r.e_ref - o.ref.energy) < TOL
    assert abs(r.total - o.total) < TOL


@pytest.mark.parametrize("method", ["gb", "pb"])
def test_frame_energy(cfg0, oracle, method):
    top, x, rec, lig, _ = cfg0
    kw = dict(pb_params=rst.PbParams(**PB_CFG0)) if method == "pb" else \
        dict(gb_params=rst.GbParams(model=2))
    f = rst.compute_binding_energy_single_frame(top, as_nm(x), rec, lig, **kw)
    o = oracle.frame_energy(top, as_seen(x), rec, lig, method=1 if method == "pb" else 0,
                            gb=oracle.gb_params(model=2), pb=oracle.pb_params(**PB_CFG0))
    for part in ("complex", "receptor", "ligand"):
        for term in ("mm", "polar", "sa"):
            a, b = getattr(f, f"{part}_{term}"), o[part][term]
            assert abs(a - b) < TOL, (part, term, a, b)
    for term in ("delta_mm", "delta_polar", "delta_sa", "delta_total"):
        assert abs(getattr(f, term) - o[term]) < TOL, term


def test_decomposition(cfg0, oracle):
    top, x, rec, lig, _ = cfg0
    d = rst.decompose_binding_energy(top, as_nm(x), rec, lig)
    orr, ol = oracle.decompose(top, as_seen(x), rec, lig)
    got = d.receptor_residues() + d.ligand_residues()
    ref = orr + ol
    assert [g.residue_index for g in got] == [r["residue_index"] for r in ref]
    worst = 0.0
    for g, r in zip(got, ref):
        for k in ("vdw", "elec", "polar", "sa"):
            worst = max(worst, abs(getattr(g, k) - r[k]))
    print("decomposition worst abs err", worst)
    assert worst < TOL


def test_batched_frames_match_single(cfg0):
    """Each frame's result is the same alone and inside a batch (shared
    launches, graph-replayed cycles)."""
    from rust_simulation_tools import synthetic
    top, _, rec, lig, seed = cfg0
    base = synthetic.config_system(0)[1]
    frames = np.stack([synthetic.noisy_frame(base, seed, f) for f in range(3)]) / 10.0
    pb = rst.PbParams(**PB_CFG0)
    batch = rst.compute_binding_energy_frames(top, frames, rec, lig, pb_params=pb, batched=True)
    plain = rst.compute_binding_energy_frames(top, frames, rec, lig, pb_params=pb, batched=False)
    for fb, fp in zip(batch.frames(), plain.frames()):
        for k in ("complex_polar", "receptor_polar", "ligand_polar", "delta_total", "delta_mm",
                  "delta_sa"):
            assert abs(getattr(fb, k) - getattr(fp, k)) <= 1e-9, k


def test_device_frames_match_host_frames(cfg0):
    """Frames already in HBM give exactly the results of host frames (the
    bounding boxes then come from a device reduction; min/max are exact)."""
    from rust_simulation_tools import synthetic
    top, _, rec, lig, seed = cfg0
    base = synthetic.config_system(0)[1]
    frames = np.stack([synthetic.noisy_frame(base, seed, f) for f in range(2)]) / 10.0
    pb = rst.PbParams(**PB_CFG0)
    dev = rst.DeviceArray(frames * 10.0)
    assert np.array_equal(dev.numpy(), frames * 10.0)
    for method in ("pb", "gb"):
        kw = dict(pb_params=pb) if method == "pb" else {}
        host = rst.compute_binding_energy_frames(top, frames, rec, lig, batched=True, **kw)
        on_dev = rst.compute_binding_energy_device_frames(top, dev, rec, lig, batched=True, **kw)
        one = rst.compute_binding_energy_device_frames(top, dev[1:2], rec, lig, batched=False,
                                                       **kw)
        for fh, fd in zip(host.frames(), on_dev.frames()):
            for k in ("complex_polar", "receptor_polar", "ligand_polar", "complex_mm",
                      "complex_sa", "delta_total"):
                assert getattr(fh, k) == getattr(fd, k), (method, k)
        assert one.frames()[0].delta_total == host.frames()[1].delta_total
