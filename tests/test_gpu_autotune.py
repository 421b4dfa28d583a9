"""Tuned launch shapes on the GPU: smoother regions, z-chunking and the host
check interval change no result bit, a Born-radius chunk count changes only the
rounding of its partial sums, and the search loop runs end to end on a small
workload (results unchanged, the table left applied and persistable)."""
import numpy as np
import pytest

import rust_simulation_tools as rst
from rust_simulation_tools import autotune, synthetic

pytestmark = pytest.mark.gpu

PB = dict(grid_spacing=0.6, solute_dielectric=4.0, solvent_dielectric=80.0,
          fine_dims=(65, 65, 65))
KEYS = ("complex_polar", "receptor_polar", "ligand_polar", "delta_total")


@pytest.fixture(autouse=True)
def restore():
    yield
    autotune.apply(None)


def _frames(cfg0, n=2):
    top, _, rec, lig, seed = cfg0
    base = synthetic.config_system(0)[1]
    return top, np.stack([synthetic.noisy_frame(base, seed, f) for f in range(n)]) / 10.0, rec, lig


def test_tuned_chunks_and_checks_bitwise(cfg0):
    top, fr, rec, lig = _frames(cfg0)
    pb = rst.PbParams(**PB)
    ref = rst.compute_binding_energy_frames(top, fr, rec, lig, pb_params=pb, batched=True)
    # regions: 1 = 32 x 32 nodes per workgroup, 2 = 44 x 44 (0: built-in)
    for chunks, checks, region in ((1, 2, 1), (5, 16, 2), (2, 8, 0), (3, 8, 1), (1, 4, 2)):
        autotune.apply(autotune.TuneTable([{"nodes_min": 0, "nodes_max": 1 << 40, "solves_min": 0,
                                            "solves_max": 1 << 20, "smoother_chunks": chunks,
                                            "smoother_region": region}],
                                          check_cycles=checks))
        got = rst.compute_binding_energy_frames(top, fr, rec, lig, pb_params=pb, batched=True)
        for a, b in zip(got.frames(), ref.frames()):
            for k in KEYS:
                assert getattr(a, k) == getattr(b, k), (chunks, checks, region, k)


def test_gb_chunks_change_rounding_only(cfg0):
    top, x, _, _, _ = cfg0
    ref = rst.compute_gb_energy(top, x / 10.0, rst.GbParams(model=2))
    for c in (1, 3, 17):
        autotune.apply(autotune.TuneTable(gb_chunks=c))
        g = rst.compute_gb_energy(top, x / 10.0, rst.GbParams(model=2))
        assert np.max(np.abs(g.born_radii() - ref.born_radii()) / ref.born_radii()) < 1e-12
        assert abs(g.total - ref.total) < 1e-9 * abs(ref.total)


def test_search_end_to_end(cfg0, tmp_path):
    top, fr, rec, lig = _frames(cfg0, 1)
    pb = rst.PbParams(**PB)
    out = {}

    def run():
        out["r"] = rst.compute_binding_energy_frames(top, fr, rec, lig, pb_params=pb, batched=True)

    run()
    ref = out["r"].frames()[0].delta_total
    table, report = autotune.tune(run, max_buckets=2, candidates=(1, 2), repeats=1,
                                  check_cycles=(8, 4))
    assert report["buckets"] and report["buckets"][0]["weight"] > 0
    assert autotune.current().to_dict()["entries"] == table.to_dict()["entries"]
    run()
    assert out["r"].frames()[0].delta_total == ref
    p = autotune.save(table, str(tmp_path / "t.json"), workload="cfg0")
    assert autotune.load(p, workload="cfg0").to_dict() == table.to_dict()
