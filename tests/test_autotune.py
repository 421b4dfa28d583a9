"""The tuning table's plumbing (host only, no device): set / get round trip,
the solver's chunk lookup, JSON persistence, and the bucketing of observed
level shapes.  Tuned launch shapes never change results; the GPU tests
compare tuned and untuned runs bit for bit."""
import pytest

from rust_simulation_tools import _abi, autotune


@pytest.fixture(autouse=True)
def restore():
    yield
    autotune.apply(None)


def test_table_round_trip_and_lookup():
    lib = _abi.load_library()
    base = lib.mmpbsa_tuning_smoother_zc(161, 161, 161, 48)
    base81 = lib.mmpbsa_tuning_smoother_zc(81, 81, 81, 48)
    assert base > 0
    # built-in regions: 44 x 44 on 161^3 (25 tiles of 512 threads per plane
    # against 64 of 256), 32 x 32 on 81^3 (9 x 512 against 16 x 256)
    assert lib.mmpbsa_tuning_smoother_region(161, 161, 161, 48) == 2
    assert lib.mmpbsa_tuning_smoother_region(81, 81, 81, 48) == 1
    t = autotune.TuneTable([{"nodes_min": 1 << 21, "nodes_max": (1 << 22) - 1, "solves_min": 1,
                             "solves_max": 64, "smoother_chunks": 3, "smoother_region": 1}],
                           check_cycles=4, gb_chunks=8)
    autotune.apply(t)
    got = autotune.current()
    assert got.entries == t.entries and got.check_cycles == 4 and got.gb_chunks == 8
    # 161^3 = 4.17e6 nodes: 3 chunks over the 159 interior planes, small region
    assert lib.mmpbsa_tuning_smoother_zc(161, 161, 161, 48) == 53
    assert lib.mmpbsa_tuning_smoother_region(161, 161, 161, 48) == 1
    # outside the entry's solve range, or another size: the built-in choice
    assert lib.mmpbsa_tuning_smoother_zc(161, 161, 161, 96) == base
    assert lib.mmpbsa_tuning_smoother_region(161, 161, 161, 96) == 2
    assert lib.mmpbsa_tuning_smoother_zc(81, 81, 81, 48) == base81
    autotune.apply(None)
    assert autotune.current().entries == [] and autotune.current().check_cycles == 8
    assert lib.mmpbsa_tuning_smoother_zc(161, 161, 161, 48) == base
    with pytest.raises(ValueError):
        autotune.apply(autotune.TuneTable(check_cycles=-1))
    with pytest.raises(ValueError):
        autotune.apply(autotune.TuneTable([{"nodes_min": 0, "nodes_max": 9, "solves_min": 1,
                                            "solves_max": 9, "smoother_region": 3}]))


def test_save_load(tmp_path):
    p = tmp_path / "t.json"
    t = autotune.TuneTable([{"nodes_min": 8, "nodes_max": 15, "solves_min": 2, "solves_max": 9,
                             "smoother_chunks": 2}], 16, 0, round_frames=32)
    autotune.save(t, str(p), workload="w")
    autotune.save(autotune.TuneTable(), str(p), workload="other")
    u = autotune.load(str(p), workload="w")
    assert u.to_dict() == t.to_dict()
    assert autotune.load(str(p), workload="missing") is None


def test_buckets():
    shapes = [(161 ** 3, 48, 176), (81 ** 3, 48, 176), (4_000_000, 12, 10), (41 ** 3, 48, 176)]
    b = autotune.buckets(shapes)
    assert b[0]["nodes_min"] == 1 << 21 and b[0]["nodes_max"] == (1 << 22) - 1
    assert b[0]["solves_min"] == 12 and b[0]["solves_max"] == 48
    assert [x["weight"] for x in b] == sorted([x["weight"] for x in b], reverse=True)
    assert all(x["nodes_min"] <= n <= x["nodes_max"] for n, _, _ in shapes[:1] for x in b[:1])
