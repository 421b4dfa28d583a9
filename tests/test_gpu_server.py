"""Serving loop (continuous batching under an HBM budget): a request's
results are identical alone and mixed with other complexes and methods, the
budget is enforced at admission, and every ticket completes."""
import time

import numpy as np
import pytest

import rust_simulation_tools as rst
from rust_simulation_tools import synthetic

pytestmark = pytest.mark.gpu

PB_SMALL = dict(grid_spacing=0.8, grid_buffer=4.0, solute_dielectric=4.0,
                solvent_dielectric=80.0, salt_concentration=0.15)
KEYS = ("complex_mm", "complex_polar", "receptor_polar", "ligand_polar", "complex_sa",
        "receptor_sa", "ligand_sa", "delta_total")


def _system(n, n_lig, idx, frames):
    seed = synthetic.seed_for(3, idx)
    top, base, rec, lig = synthetic.build_complex(n, n_lig, None, seed)
    f = np.stack([synthetic.noisy_frame(base, seed, k) for k in range(frames)]) / 10.0
    return top, f, rec, lig


@pytest.fixture(scope="module")
def systems():
    return [_system(900, 80, 0, 3), _system(1400, 120, 1, 2), _system(600, 60, 2, 2)]


def _same(a, b):
    for fa, fb in zip(a.frames(), b.frames()):
        for k in KEYS:
            assert getattr(fa, k) == getattr(fb, k), k


def test_alone_vs_mix_identity(systems):
    (ta, fa, ra, la), (tb, fb, rb, lb), (tc, fc, rc, lc) = systems
    pb = rst.PbParams(**PB_SMALL)
    srv = rst.Server(max_round_frames=16)
    try:
        alone, st = srv.wait(srv.submit(ta, fa, ra, la, pb_params=pb))
        assert st.n_frames == 3 and st.total_ms >= st.queue_ms >= 0.0
        # the same request among a PB request of another complex and a GB one
        t1 = srv.submit(tb, fb, rb, lb, pb_params=pb)
        t2 = srv.submit(ta, fa, ra, la, pb_params=pb)
        t3 = srv.submit(tc, fc, rc, lc)
        mixed, _ = srv.wait(t2)
        _same(alone, mixed)
        # and both equal the direct batched path
        direct = rst.compute_binding_energy_frames(ta, fa, ra, la, pb_params=pb, batched=True)
        _same(alone, direct)
        other, _ = srv.wait(t1)
        _same(other, rst.compute_binding_energy_frames(tb, fb, rb, lb, pb_params=pb))
        gb, _ = srv.wait(t3)
        _same(gb, rst.compute_binding_energy_frames(tc, fc, rc, lc))
    finally:
        srv.close()


def test_direct_calls_while_server_captures(systems):
    """Direct calls that create and destroy device state while the server
    thread captures its cycle graphs: every direct call uses a topology built
    afresh (context and split creation: stream, pinned buffers, synchronous
    uploads; destruction: frees), and device arrays are allocated, filled,
    read back and freed in between.  HIP rejects synchronising calls while a
    capture is open; the library's capture gate keeps them apart, so neither
    side fails and the results stay equal."""
    import gc
    (ta, fa, ra, la), (tb, fb, rb, lb), _ = systems
    pb = rst.PbParams(**PB_SMALL)
    ref = rst.compute_binding_energy_frames(tb, fb, rb, lb, pb_params=pb, batched=True)
    srv = rst.Server(max_round_frames=16)
    try:
        tickets = [srv.submit(ta, fa, ra, la, pb_params=pb) for _ in range(6)]
        direct = []
        for k in range(4):
            fresh, _, rr, ll = synthetic.build_complex(1400, 120, None, synthetic.seed_for(3, 1))
            direct.append(rst.compute_binding_energy_frames(fresh, fb, rr, ll, pb_params=pb,
                                                            batched=True))
            del fresh            # context and split destroyed while tickets are pending
            gc.collect()
            arr = rst.DeviceArray(fb * (k + 1.0))
            assert np.array_equal(arr.numpy(), fb * (k + 1.0))
            del arr
        served = [srv.wait(t)[0] for t in tickets]
        for r in direct:
            _same(ref, r)
        for r in served[1:]:
            _same(served[0], r)
    finally:
        srv.close()


def test_budget_admission(systems):
    ta, fa, ra, la = systems[0]
    pb = rst.PbParams(**PB_SMALL)
    tiny = rst.Server(hbm_budget_bytes=1024)
    try:
        with pytest.raises(RuntimeError, match="budget"):
            tiny.submit(ta, fa, ra, la, pb_params=pb)
    finally:
        tiny.close()
    # a budget of a few frames (each ~0.1 GB by the admission estimate) still
    # completes every request, with the same results
    one = rst.Server(hbm_budget_bytes=250 << 20)
    ref = rst.compute_binding_energy_frames(ta, fa, ra, la, pb_params=pb)
    try:
        tickets = [one.submit(ta, fa, ra, la, pb_params=pb) for _ in range(2)]
        for t in tickets:
            res = None
            while res is None:
                res = one.poll(t)
                if res is None:
                    time.sleep(0.01)
            _same(res[0], ref)
    finally:
        one.close()


def _close(a, b, rel=1e-12):
    assert abs(a - b) <= rel * max(1.0, abs(a), abs(b)), (a, b)


def test_request_summary_on_device(systems):
    """Each finished request's statistics (binding.rs:704-743) and interaction
    entropy (entropy.rs:53-78) come from the device epilogue: equal, to
    rounding, to the host functions over the frames the ticket returns; a
    single frame has std 0 and no entropy.  The temperature is the request's
    own (PB parameters for PB, GB parameters otherwise)."""
    (ta, fa, ra, la), (tb, fb, rb, lb), (tc, fc, rc, lc) = systems
    srv = rst.Server(max_round_frames=4)
    try:
        gb = rst.GbParams(model=2, temperature=310.0)
        many = np.concatenate([fc] * 4)  # 8 frames over two rounds
        t1 = srv.submit(tc, many, rc, lc, gb_params=gb)
        t2 = srv.submit(ta, fa[:1], ra, la)
        t3 = srv.submit(tb, fb, rb, lb, pb_params=rst.PbParams(**PB_SMALL, temperature=290.0))
        for t, temp, n in ((t1, 310.0, 8), (t2, 300.0, 1), (t3, 290.0, 2)):
            res, st, sm = srv.wait_summary(t)
            assert st.n_frames == n and sm.temperature == temp
            for k in ("mean_delta_mm", "mean_delta_polar", "mean_delta_sa", "mean_delta_total",
                      "std_delta_total", "std_delta_mm", "std_delta_polar", "std_delta_sa",
                      "sem_delta_total"):
                _close(getattr(sm, k), getattr(res, k))
            if n == 1:
                assert sm.std_delta_total == 0.0 and sm.entropy is None
            else:
                ie = rst.interaction_entropy(res.frames(), temperature=temp)
                _close(sm.entropy.minus_tds, ie.minus_tds)
        # poll_summary on a finished request, and NOT_READY handling
        t4 = srv.submit(tc, fc, rc, lc)
        got = None
        for _ in range(2000):
            got = srv.poll_summary(t4)
            if got is not None:
                break
            time.sleep(0.01)
        assert got is not None and got[2].temperature == 300.0
    finally:
        srv.close()
