"""MM-PBSA frames/s on one MI355X.

Headline workload: the 25,000-atom synthetic complex (23,500 receptor + 1,500
ligand atoms, 45 A sphere), frames = base + N(0, 0.3^2) A noise, PB on a 161^3
grid at 0.5 A (eps_in 4, eps_out 80, 0.15 M, two-level focusing), MM and SA
for the complex, receptor and ligand of every frame.  The frames are placed in
HBM before the timed region; one step = a batch of frames through the batched
path (every PB solve of the batch advances in shared, graph-replayed launches).

Beside it, in the same run:
  - the same steps from host buffers (PCIe included) and the plain path (one
    frame at a time, host-synchronous, no graphs);
  - every term's error against the FP64 oracle (13 frame terms, Born radii,
    the oracle's own convergence) and the oracle timed on the host cores;
  - the smoother's roofline from HIP events in one extra profiled step, with
    hardware traffic from two rocprofv3 counter runs of that step; rooflines
    of the pair kernels and the face-dielectric kernel from HIP events;
  - the same complex by MM-GBSA (GB polar term), batched;
  - the 2,500-atom complex's batched / plain ratio;
  - uneven batching (8 complexes, 3k-60k atoms, mixed vs one at a time);
  - the many-complex workload (64 complexes, 3k-60k atoms, own auto_grid
    grids) through the serving loop;
  - Poisson-arrival serving of complexes of 5k-100k atoms with 1-100 frames;
  - the 150,000-atom complex with focusing 193^3 -> 321^3 and per-residue
    decomposition.
"""
import argparse
import concurrent.futures as cf
import csv
import json
import math
import multiprocessing
import os
import platform
import shutil
import subprocess
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(ROOT, "rust-simulation-tools_amd")
for p in (ROOT, PKG):
    if p not in sys.path:
        sys.path.insert(0, p)

import ctypes as C  # noqa: E402

import numpy as np  # noqa: E402

HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFS = 78.6
FP32_PEAK_TFS = 157.3
CFG1_PB = dict(grid_spacing=0.5, solute_dielectric=4.0, solvent_dielectric=80.0,
               salt_concentration=0.15, fine_dims=(161, 161, 161))
CFG0_PB = dict(grid_spacing=0.6, solute_dielectric=4.0, solvent_dielectric=80.0,
               salt_concentration=0.15, fine_dims=(65, 65, 65))
# many-complex workloads: auto_grid at 0.5 A with a 10 A buffer per subsystem
AUTO_PB = dict(grid_spacing=0.5, grid_buffer=10.0, solute_dielectric=4.0,
               solvent_dielectric=80.0, salt_concentration=0.15)
GB_MODEL = 2  # OBC-II
# energy terms at or above this magnitude (one FP64 ulp >= 1.2e-4 kcal/mol) are
# compared in ulps of the oracle's value as well
LARGE_TERM = 1e12


def _note(msg):
    print(msg, file=sys.stderr, flush=True)


def cpu_share():
    """Host cores this job may use: the affinity mask, capped by
    OMP_NUM_THREADS when the environment sets it (the GPU host allots a share
    of its cores to each GPU and says so there)."""
    try:
        aff = len(os.sched_getaffinity(0))
    except AttributeError:
        aff = os.cpu_count() or 1
    omp = os.environ.get("OMP_NUM_THREADS")
    share = min(aff, int(omp)) if omp and omp.isdigit() and int(omp) > 0 else aff
    return share, aff


def cpu_model():
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


# ---- synthetic complexes, built in worker processes -----------------------------
def _build_one(job):
    """(config, index, n_atoms, n_ligand, n_frames) -> (topology, frames nm,
    receptor residues, ligand residues); runs in a spawned worker."""
    from rust_simulation_tools import synthetic
    config, idx, n, n_lig, nf = job
    seed = synthetic.seed_for(config, idx)
    top, base, rec, lig = synthetic.build_complex(int(n), int(n_lig), None, seed)
    f = np.stack([synthetic.noisy_frame(base, seed, k) for k in range(nf)]) / 10.0
    return top, f, rec, lig


def build_complexes(jobs, workers):
    if workers <= 1 or len(jobs) <= 1:
        return [_build_one(j) for j in jobs]
    ctx = multiprocessing.get_context("spawn")
    with cf.ProcessPoolExecutor(max_workers=min(workers, len(jobs)), mp_context=ctx) as ex:
        return list(ex.map(_build_one, jobs, chunksize=1))


def _lig_atoms(n, frac):
    return max(16, int(round(n * frac / 16.0)) * 16)


def _frame_errors(f, o, ulps=False):
    """|GPU - oracle| per frame term (kcal/mol); with ulps=True also each error
    in units of the last place of the oracle's FP64 value."""
    err, ulp = {}, {}
    for part in ("complex", "receptor", "ligand"):
        for term in ("mm", "polar", "sa"):
            ref = o[part][term]
            err[f"{part}_{term}"] = abs(getattr(f, f"{part}_{term}") - ref)
            ulp[f"{part}_{term}"] = err[f"{part}_{term}"] / math.ulp(ref) if ref else 0.0
    for term in ("delta_mm", "delta_polar", "delta_sa", "delta_total"):
        err[term] = abs(getattr(f, term) - o[term])
        ulp[term] = err[term] / math.ulp(o[term]) if o[term] else 0.0
    return (err, ulp) if ulps else err


def _worst(err, ulp, o):
    """The largest error of a frame's terms: kcal/mol, the term, its oracle value
    and the error in units of that value's last place."""
    k = max(err, key=err.get)
    part, _, term = k.partition("_")
    ref = o[k] if k.startswith("delta") else o[part][term]
    return {"term": k, "err_kcal": err[k], "oracle_value": ref, "err_ulps": ulp[k]}


# ---- legs -------------------------------------------------------------------------
def uneven_leg(rst, synthetic, workers):
    """Eight complexes, N log-spaced over [3k, 60k], two frames each: all in one
    server (their solves share rounds) against each complex batched alone."""
    rng = np.random.Generator(np.random.PCG64(synthetic.seed_for(2, 900)))
    sizes = np.round(np.exp(np.linspace(math.log(3000), math.log(60000), 8))).astype(int)
    jobs = [(2, i, int(n), _lig_atoms(int(n), rng.uniform(0.03, 0.10)), 2)
            for i, n in enumerate(sizes)]
    cx = build_complexes(jobs, workers)
    pb = rst.PbParams(**AUTO_PB)
    for top, f, rec, lig in cx:  # contexts, splits, first-use costs
        rst.compute_binding_energy_frames(top, f[:1], rec, lig, pb_params=pb)
    ts = time.perf_counter()
    for top, f, rec, lig in cx:
        rst.compute_binding_energy_frames(top, f, rec, lig, pb_params=pb, batched=True)
    alone = time.perf_counter() - ts
    srv = rst.Server(hbm_budget_bytes=200 << 30, max_round_frames=16)
    try:
        top, f, rec, lig = cx[-1]  # warm the server's device pool with the largest frame
        srv.wait(srv.submit(top, f[:1], rec, lig, pb_params=pb))
        ts = time.perf_counter()
        tickets = [srv.submit(top, f, rec, lig, pb_params=pb) for top, f, rec, lig in cx]
        for t in tickets:
            srv.wait(t)
        mixed = time.perf_counter() - ts
    finally:
        srv.close()
    nfr = 2 * len(cx)
    return {"workload": "8 complexes, N = " + ", ".join(str(int(n)) for n in sizes) +
                        "; 2 frames each; auto_grid h 0.5 A, buffer 10 A",
            "mixed_frames_per_s": nfr / mixed, "one_complex_at_a_time_frames_per_s": nfr / alone,
            "mixed_over_one_at_a_time": alone / mixed}


def _serve_all(rst, cx, pb, round_frames, frames=None, progress=None):
    """Submit every complex's frames at once to one server; (wall s, request
    latencies ms).  With a progress label, a line per finished request."""
    srv = rst.Server(hbm_budget_bytes=200 << 30, max_round_frames=round_frames)
    try:
        t0 = time.perf_counter()
        tickets = [srv.submit(top, f if frames is None else f[:frames], rec, lig, pb_params=pb)
                   for top, f, rec, lig in cx]
        lat = []
        for k, t in enumerate(tickets):
            _, st = srv.wait(t)
            lat.append(st.total_ms)
            if progress:
                _note(f"{progress}: request {k + 1}/{len(tickets)} done at "
                      f"{time.perf_counter() - t0:.1f} s")
        return time.perf_counter() - t0, lat
    finally:
        srv.close()


def many_complex_leg(rst, synthetic, n_cx, n_frames, workers, tune):
    """n_cx complexes, N log-uniform over [3k, 60k], ligand fraction U(3%, 10%),
    n_frames noisy frames each, each on its own auto_grid grids (0.5 A, 10 A
    buffer): every frame submitted at once to one server under a 200 GB
    budget, which batches frames of different complexes into shared solver
    launches.  Run twice: with the built-in launch choices and the server's
    default round size, then with the autotuner's choices (smoother z-chunks
    per level bucket, check interval, round size), tuned on one frame of the
    eight complexes at the size octiles.  The plain path (one frame at a time)
    is timed on the first frame of four complexes at the size quartiles."""
    from rust_simulation_tools import autotune
    rng = np.random.Generator(np.random.PCG64(synthetic.seed_for(2, 0)))
    sizes = np.round(np.exp(rng.uniform(math.log(3000), math.log(60000), n_cx))).astype(int)
    fracs = rng.uniform(0.03, 0.10, n_cx)
    jobs = [(2, i, int(n), _lig_atoms(int(n), fr), n_frames)
            for i, (n, fr) in enumerate(zip(sizes, fracs))]
    ts = time.perf_counter()
    cx = build_complexes(jobs, workers)
    _note(f"many: {n_cx} complexes built in {time.perf_counter() - ts:.1f} s")
    pb = rst.PbParams(**AUTO_PB)
    for top, f, rec, lig in cx:  # contexts and splits (GB frame: no PB first-use cost)
        rst.compute_binding_energy_frames(top, f[:1], rec, lig)
    order = np.argsort(sizes)
    quart = [int(order[int(q * (n_cx - 1))]) for q in (0.125, 0.375, 0.625, 0.875)]
    plain_t = []
    for k in quart:
        top, f, rec, lig = cx[k]
        rst.compute_binding_energy_frames(top, f[:1], rec, lig, pb_params=pb, batched=False)
        t0 = time.perf_counter()
        rst.compute_binding_energy_frames(top, f[:1], rec, lig, pb_params=pb, batched=False)
        plain_t.append(time.perf_counter() - t0)
    big = int(order[-1])
    _serve_all(rst, [cx[big]], pb, 16, frames=1)  # the server path's first-use costs
    nfr = n_cx * n_frames
    autotune.apply(None)
    default_round = 16
    span0, lat0 = _serve_all(rst, cx, pb, default_round, progress="many built-in")
    _note(f"many: built-in {nfr / span0:.2f} frames/s")
    out = {"workload": f"{n_cx} complexes, N log-uniform [3k, 60k] (min {int(sizes.min())}, "
                       f"median {int(np.median(sizes))}, max {int(sizes.max())}), ligand "
                       f"3-10%, {n_frames} frames each ({nfr} frames), auto_grid h 0.5 A "
                       "buffer 10 A, one server, 200 GB budget",
           "frames": nfr,
           "builtin": {"frames_per_s": nfr / span0, "wall_s": span0, "round_frames": default_round,
                       "request_latency_ms": {"p50": float(np.percentile(lat0, 50)),
                                              "p99": float(np.percentile(lat0, 99))}},
           "plain_sample": {"complexes_atoms": [int(sizes[k]) for k in quart],
                            "seconds_per_frame": plain_t,
                            "frames_per_s": len(plain_t) / sum(plain_t)}}
    if tune:
        octiles = [cx[int(order[int(q * (n_cx - 1))])] for q in np.linspace(0.0625, 0.9375, 8)]
        rounds = {"n": default_round}
        ts = time.perf_counter()
        table, rep = autotune.tune(lambda: _serve_all(rst, octiles, pb, rounds["n"], frames=1),
                                   max_buckets=3, candidates=(1, 2, 3, 4), repeats=1,
                                   check_cycles=(8, 4, 16), log=_note)

        def with_round(r):
            rounds["n"] = r
            _serve_all(rst, octiles, pb, r, frames=1)

        best_round, tried = autotune.choose(with_round, (16, 8, 32, 64), repeats=1, log=_note,
                                            name="round_frames")
        table.round_frames = best_round
        t_tune = time.perf_counter() - ts
        try:
            path = autotune.save(table, workload="many-complex")
        except OSError:
            path = None
        span1, lat1 = _serve_all(rst, cx, pb, best_round, progress="many tuned")
        _note(f"many: tuned {nfr / span1:.2f} frames/s")
        autotune.apply(None)
        out["tuned"] = {"frames_per_s": nfr / span1, "wall_s": span1, "round_frames": best_round,
                        "request_latency_ms": {"p50": float(np.percentile(lat1, 50)),
                                               "p99": float(np.percentile(lat1, 99))},
                        "tuned_over_builtin": span0 / span1,
                        "table": table.to_dict(), "search": rep,
                        "round_tried_ms": {str(k): v * 1e3 for k, v in tried.items()},
                        "tuning_s": t_tune, "persisted_to": path,
                        "tuned_on": "first frame of the 8 complexes at the size octiles"}
    best = out.get("tuned", out["builtin"])
    out["frames_per_s"] = best["frames_per_s"]
    return out


def serving_leg(rst, synthetic, n_req, workers, oracle_max_atoms, check_oracle):
    """Poisson arrivals (50/s) of n_req requests into one server under a 200 GB
    budget; each request is one complex of N log-uniform over [5k, 100k] atoms
    with U{1..100} frames.  Request latency (submission to last frame),
    throughput, and a seeded 5% sample of frames checked against the FP64
    oracle where the oracle is affordable (complexes up to oracle_max_atoms;
    the smallest request's first frame always)."""
    rng = np.random.Generator(np.random.PCG64(synthetic.seed_for(3, 0)))
    sizes = np.round(np.exp(rng.uniform(math.log(5000), math.log(100000), n_req))).astype(int)
    nframes = rng.integers(1, 101, n_req)
    fracs = rng.uniform(0.03, 0.10, n_req)
    arrivals = np.cumsum(rng.exponential(1.0 / 50.0, n_req))
    jobs = [(3, i, int(n), _lig_atoms(int(n), fr), int(nf))
            for i, (n, fr, nf) in enumerate(zip(sizes, fracs, nframes))]
    ts = time.perf_counter()
    cx = build_complexes(jobs, workers)
    _note(f"serving: {n_req} complexes built in {time.perf_counter() - ts:.1f} s")
    pb = rst.PbParams(**AUTO_PB)
    for top, f, rec, lig in cx:  # register topologies (contexts, splits) up front
        rst.compute_binding_energy_frames(top, f[:1], rec, lig)
    srv = rst.Server(hbm_budget_bytes=200 << 30, max_round_frames=16)
    lat, results = [], []
    try:
        big = int(np.argmax(sizes))  # warm the server's device pool
        srv.wait(srv.submit(cx[big][0], cx[big][1][:1], cx[big][2], cx[big][3], pb_params=pb))
        t0 = time.perf_counter()
        tickets = []
        for k, (top, f, rec, lig) in enumerate(cx):
            dt = arrivals[k] - (time.perf_counter() - t0)
            if dt > 0:
                time.sleep(dt)
            tickets.append(srv.submit(top, f, rec, lig, pb_params=pb))
        for k, t in enumerate(tickets):
            r, st = srv.wait(t)
            lat.append(st.total_ms)
            results.append(r)
            _note(f"serving: request {k + 1}/{n_req} done at {time.perf_counter() - t0:.1f} s")
        span = time.perf_counter() - t0
    finally:
        srv.close()
    total = int(nframes.sum())
    out = {"workload": f"{n_req} requests, Poisson 50/s, N log-uniform [5k, 100k] "
                       f"(drawn: {', '.join(str(int(n)) for n in sizes)}), frames U{{1..100}} "
                       f"(drawn: {', '.join(str(int(n)) for n in nframes)}), auto_grid h 0.5 A "
                       "buffer 10 A, HBM budget 200 GB",
           "frames": total, "frames_per_s": total / span, "wall_s": span,
           "latency_ms": {"p50": float(np.percentile(lat, 50)), "p99": float(np.percentile(lat, 99)),
                          "max": float(np.max(lat))}}
    if check_oracle:
        from oracle import oracle as O
        srng = np.random.Generator(np.random.PCG64(synthetic.seed_for(3, 1)))
        pairs = [(k, f) for k in range(n_req) for f in range(int(nframes[k]))]
        sample = [pairs[i] for i in sorted(srng.choice(len(pairs), max(1, len(pairs) // 20),
                                                       replace=False))]
        small = int(np.argmin(sizes))
        checked = [(k, f) for k, f in sample if sizes[k] <= oracle_max_atoms]
        if (small, 0) not in checked:
            checked.insert(0, (small, 0))
        errs, delta_errs, worst, small_errs, big_ulps = [], [], [], [], [0.0]
        for k, f in checked:
            top, fr, rec, lig = cx[k]
            o = O.frame_energy(top, fr[f] * 10.0, rec, lig, method=1, pb=O.pb_params(**AUTO_PB))
            e, u = _frame_errors(results[k].frames()[f], o, ulps=True)
            errs.append(max(e.values()))
            delta_errs.append(max(v for t, v in e.items() if t.startswith("delta")))
            worst.append(_worst(e, u, o))
            for t in e:
                part, _, term = t.partition("_")
                ref = o[t] if t.startswith("delta") else o[part][term]
                if abs(ref) < LARGE_TERM:
                    small_errs.append(e[t])
                else:
                    big_ulps.append(u[t])
            _note(f"serving: oracle check {len(errs)}/{len(checked)} ({int(sizes[k])} atoms, "
                  f"frame {f}): max {errs[-1]:.2e} kcal/mol ({worst[-1]['term']}, "
                  f"{worst[-1]['err_ulps']:.1f} ulp of {worst[-1]['oracle_value']:.6g}), "
                  f"delta terms {delta_errs[-1]:.2e}")
        out["oracle_check"] = {"sample_frames": len(sample), "checked_frames": len(checked),
                               "checked": [[int(sizes[k]), int(f)] for k, f in checked],
                               "rule": f"seeded 5% sample of frames, oracle run on complexes of "
                                       f"at most {oracle_max_atoms} atoms, plus the smallest "
                                       "request's first frame",
                               "max_term_err_kcal": float(max(errs)),
                               "max_delta_term_err_kcal": float(max(delta_errs)),
                               "worst_frame": worst[int(np.argmax(errs))],
                               "max_err_kcal_terms_below_1e12": float(max(small_errs)),
                               "max_err_ulps_terms_above_1e12": float(max(big_ulps)),
                               "terms_above_1e12": len(big_ulps) - 1,
                               "note": "synthetic frames put a few non-bonded pairs within "
                                       "~0.1 A, so some subsystem MM totals reach 1e13-1e15 "
                                       "kcal/mol, where one FP64 ulp is 0.002-0.1 kcal/mol; "
                                       "terms of magnitude 1e12 or more are compared in ulps of "
                                       "the oracle's value (the delta terms cancel the "
                                       "intra-partner pairs)"}
    return out


def _lj_coul(top, xa, ia, xb, ib):
    """FP64 brute force over a x b: sum of A/r^12 - B/r^6 and qa qb / r."""
    qa = np.asarray(top.charges_amber)
    ty = np.asarray(top.atom_type_indices)
    nbi = np.asarray(top.nb_parm_index)
    A, B = np.asarray(top.lj_acoef), np.asarray(top.lj_bcoef)
    d = np.linalg.norm(xa[:, None, :] - xb[None, :, :], axis=2)
    k = nbi[top.n_types * ty[ia][:, None] + ty[ib][None, :]] - 1
    a = np.where(k >= 0, A[np.maximum(k, 0)], 0.0)
    b = np.where(k >= 0, B[np.maximum(k, 0)], 0.0)
    r6 = d ** 6
    return float((a / (r6 * r6) - b / r6).sum()), float((qa[ia][:, None] * qa[ib][None, :] / d).sum())


def quasi_harmonic_leg(rst, synthetic, with_cpu):
    """Quasi-harmonic entropy (entropy.rs:97-217) of the 25,000-atom complex's
    1,500-atom ligand over 200 noisy frames (a 4,500 x 4,500 covariance) on
    the GPU; the oracle's cyclic Jacobi is timed on a 200-atom sample of the
    same ligand (its cost grows as the cube of the coordinates)."""
    top, base, _, lig, seed = synthetic.config_system(1)
    sel = top.build_selection(lig)
    masses = np.random.default_rng(seed).choice([1.008, 12.011, 14.007, 15.999, 32.06], len(sel))
    traj = np.stack([synthetic.noisy_frame(base, seed, f)[sel] for f in range(200)])
    rst.quasi_harmonic_entropy(traj[:, :50], masses[:50])  # first call: solver set-up
    ts = time.perf_counter()
    q = rst.quasi_harmonic_entropy(traj, masses)
    out = {"workload": f"{len(sel)} atoms x 200 frames ({3 * len(sel)} coordinates)",
           "gpu_s": time.perf_counter() - ts, "minus_tds": q.minus_tds}
    if with_cpu:
        from oracle import oracle as O
        n = 200
        ts = time.perf_counter()
        o = O.quasi_harmonic_entropy(traj[:, :n], masses[:n])
        cpu_s = time.perf_counter() - ts
        ts = time.perf_counter()
        g = rst.quasi_harmonic_entropy(traj[:, :n], masses[:n]).minus_tds
        out["sample"] = {"workload": f"{n} atoms x 200 frames", "cpu_oracle_s": cpu_s,
                         "gpu_s": time.perf_counter() - ts, "rel_err": abs(g - o) / abs(o)}
    return out


def large_decomposition_leg(rst, synthetic, n_frames):
    """The 150,000-atom complex (142,500 + 7,500 atoms; 8,907 + 469 residues of
    16 atoms): per frame its MM-PBSA and its per-residue decomposition with PB
    polar terms, on explicit focusing grids shared by the three subsystems
    (193^3 at 1.0 A centred on the complex, then 321^3 at 0.4 A centred on the
    interface), beside the GB decomposition of the first frame.  Errors: the
    size-independent identities, and the vdW / elec terms of 12 sampled ligand
    residues against an FP64 brute-force pair sum."""
    top, base, rec, lig, seed = synthetic.config_system(4)
    _note("large: system built")
    ra, la = top.build_selection(rec), top.build_selection(lig)
    xn0 = synthetic.noisy_frame(base, seed, 0) / 10.0
    x0 = xn0 * 10.0
    kw = dict(solute_dielectric=4.0, solvent_dielectric=80.0, salt_concentration=0.15,
              **synthetic.focusing_grids(x0, ra, la))
    pb = rst.PbParams(**kw)
    ts = time.perf_counter()
    g = rst.decompose_binding_energy(top, xn0, rec, lig)  # upload, split, first use
    first = time.perf_counter() - ts
    ts = time.perf_counter()
    g = rst.decompose_binding_energy(top, xn0, rec, lig)
    t_gb = time.perf_counter() - ts
    t_frame, t_dec, ident = [], [], []
    dpb0 = None
    for k in range(n_frames):
        xn = synthetic.noisy_frame(base, seed, k) / 10.0
        ts = time.perf_counter()
        e = rst.compute_binding_energy_single_frame(top, xn, rec, lig, pb_params=pb)
        t_frame.append(time.perf_counter() - ts)
        ts = time.perf_counter()
        d, dpb = rst.decompose_binding_energy_pb(top, xn, rec, lig, pb_params=pb)
        t_dec.append(time.perf_counter() - ts)
        pol = math.fsum(c.polar for c in d.receptor_residues() + d.ligand_residues())
        ident.append(max(abs(pol - dpb), abs(e.delta_polar - dpb)))
        if k == 0:
            dpb0, e0 = dpb, e
        _note(f"large: frame {k}: MM-PBSA {t_frame[-1]:.2f} s, PB decomposition {t_dec[-1]:.2f} s")
    rng = np.random.default_rng(3)
    worst = 0.0
    for k in rng.choice(len(lig), 12, replace=False):
        res = lig[k]
        ptr = np.append(np.asarray(top.residue_pointers), top.n_atoms)
        ia = np.arange(ptr[res], ptr[res + 1])
        v, el = _lj_coul(top, x0[ia], ia, x0[ra], ra)
        c = g.ligand_residues()[k]
        worst = max(worst, abs(c.vdw - 0.5 * v), abs(c.elec - 0.5 * el))
    return {"workload": f"150,000-atom complex, 9,376 residues; PB 193^3 at 1.0 A then 321^3 at "
                        f"0.4 A on the interface centroid; {n_frames} frames",
            "frame_mmpbsa_s": float(np.mean(t_frame)), "pb_decomposition_s": float(np.mean(t_dec)),
            "gb_decomposition_s": t_gb, "first_call_s": first,
            "delta_total": e0.delta_total, "delta_pb": dpb0,
            "max_identity_err_kcal": float(max(ident)),
            "identities": "sum of residue PB terms = dG_PB = the frame's dPolar",
            "sampled_residue_vdw_elec_err_vs_fp64_kcal": worst,
            "sampled_residues": 12}


def measure_traffic(frames_per_step, grid_threads):
    """HBM bytes per launch of the timed smoother passes from two rocprofv3
    counter runs (FETCH_SIZE, then WRITE_SIZE, KiB per dispatch) of one batched
    step in a child process, each scaled by what the same counter reads for a
    1 GiB plain copy kernel run first in that process.  The timed launches are
    the smoother dispatches of that launch size."""
    prof = shutil.which("rocprofv3")
    if not prof or grid_threads <= 0:
        return None
    copy_mib = 1024
    child = [sys.executable, os.path.abspath(__file__), "--steps", "1", "--warmup", "0",
             "--frames-per-step", str(frames_per_step), "--no-cpu", "--no-extras",
             "--plain-frames", "0", "--no-traffic", "--copy-calibration-mib", str(copy_mib)]
    vals = {}
    with tempfile.TemporaryDirectory(dir="/tmp") as tmp:
        for counter in ("FETCH_SIZE", "WRITE_SIZE"):
            out = os.path.join(tmp, counter)
            cmd = [prof, "--pmc", counter, "--kernel-include-regex", "smooth3?_kernel|copy_kernel",
                   "--output-format", "csv", "-d", out, "-o", "run", "--"] + child
            r = subprocess.run(cmd, stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL,
                               timeout=300, cwd=tmp, env=dict(os.environ, TMPDIR="/tmp"))
            path = os.path.join(out, "run_counter_collection.csv")
            if r.returncode != 0 or not os.path.exists(path):
                return None
            copy, smooth = [], []
            with open(path) as fh:
                for row in csv.DictReader(fh):
                    v = float(row["Counter_Value"])
                    if "copy_kernel" in row["Kernel_Name"]:
                        copy.append(v)
                    elif int(row["Grid_Size"]) == grid_threads:
                        smooth.append(v)
            if not copy or not smooth:
                return None
            scale = (copy_mib * 1024.0) / (sum(copy) / len(copy))  # KiB moved / KiB counted
            vals[counter] = (sum(smooth) / len(smooth), scale, len(smooth))
    fetch = vals["FETCH_SIZE"][0] * vals["FETCH_SIZE"][1] * 1024.0
    write = vals["WRITE_SIZE"][0] * vals["WRITE_SIZE"][1] * 1024.0
    return {"bytes_per_launch": fetch + write, "fetch_bytes": fetch, "write_bytes": write,
            "fetch_scale": vals["FETCH_SIZE"][1], "write_scale": vals["WRITE_SIZE"][1],
            "dispatches": vals["FETCH_SIZE"][2], "copy_calibration_bytes": copy_mib << 20}


def kernel_rooflines(lib, names):
    """Achieved rate of named kernels from the library's HIP-event timers."""
    out = {}
    for name, bound in names.items():
        ms, n, work = C.c_double(), C.c_int64(), C.c_double()
        lib.mmpbsa_profile_kernel(name.encode(), C.byref(ms), C.byref(n), C.byref(work))
        if n.value <= 0 or ms.value <= 0:
            continue
        rate = work.value / (ms.value * 1e-3)
        if bound == "hbm":
            ach, peak, unit = rate / 1e9, HBM_PEAK_GBS, "GB/s"
        else:
            ach, peak, unit = rate / 1e12, FP64_PEAK_TFS if bound == "fp64" else FP32_PEAK_TFS, \
                "TFLOP/s"
        out[name] = {"bound": bound, "achieved": ach, "peak": peak, "unit": unit,
                     "frac": ach / peak, "launches": n.value, "avg_launch_us": ms.value * 1e3 / n.value,
                     "work_per_launch": work.value / n.value}
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=4)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--frames-per-step", type=int, default=8)
    ap.add_argument("--plain-frames", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extras", action="store_true", help="skip every leg beside the headline")
    ap.add_argument("--cpu-threads", type=int, default=0,
                    help="oracle threads (default: this job's CPU share)")
    ap.add_argument("--no-many", action="store_true",
                    help="skip the many-complex legs (uneven batch, 64 complexes, serving)")
    ap.add_argument("--no-large", action="store_true",
                    help="skip the 150,000-atom decomposition leg")
    ap.add_argument("--large-only", action="store_true",
                    help="run only the 150,000-atom decomposition leg and print its JSON")
    ap.add_argument("--many-complexes", type=int, default=64)
    ap.add_argument("--many-frames", type=int, default=2,
                    help="frames per complex of the many-complex leg (the workload names 20)")
    ap.add_argument("--only", choices=("uneven", "many", "serving"),
                    help="run only this many-complex leg after the headline (for profiling)")
    ap.add_argument("--no-tune", action="store_true",
                    help="skip the autotuner runs (many-complex leg, GB chunking)")
    ap.add_argument("--serving-requests", type=int, default=4)
    ap.add_argument("--large-frames", type=int, default=2,
                    help="frames of the 150,000-atom leg (the workload names 50)")
    ap.add_argument("--oracle-max-atoms", type=int, default=12000)
    ap.add_argument("--no-traffic", action="store_true",
                    help="skip the two rocprofv3 counter runs behind roofline.traffic")
    ap.add_argument("--copy-calibration-mib", type=int, default=0,
                    help="first copy a buffer of this size on the GPU (a plain copy kernel of "
                         "known bytes, to scale hardware traffic counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU; --gpus {args.gpus} is not supported "
              "(frames would shard across GPUs with only per-frame results exchanged)",
              file=sys.stderr)
        sys.exit(2)

    import rust_simulation_tools as rst
    from rust_simulation_tools import _abi, synthetic
    share, aff = cpu_share()
    workers = min(16, share)
    lib = _abi.load_library()
    if lib.mmpbsa_init(None) != 0:
        raise SystemExit(f"no usable gfx950 device: {lib.mmpbsa_last_error(None)}")
    if args.copy_calibration_mib > 0:
        # one plain copy kernel of known bytes (read + write) for counter scaling
        nb = args.copy_calibration_mib << 20
        src, dst = C.c_void_p(), C.c_void_p()
        lib.mmpbsa_device_alloc(0, nb, C.byref(src))
        lib.mmpbsa_device_alloc(0, nb, C.byref(dst))
        if lib.mmpbsa_device_copy(dst, src, nb) != 0:
            raise SystemExit("copy calibration failed")
        lib.mmpbsa_device_free(src)
        lib.mmpbsa_device_free(dst)

    if args.large_only:
        print(json.dumps({"large_decomposition": large_decomposition_leg(rst, synthetic,
                                                                         args.large_frames)}))
        return

    top, base, rec, lig, seed = synthetic.config_system(1)
    B = args.frames_per_step
    nfr = (args.warmup + args.steps + 1) * B
    frames_nm = np.stack([synthetic.noisy_frame(base, seed, f) for f in range(nfr)]) / 10.0
    pb = rst.PbParams(**CFG1_PB)

    dev = rst.DeviceArray(frames_nm * 10.0)  # frames resident in HBM, Angstrom [F, N, 3]

    def step(k, batched=True, **kw):
        kw = kw or dict(pb_params=pb)
        return rst.compute_binding_energy_device_frames(top, dev[k * B:(k + 1) * B], rec, lig,
                                                        batched=batched, **kw)

    def step_host(k):
        return rst.compute_binding_energy_frames(top, frames_nm[k * B:(k + 1) * B], rec, lig,
                                                 pb_params=pb, batched=True)

    for k in range(args.warmup):
        step(k)
    lat = []
    first = None
    t0 = time.perf_counter()
    for k in range(args.warmup, args.warmup + args.steps):
        ts = time.perf_counter()
        r = step(k)
        lat.append((time.perf_counter() - ts) * 1e3)
        if first is None:
            first = r
    elapsed = time.perf_counter() - t0
    value = args.steps * B / elapsed
    _note(f"headline: {value:.2f} frames/s")

    roofline = None
    extras = {}
    if not args.no_extras:
        # the same steps from host buffers: PCIe transfers included
        ts = time.perf_counter()
        for k in range(args.warmup, args.warmup + args.steps):
            step_host(k)
        extras["host_buffer_frames_per_s"] = args.steps * B / (time.perf_counter() - ts)
        # dominant kernel: finest-level smoother passes, timed with HIP events;
        # the pair kernels and face maps of the same step with their own events
        lib.mmpbsa_profile_enable(1)
        step(args.warmup + args.steps)
        ms, n, by, gt = C.c_double(), C.c_int64(), C.c_double(), C.c_int64()
        lib.mmpbsa_profile_read(C.byref(ms), C.byref(n), C.byref(by), C.byref(gt))
        kr = kernel_rooflines(lib, {"mm_pair_kernel": "fp64", "face_dielectrics": "hbm",
                                    "dh_boundary_kernel": "fp32"})
        lib.mmpbsa_profile_enable(0)
        # the same frames by MM-GBSA (OBC-II): batched throughput, then one
        # profiled step for the GB pair kernels
        gbp = rst.GbParams(model=GB_MODEL)
        step(0, gb_params=gbp)
        ts = time.perf_counter()
        for k in range(args.warmup, args.warmup + args.steps):
            step(k, gb_params=gbp)
        extras["gbsa_batched_frames_per_s"] = args.steps * B / (time.perf_counter() - ts)
        lib.mmpbsa_profile_enable(1)
        step(0, gb_params=gbp)
        kr.update(kernel_rooflines(lib, {"gb_radii": "fp64", "gb_energy_kernel": "fp64"}))
        lib.mmpbsa_profile_enable(0)
        if not args.no_tune:
            # pair-kernel chunking of the Born radii, tuned on the same GB step
            from rust_simulation_tools import autotune
            gtab, gtried = autotune.tune_gb(lambda: step(0, gb_params=gbp), log=_note)
            ts = time.perf_counter()
            for k in range(args.warmup, args.warmup + args.steps):
                step(k, gb_params=gbp)
            extras["gbsa_tuned"] = {"frames_per_s": args.steps * B / (time.perf_counter() - ts),
                                    "gb_chunks": gtab.gb_chunks, "tried_ms": gtried}
            lib.mmpbsa_profile_enable(1)
            step(0, gb_params=gbp)
            extras["gbsa_tuned"]["gb_radii"] = kernel_rooflines(lib, {"gb_radii": "fp64"}).get(
                "gb_radii")
            lib.mmpbsa_profile_enable(0)
            autotune.apply(None)
        extras["kernel_rooflines"] = kr
        if n.value > 0 and ms.value > 0:
            achieved = by.value / (ms.value * 1e-3) / 1e9
            traffic = None if args.no_traffic else measure_traffic(B, gt.value)
            roofline = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": achieved / HBM_PEAK_GBS,
                        "traffic": traffic["bytes_per_launch"] if traffic else None,
                        "traffic_detail": traffic,
                        "kernel": "smooth3_kernel (finest level: one smoothing phase, three "
                                  "red-black sweeps, per launch, batched over solves)",
                        "launches": n.value, "avg_launch_us": ms.value * 1e3 / n.value,
                        "alg_bytes_per_launch": by.value / n.value}
        # small system: batched vs plain where launches dominate
        t0s, b0, r0, l0, s0 = synthetic.config_system(0)
        f0 = np.stack([synthetic.noisy_frame(b0, s0, f) for f in range(16)]) / 10.0
        p0 = rst.PbParams(**CFG0_PB)
        rst.compute_binding_energy_frames(t0s, f0[:2], r0, l0, pb_params=p0, batched=True)
        ts = time.perf_counter()
        rst.compute_binding_energy_frames(t0s, f0, r0, l0, pb_params=p0, batched=True)
        fb = 16 / (time.perf_counter() - ts)
        ts = time.perf_counter()
        rst.compute_binding_energy_frames(t0s, f0[:4], r0, l0, pb_params=p0, batched=False)
        fp = 4 / (time.perf_counter() - ts)
        extras["small_complex_2500_atoms"] = {"batched_frames_per_s": fb,
                                              "plain_frames_per_s": fp,
                                              "batched_over_plain": fb / fp}
        extras["quasi_harmonic_entropy"] = quasi_harmonic_leg(rst, synthetic, not args.no_cpu)
        _note("quasi-harmonic entropy done")

    if args.only:
        if args.only == "uneven":
            extras["uneven_batch"] = uneven_leg(rst, synthetic, workers)
        elif args.only == "many":
            extras["many_complexes"] = many_complex_leg(rst, synthetic, args.many_complexes,
                                                        args.many_frames, workers,
                                                        not args.no_tune)
        else:
            extras["serving"] = serving_leg(rst, synthetic, args.serving_requests, workers,
                                            args.oracle_max_atoms, not args.no_cpu)
    elif not args.no_extras and not args.no_many:
        extras["uneven_batch"] = uneven_leg(rst, synthetic, workers)
        _note("uneven batch done")
        extras["many_complexes"] = many_complex_leg(rst, synthetic, args.many_complexes,
                                                    args.many_frames, workers, not args.no_tune)
        _note("many-complex leg done")
        extras["serving"] = serving_leg(rst, synthetic, args.serving_requests, workers,
                                        args.oracle_max_atoms, not args.no_cpu)
        _note("serving leg done")
    if not args.no_extras and not args.no_large and not args.only:
        extras["large_decomposition"] = large_decomposition_leg(rst, synthetic, args.large_frames)

    plain_lat = []
    for f in range(args.plain_frames):
        ts = time.perf_counter()
        rst.compute_binding_energy_device_frames(top, dev[f:f + 1], rec, lig, pb_params=pb,
                                                 batched=False)
        plain_lat.append(time.perf_counter() - ts)
    plain_fps = 1.0 / float(np.median(plain_lat)) if plain_lat else None

    cpu = None
    errors = None
    if not args.no_cpu:
        from oracle import oracle as O
        threads = args.cpu_threads or share
        O.lib().orc_set_threads(threads)
        x = frames_nm[args.warmup * B] * 10.0
        ts = time.perf_counter()
        o = O.frame_energy(top, x, rec, lig, method=1, pb=O.pb_params(**CFG1_PB))
        cpu_s = time.perf_counter() - ts
        g = first.frames()[0]
        terms, term_ulps = _frame_errors(g, o, ulps=True)
        # Born radii of the complex (OBC-II) and the oracle's own PB convergence
        ge = rst.compute_gb_energy(top, x / 10.0, rst.GbParams(model=GB_MODEL))
        og, oR = O.gb_energy(top, x, O.gb_params(model=GB_MODEL))
        gp = rst.compute_pb_energy(top, x / 10.0, pb)
        oc = o["complex"]["pb"]
        pbinfo = {}
        for part in ("complex", "receptor", "ligand"):
            r_ = o[part]["pb"]
            pbinfo[part] = {"oracle_converged": [bool(r_.solv.converged), bool(r_.ref.converged)],
                            "oracle_final_residual": [r_.solv.final_residual,
                                                      r_.ref.final_residual],
                            "oracle_cycles": [r_.solv.iterations, r_.ref.iterations]}
        pbinfo["complex"]["gpu_cycles"] = list(gp.iterations)
        pbinfo["complex"]["gpu_final_residual"] = list(gp.final_residual)
        errors = {"frame_terms_kcal": terms, "max_term_kcal": max(terms.values()),
                  "frame_terms_ulps": term_ulps, "worst_term": _worst(terms, term_ulps, o),
                  "born_radii_max_rel": float(np.max(np.abs(ge.born_radii() - oR) / oR)),
                  "gb_complex_kcal": abs(ge.total - og),
                  "pb_complex_dG_kcal": abs(gp.total - oc.total),
                  "pb_solves": pbinfo,
                  "oracle_vs_exact": "no scipy or long-double pin at this size; the oracle's "
                                     "tol-1e-6 answer is within 6e-4 kcal/mol of its tol-1e-10 "
                                     "answer on this grid"}
        cpu = {"value": 1.0 / cpu_s, "unit": "frames/s", "cores": threads, "kind": "port",
               "nproc": os.cpu_count(), "affinity_cpus": aff, "threads": threads,
               "cpu_model": cpu_model(),
               "sample": "1 frame of the 25,000-atom complex (MM + PB + SA for complex, "
                         "receptor, ligand) on the FP64 C oracle, OpenMP over this job's CPU "
                         "share (affinity mask capped by OMP_NUM_THREADS)"}

    out = {
        "metric": "MM-PBSA frames/s (batched) at dG err <= 0.01 kcal/mol",
        "value": value, "unit": "frames/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": elapsed / args.steps * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "fp32 smoother + fp64 residual/energies", "data": "synthetic",
        "config": {"workload": "25,000-atom complex (23,500 + 1,500), 161^3 PB grid at 0.5 A, "
                               "focusing, MM + PB + SA per subsystem",
                   "frames_per_step": B},
        "plain_frames_per_s": plain_fps,
        "batched_over_plain": (value / plain_fps) if plain_fps else None,
        "latency_ms": {"p50": float(np.percentile(lat, 50)), "p99": float(np.percentile(lat, 99)),
                       "of": f"one step of {B} frames"},
        "dG_err_vs_fp64_oracle": errors["frame_terms_kcal"]["delta_total"] if errors else None,
        "errors_vs_fp64_oracle": errors,
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    if cpu:
        out["batched_over_cpu"] = value / cpu["value"]
    out.update(extras)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
