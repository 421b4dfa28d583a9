"""MM-PBSA frames/s on one MI355X.

Workload: the 25,000-atom synthetic complex (23,500 receptor + 1,500 ligand
atoms, 45 A sphere), frames = base + N(0, 0.3^2) A noise, PB on a 161^3 grid
at 0.5 A (eps_in 4, eps_out 80, 0.15 M, two-level focusing), MM and SA for the
complex, receptor and ligand of every frame.  The frames are placed in HBM
before the timed region; one step = a batch of frames through the batched
path (every PB solve of the batch advances in shared, graph-replayed
launches).  Beside it: the same steps from host buffers (PCIe included), the
plain path (one frame at a time, host-synchronous, no graphs), the smoother's
roofline from HIP events in one extra profiled step of the same workload, the
2,500-atom complex's batched / plain ratio, and the FP64 CPU oracle on the
host cores.
"""
import argparse
import csv
import json
import math
import os
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
CFG1_PB = dict(grid_spacing=0.5, solute_dielectric=4.0, solvent_dielectric=80.0,
               salt_concentration=0.15, fine_dims=(161, 161, 161))
CFG0_PB = dict(grid_spacing=0.6, solute_dielectric=4.0, solvent_dielectric=80.0,
               salt_concentration=0.15, fine_dims=(65, 65, 65))
# many-complex workloads: auto_grid at 0.5 A with a 10 A buffer per subsystem
AUTO_PB = dict(grid_spacing=0.5, grid_buffer=10.0, solute_dielectric=4.0,
               solvent_dielectric=80.0, salt_concentration=0.15)


def _note(msg):
    print(msg, file=sys.stderr, flush=True)


def _complexes(rst, synthetic, config, sizes, frames, rng):
    """(topology, frames nm, receptor residues, ligand residues) per size, with
    a ligand fraction drawn from U(3%, 10%)."""
    out = []
    for i, (n, nf) in enumerate(zip(sizes, frames)):
        seed = synthetic.seed_for(config, i)
        n_lig = max(16, int(round(n * rng.uniform(0.03, 0.10) / 16.0)) * 16)
        top, base, rec, lig = synthetic.build_complex(int(n), n_lig, None, seed)
        f = np.stack([synthetic.noisy_frame(base, seed, k) for k in range(nf)]) / 10.0
        out.append((top, f, rec, lig))
    return out


def uneven_leg(rst, synthetic):
    """Eight complexes, N log-spaced over [3k, 60k], two frames each: all in one
    server (their solves share rounds) against each complex batched alone."""
    rng = np.random.Generator(np.random.PCG64(synthetic.seed_for(2, 900)))
    sizes = np.round(np.exp(np.linspace(math.log(3000), math.log(60000), 8))).astype(int)
    cx = _complexes(rst, synthetic, 2, sizes, [2] * 8, rng)
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
            "mixed_frames_per_s": nfr / mixed, "one_complex_at_a_time_frames_per_s": nfr / alone}


def serving_leg(rst, synthetic, check_oracle):
    """Poisson arrivals (50/s) of 24 requests, N log-uniform over [5k, 25k],
    1-8 frames each, into one server under a 200 GB budget: request latency
    (submission to last frame) and throughput; one sampled frame is checked
    against the FP64 oracle."""
    rng = np.random.Generator(np.random.PCG64(synthetic.seed_for(3, 900)))
    n_req = 24
    sizes = np.round(np.exp(rng.uniform(math.log(5000), math.log(25000), n_req))).astype(int)
    nframes = rng.integers(1, 9, n_req)
    arrivals = np.cumsum(rng.exponential(1.0 / 50.0, n_req))
    cx = _complexes(rst, synthetic, 3, sizes, nframes, rng)
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
        for t in tickets:
            r, st = srv.wait(t)
            lat.append(st.total_ms)
            results.append(r)
        span = time.perf_counter() - t0
    finally:
        srv.close()
    out = {"workload": f"{n_req} requests, Poisson 50/s, N log-uniform [5k, 25k], 1-8 frames, "
                       "auto_grid h 0.5 A buffer 10 A, HBM budget 200 GB",
           "frames": int(nframes.sum()), "frames_per_s": float(nframes.sum()) / span,
           "latency_ms": {"p50": float(np.percentile(lat, 50)), "p99": float(np.percentile(lat, 99)),
                          "max": float(np.max(lat))}}
    if check_oracle:
        from oracle import oracle as O
        k = int(np.argmin(sizes))  # a sampled request: the smallest system, its first frame
        top, f, rec, lig = cx[k]
        o = O.frame_energy(top, f[0] * 10.0, rec, lig, method=1, pb=O.pb_params(**AUTO_PB))
        out["sampled_dG_err_vs_fp64_oracle"] = abs(results[k].frames()[0].delta_total -
                                                   o["delta_total"])
    return out


def large_decomposition_leg(rst, synthetic):
    """The 150,000-atom complex (142,500 + 7,500 atoms; 8,907 + 469 residues of
    16 atoms): one frame's MM-PBSA and its per-residue decomposition with PB
    polar terms, on explicit focusing grids shared by the three subsystems
    (193^3 at 1.0 A centred on the complex, then 321^3 at 0.4 A centred on the
    interface), beside the GB decomposition of the same frame."""
    top, base, rec, lig, seed = synthetic.config_system(4)
    x = synthetic.noisy_frame(base, seed, 0)
    _note("large: system built")
    ra, la = top.build_selection(rec), top.build_selection(lig)
    kw = dict(solute_dielectric=4.0, solvent_dielectric=80.0, salt_concentration=0.15,
              **synthetic.focusing_grids(x, ra, la))
    pb = rst.PbParams(**kw)
    xn = x / 10.0
    ts = time.perf_counter()
    rst.decompose_binding_energy(top, xn, rec, lig)  # topology upload, split, first use
    first = time.perf_counter() - ts
    _note(f"large: first GB decomposition {first:.2f} s")
    ts = time.perf_counter()
    rst.decompose_binding_energy(top, xn, rec, lig)
    t_gb = time.perf_counter() - ts
    ts = time.perf_counter()
    e = rst.compute_binding_energy_single_frame(top, xn, rec, lig, pb_params=pb)
    t_frame = time.perf_counter() - ts
    _note(f"large: PB frame {t_frame:.2f} s")
    ts = time.perf_counter()
    d, dpb = rst.decompose_binding_energy_pb(top, xn, rec, lig, pb_params=pb)
    t_dec = time.perf_counter() - ts
    _note(f"large: PB decomposition {t_dec:.2f} s")
    pol = sum(c.polar for c in d.receptor_residues() + d.ligand_residues())
    return {"workload": "150,000-atom complex, 9,376 residues; PB 193^3 at 1.0 A then 321^3 "
                        "at 0.4 A on the interface centroid; 1 frame",
            "frame_mmpbsa_s": t_frame, "pb_decomposition_s": t_dec, "gb_decomposition_s": t_gb,
            "first_call_s": first, "delta_total": e.delta_total, "delta_pb": dpb,
            "residue_sum_minus_delta_pb": pol - dpb,
            "frame_delta_polar_minus_decomposition_delta_pb": e.delta_polar - dpb}


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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=4)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--frames-per-step", type=int, default=8)
    ap.add_argument("--plain-frames", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extras", action="store_true", help="skip roofline and small-system legs")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--no-many", action="store_true",
                    help="skip the many-complex legs (uneven batch, serving)")
    ap.add_argument("--no-large", action="store_true",
                    help="skip the 150,000-atom decomposition leg")
    ap.add_argument("--large-only", action="store_true",
                    help="run only the 150,000-atom decomposition leg and print its JSON")
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
        print(json.dumps({"large_decomposition": large_decomposition_leg(rst, synthetic)}))
        return

    top, base, rec, lig, seed = synthetic.config_system(1)
    B = args.frames_per_step
    nfr = (args.warmup + args.steps + 1) * B
    frames_nm = np.stack([synthetic.noisy_frame(base, seed, f) for f in range(nfr)]) / 10.0
    pb = rst.PbParams(**CFG1_PB)

    dev = rst.DeviceArray(frames_nm * 10.0)  # frames resident in HBM, Angstrom [F, N, 3]

    def step(k, batched=True):
        return rst.compute_binding_energy_device_frames(top, dev[k * B:(k + 1) * B], rec, lig,
                                                        pb_params=pb, batched=batched)

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

    roofline = None
    extras = {}
    if not args.no_extras:
        # the same steps from host buffers: PCIe transfers included
        ts = time.perf_counter()
        for k in range(args.warmup, args.warmup + args.steps):
            step_host(k)
        extras["host_buffer_frames_per_s"] = args.steps * B / (time.perf_counter() - ts)
        # dominant kernel: finest-level smoother passes, timed with HIP events
        lib.mmpbsa_profile_enable(1)
        step(args.warmup + args.steps)
        ms, n, by, gt = C.c_double(), C.c_int64(), C.c_double(), C.c_int64()
        lib.mmpbsa_profile_read(C.byref(ms), C.byref(n), C.byref(by), C.byref(gt))
        lib.mmpbsa_profile_enable(0)
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

    if not args.no_extras and not args.no_many:
        extras["uneven_batch"] = uneven_leg(rst, synthetic)
        extras["serving"] = serving_leg(rst, synthetic, not args.no_cpu)
    if not args.no_extras and not args.no_large:
        extras["large_decomposition"] = large_decomposition_leg(rst, synthetic)

    plain_lat = []
    for f in range(args.plain_frames):
        ts = time.perf_counter()
        rst.compute_binding_energy_device_frames(top, dev[f:f + 1], rec, lig, pb_params=pb,
                                                 batched=False)
        plain_lat.append(time.perf_counter() - ts)
    plain_fps = 1.0 / float(np.median(plain_lat)) if plain_lat else None

    cpu = None
    dg_err = None
    if not args.no_cpu:
        from oracle import oracle as O
        threads = args.cpu_threads or min(16, os.cpu_count() or 1)
        O.lib().orc_set_threads(threads)
        x = frames_nm[args.warmup * B] * 10.0
        ts = time.perf_counter()
        o = O.frame_energy(top, x, rec, lig, method=1, pb=O.pb_params(**CFG1_PB))
        cpu_s = time.perf_counter() - ts
        g = first.frames()[0]
        dg_err = abs(g.delta_total - o["delta_total"])
        cpu = {"value": 1.0 / cpu_s, "unit": "frames/s", "cores": threads, "kind": "port",
               "sample": "1 frame of the 25,000-atom complex (MM + PB + SA for complex, "
                         "receptor, ligand) on the FP64 C oracle, OpenMP"}

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
        "dG_err_vs_fp64_oracle": dg_err,
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    out.update(extras)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
