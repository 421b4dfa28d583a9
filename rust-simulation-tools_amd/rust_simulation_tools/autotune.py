"""Autotuner for the batched solver's launch shapes (no counterpart in the
reference, which fixes 8^3 SOR tiles and its thread chunking:
pb_solver.rs:214-327, pb_grid.rs:233).

A table maps (multigrid level size, solves sharing the launch) to the phase
smoother's workgroup region (32 x 32 or 44 x 44 nodes) and z-chunk count,
together with the host check interval, the Born-radius kernel's j-chunking
and the serving loop's round size.  None of the smoother choices changes a
result bit (the smoother, the streaming restriction and prolongation are
exact for any tiling and z-chunking), so a tuned run returns the energies of
an untuned one.

Search: the workload is run once with observation on, which records the
level shapes the solver lays out; shapes are grouped into size buckets
(powers of two of the node count) weighted by node count x occurrences;
then, heaviest bucket first, each region and then each chunk count is timed
on the workload and kept only when it beats the incumbent by more than the
noise margin.  Tables persist as JSON keyed by device.
"""
from __future__ import annotations

import ctypes as C
import json
import math
import os
import time

from . import _abi
from ._abi import TuneEntryC, TuningC

DEFAULT_CHECK_CYCLES = 8


def _lib():
    return _abi.load_library()


class TuneTable:
    """entries: [{"nodes_min", "nodes_max", "solves_min", "solves_max",
    "smoother_chunks" (0 = built-in), "smoother_region" (0 = built-in, 1 = 32 x
    32, 2 = 44 x 44)}], check_cycles, gb_chunks (0 = built-in), round_frames
    (the serving loop's round size; None = its default)."""

    def __init__(self, entries=None, check_cycles=DEFAULT_CHECK_CYCLES, gb_chunks=0,
                 round_frames=None):
        self.entries = list(entries or [])
        self.check_cycles = int(check_cycles)
        self.gb_chunks = int(gb_chunks)
        self.round_frames = round_frames

    def to_dict(self):
        return {"entries": self.entries, "check_cycles": self.check_cycles,
                "gb_chunks": self.gb_chunks, "round_frames": self.round_frames}

    @classmethod
    def from_dict(cls, d):
        return cls(d.get("entries", []), d.get("check_cycles", DEFAULT_CHECK_CYCLES),
                   d.get("gb_chunks", 0), d.get("round_frames"))

    def copy(self):
        return TuneTable.from_dict(json.loads(json.dumps(self.to_dict())))


def apply(table: TuneTable | None):
    """Install a table in the library (None restores the built-in choices)."""
    lib = _lib()
    if table is None:
        rc = lib.mmpbsa_tuning_set(None)
    else:
        arr = (TuneEntryC * max(1, len(table.entries)))()
        for k, e in enumerate(table.entries):
            arr[k] = TuneEntryC(int(e["nodes_min"]), int(e["nodes_max"]), int(e["solves_min"]),
                                int(e["solves_max"]), int(e.get("smoother_chunks", 0)),
                                int(e.get("smoother_region", 0)))
        t = TuningC(table.check_cycles, table.gb_chunks, len(table.entries), arr)
        rc = lib.mmpbsa_tuning_set(C.byref(t))
    if rc != 0:
        raise ValueError("invalid tuning table")


def current() -> TuneTable:
    lib = _lib()
    t = TuningC()
    lib.mmpbsa_tuning_get(C.byref(t), None, 0)
    n = t.n_entries
    arr = (TuneEntryC * max(1, n))()
    lib.mmpbsa_tuning_get(C.byref(t), arr, n)
    entries = [{"nodes_min": arr[k].nodes_min, "nodes_max": arr[k].nodes_max,
                "solves_min": arr[k].solves_min, "solves_max": arr[k].solves_max,
                "smoother_chunks": arr[k].smoother_chunks,
                "smoother_region": arr[k].smoother_region} for k in range(n)]
    return TuneTable(entries, t.check_cycles, t.gb_chunks)


def default_path():
    return os.environ.get("MMPBSA_TUNE_FILE") or os.path.join(
        os.path.expanduser("~"), ".cache", "mmpbsa-hip", "tuning.json")


def save(table: TuneTable, path=None, device="gfx950", workload=None):
    path = path or default_path()
    os.makedirs(os.path.dirname(path) or ".", exist_ok=True)
    data = {}
    if os.path.exists(path):
        with open(path) as fh:
            data = json.load(fh)
    data.setdefault(device, {})[workload or "default"] = table.to_dict()
    with open(path, "w") as fh:
        json.dump(data, fh, indent=1)
    return path


def load(path=None, device="gfx950", workload=None) -> TuneTable | None:
    path = path or default_path()
    if not os.path.exists(path):
        return None
    with open(path) as fh:
        data = json.load(fh)
    d = data.get(device, {}).get(workload or "default")
    return TuneTable.from_dict(d) if d else None


def observe(run):
    """Run the workload once with observation on: [(level nodes, solves in the
    group, occurrences)] of the batched smoother levels it laid out."""
    lib = _lib()
    lib.mmpbsa_tuning_observe(1)
    try:
        run()
        n = lib.mmpbsa_tuning_observed(None, None, None, 0)
        nodes = (C.c_int64 * max(1, n))()
        solves = (C.c_int32 * max(1, n))()
        cnt = (C.c_int64 * max(1, n))()
        lib.mmpbsa_tuning_observed(nodes, solves, cnt, n)
    finally:
        lib.mmpbsa_tuning_observe(0)
    return [(int(nodes[k]), int(solves[k]), int(cnt[k])) for k in range(n)]


def buckets(shapes):
    """Group observed shapes into node-count buckets [2^b, 2^(b+1)), each with
    the solve range it was seen with and its weight (nodes x occurrences),
    heaviest first."""
    agg = {}
    for nodes, solves, cnt in shapes:
        b = int(math.floor(math.log2(max(1, nodes))))
        w, lo, hi = agg.get(b, (0, 1 << 30, 0))
        agg[b] = (w + nodes * cnt, min(lo, solves), max(hi, solves))
    out = [{"nodes_min": 1 << b, "nodes_max": (1 << (b + 1)) - 1, "solves_min": lo,
            "solves_max": hi, "weight": w} for b, (w, lo, hi) in agg.items()]
    return sorted(out, key=lambda e: -e["weight"])


def _time(run, repeats):
    best = float("inf")
    for _ in range(repeats):
        t0 = time.perf_counter()
        run()
        best = min(best, time.perf_counter() - t0)
    return best


def tune(run, max_buckets=4, candidates=(1, 2, 3, 4, 6), regions=(1, 2), repeats=2,
         margin=0.01, check_cycles=(8, 4, 16), base: TuneTable | None = None, log=None):
    """Tune the smoother region and chunk count of the heaviest level buckets
    the workload `run` (a callable) exercises, then the check interval.
    Returns (table, report).  The table is left applied."""
    table = (base or TuneTable()).copy()
    apply(table)
    run()  # first-use costs
    shapes = observe(run)
    bks = buckets(shapes)[:max_buckets]
    t_default = _time(run, repeats)
    best_t = t_default
    report = {"default_s": t_default, "buckets": []}
    for bk in bks:
        entry = {k: bk[k] for k in ("nodes_min", "nodes_max", "solves_min", "solves_max")}

        def with_entry(**kw):
            trial = table.copy()
            old = [e for e in trial.entries if e["nodes_min"] == entry["nodes_min"]]
            cur = dict(old[0]) if old else dict(entry, smoother_chunks=0, smoother_region=0)
            cur.update(kw)
            trial.entries = [e for e in trial.entries if e["nodes_min"] != entry["nodes_min"]]
            trial.entries.append(cur)
            return trial

        picked, tried = {}, {}
        for key, values in (("smoother_region", regions), ("smoother_chunks", candidates)):
            for v in values:
                trial = with_entry(**{key: v})
                apply(trial)
                t = _time(run, repeats)
                tried[f"{key.split('_')[1]} {v}"] = t
                if log:
                    log(f"autotune: nodes [{entry['nodes_min']}, {entry['nodes_max']}] "
                        f"{key} {v}: {t * 1e3:.1f} ms")
                if t < best_t * (1.0 - margin):
                    best_t, table, picked[key] = t, trial, v
            apply(table)
        report["buckets"].append(dict(entry, weight=bk["weight"],
                                      region=picked.get("smoother_region", "built-in"),
                                      chunks=picked.get("smoother_chunks", "built-in"),
                                      tried_ms={k: v * 1e3 for k, v in tried.items()}))
    tried = {}
    for cc in check_cycles:
        trial = table.copy()
        trial.check_cycles = cc
        apply(trial)
        t = _time(run, repeats)
        tried[cc] = t
        if cc != table.check_cycles and t < best_t * (1.0 - margin):
            best_t, table = t, trial
    apply(table)
    report["check_cycles"] = {"chosen": table.check_cycles,
                              "tried_ms": {str(k): v * 1e3 for k, v in tried.items()}}
    report["tuned_s"] = best_t
    return table, report


def choose(run_for, candidates, repeats=2, margin=0.01, log=None, name="value"):
    """Time run_for(c) for each candidate c (the first is the incumbent) and
    return (best candidate, {c: seconds})."""
    tried = {}
    best_c, best_t = None, float("inf")
    for c in candidates:
        t = _time(lambda: run_for(c), repeats)
        tried[c] = t
        if log:
            log(f"autotune: {name} {c}: {t * 1e3:.1f} ms")
        if best_c is None or t < best_t * (1.0 - margin):
            best_c, best_t = c, t
    return best_c, tried


def tune_gb(run, candidates=(0, 2, 4, 8, 16, 32), repeats=3, table: TuneTable | None = None,
            log=None):
    """Choose the Born-radius kernel's j-chunk count (0 = built-in) for the GB
    workload `run`; returns (table, {chunks: ms})."""
    table = (table or current()).copy()

    def with_chunks(c):
        t = table.copy()
        t.gb_chunks = c
        apply(t)
        run()

    run()
    best, tried = choose(with_chunks, candidates, repeats, log=log, name="gb_chunks")
    table.gb_chunks = best
    apply(table)
    return table, {str(k): v * 1e3 for k, v in tried.items()}
