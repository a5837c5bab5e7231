"""Benchmark: vector pairs graded per second on the 50k-gate netlist (one GPU).

A step is one full grading pass (sim -> mark -> merge -> cover over every
chunk, the reference's chunk loop main.cu:171-231) of 65,536 pairs on the
seeded 50,000-gate / 60-level netlist, with the netlist and vectors already
resident in HBM.  Prints one JSON line.

The second half of the metric, time to 95% of full path coverage with the
learned pair ranking against plain in-order grading (both to the same
fraction of the paths, in the same chunks), comes from the held-out
experiment of pdfgrade.learned (32 training and 8 held-out 50k-gate circuits,
1M candidate pairs each), run after the timed steps in both marking modes
(--learned 0 skips it).

`also` holds the same workload in the other marking mode, the cover-heavy
50k-gate circuit (many more marked paths: the merge and the cover DP under
load) and the 256-circuit batch against per-circuit grading (--extra 0 /
--batch 0 skip them).  --workload cover-heavy-50k-gates or
batch-256-circuits makes one of them the main line.

The roofline object is for the dominant kernel (the level walk with the
recorded tuning): algorithmic bytes per launch over its average duration
timed live with HIP events on the grader's stream (frac), over the average of
the committed rocprofv3 trace (frac_rocprof), the committed FETCH_SIZE +
WRITE_SIZE traffic over the same time (traffic_frac), and the bytes the
walk's own design must move (model_frac).  The CPU baseline is the
bit-sliced C grader (oracle/bitsliced.c, 64 pairs per word, OpenMP over the
host's threads) over the whole workload on this host, with its one-core rate
and the serial byte-per-pattern oracle's rate on a bounded sample beside it.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path[:0] = [ROOT, os.path.join(ROOT, "gpu-path-delay-coverage_amd")]

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, 8 TB/s


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="single-50k-gates")
    ap.add_argument("--cpu-sample", type=int, default=4096,
                    help="pairs graded by the serial oracle beside the CPU baseline (0 = skip every CPU leg)")
    ap.add_argument("--robust", type=int, default=1)
    ap.add_argument("--segs", type=int, default=1,
                    help="single workloads: segment length K (1..8); K >= 2 times pdf_grade_segs")
    ap.add_argument("--graph", type=int, default=None, help="batch workloads: force graph replay on (1) or off (0)")
    ap.add_argument("--learned", type=int, default=1, help="50k workload: run the learned-ranking experiment")
    ap.add_argument("--extra", type=int, default=1,
                    help="50k workload: also time the other marking mode, the cover-heavy input and the batch")
    ap.add_argument("--batch", type=int, default=1, help="with --extra: include the 256-circuit batch line")
    ap.add_argument("--autotune", type=int, default=None,
                    help="run the autotuner before the timed steps (default: no for single workloads, yes for "
                         "batches)")
    ap.add_argument("--tune", action="append", default=[], metavar="KNOB=VALUE",
                    help="single workloads: set a grader tuning knob after the recorded tuning, e.g. fwd_order=1 "
                         "(knobs: the pdfgrade TUNE_* names in lower case); results never change")
    return ap.parse_args()


def measured_traffic(workload: str, kernel: str):
    """HBM bytes per launch of a kernel (k_walk, k_fwd, k_cover ...), from the
    committed rocprofv3 FETCH_SIZE / WRITE_SIZE summary
    (profiles/traffic_<workload>.md, its json block), or None."""
    path = os.path.join(ROOT, "profiles", f"traffic_{workload}.md")
    try:
        with open(path) as f:
            text = f.read()
        block = text.split("```json", 1)[1].split("```", 1)[0]
        return json.loads(block)["bytes_per_launch"].get(kernel)
    except (OSError, IndexError, KeyError, ValueError):
        return None


def host_threads() -> int:
    cores = len(os.sched_getaffinity(0))
    env = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
    return max(1, min(cores, env) if env > 0 else cores)


def host_cpu() -> dict:
    """The host's processor count and model (from /proc/cpuinfo)."""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"nproc": os.cpu_count(), "usable": len(os.sched_getaffinity(0)), "model": model}


def serial_oracle_rate(net, bits, pairs, sample, robust):
    """Serial byte-per-pattern C oracle (one core) on the first `sample` pairs."""
    import numpy as np

    from oracle import oracle as orc

    f0 = np.full(net.n, 0xFFFFFFFF, np.uint32)
    f1 = f0.copy()
    gc = np.zeros(net.n, np.uint64)
    t0 = time.perf_counter()
    codes = orc.sim_chunk(net, bits, pairs, 0, sample)
    marks, _ = orc.mark_chunk(net, codes, robust)
    orc.merge_chunk(net, codes, marks, 0, f0, f1)
    orc.cover_chunk(net, marks, 0, f0, f1, gc)
    return sample / (time.perf_counter() - t0)


def cpu_baseline(nl, bits, wl, sample, robust, gpu):
    """The bit-sliced C grader over the whole workload (same pairs, chunk and
    mode as the GPU steps) on every host thread, then on one; the serial
    oracle on `sample` pairs beside it.  `gpu` holds the GPU run's total,
    first-detection arrays and per-node counts, compared when the CPU leg
    graded the whole workload."""
    import numpy as np

    from oracle import oracle as orc

    net = orc.OracleNet(nl.arrays())
    threads = host_threads()
    # whole workload up to 2^32 node-pairs (about 1 s on 16 cores), else a prefix
    pairs = min(wl.pairs, max(64, (2**32 // net.n) // 64 * 64))
    chunk = min(wl.chunk, pairs)
    whole = pairs == wl.pairs
    what = f"all {pairs}" if whole else f"first {pairs} (graded as a run of their own) of {wl.pairs}"
    rate = {}
    for t in sorted({threads, 1}, reverse=True):
        t0 = time.perf_counter()
        r = orc.grade_bitsliced(net, bits, pairs, chunk, robust, threads=t)
        rate[t] = pairs / (time.perf_counter() - t0)
    return {"value": rate[threads], "unit": "pairs/s", "cores": threads, "kind": "port",
            "sample": (f"{what} pairs of the {wl.name} netlist, chunk {chunk}, bit-sliced C grader "
                       f"(oracle/bitsliced.c), {threads} OpenMP threads"),
            "one_core": rate[1], "total_paths": r["total"], "host": host_cpu(),
            "matches_gpu": (r["total"] == gpu["total"] and np.array_equal(r["first_t0"], gpu["first"][0])
                            and np.array_equal(r["first_t1"], gpu["first"][1])
                            and np.array_equal(r["gate_counts"], gpu["gate_counts"])) if whole else None,
            "serial_oracle": {"value": serial_oracle_rate(net, bits, wl.pairs, sample, robust), "unit": "pairs/s",
                              "cores": 1, "sample": f"first {sample} pairs, byte-per-pattern oracle.c, one chunk"}}


def segs_cpu_baseline(nl, bits, wl, sample, robust, K):
    """Serial C oracle with segments of length K (one core) on the first
    `sample` pairs, graded as a vector set of their own."""
    from oracle import oracle as orc

    net = orc.OracleNet(nl.arrays())
    t0 = time.perf_counter()
    orc.grade_segs(net, bits, sample, sample, robust, K)
    return {"value": sample / (time.perf_counter() - t0), "unit": "pairs/s", "cores": 1, "kind": "port",
            "sample": f"first {sample} of {wl.pairs} pairs of the {wl.name} netlist, one chunk, serial C oracle "
                      f"with segments of length {K}", "host": host_cpu()}


def bench_batch(args, wl):
    """One batched grade of every circuit of the workload against per-circuit
    plain grading (one pdf_grade per circuit, one after another)."""
    import numpy as np

    import pdfgrade as pg

    robust = bool(args.robust)
    members = wl.members()
    nets, vecs, pairs = [], [], []
    for cfg, vseed, p in members:
        nets.append(pg.Netlist.generate(cfg))
        vecs.append(pg.gen_vectors(vseed, cfg.num_pi, p))
        pairs.append(p)
    infos = [nl.info() for nl in nets]
    gate_pairs = sum((i["n"] - i["num_pi"]) * p for i, p in zip(infos, pairs))
    algo_bytes = sum((10 * i["n"] + i["num_pi"]) / 8.0 * p for i, p in zip(infos, pairs))
    fwd_bytes = sum((5 * i["n"] + i["num_pi"]) / 8.0 * p for i, p in zip(infos, pairs))  # sim + forward mark
    # per-circuit plain grading, one circuit at a time (each warmed up first)
    plain_ms, plain_tot = 0.0, []
    for nl, v, p in zip(nets, vecs, pairs):
        g = pg.Grader(nl)
        g.set_vectors(v, p)
        g.set_chunk(p)
        if args.graph is not None:
            g.set_tuning(pg.TUNE_USE_GRAPH, args.graph)
        g.grade(robust)
        t0 = time.perf_counter()
        plain_tot.append(g.grade(robust))
        plain_ms += (time.perf_counter() - t0) * 1e3
        g.close()
    b = pg.Batch(nets, vecs, pairs)
    tune = b.autotune(robust, reps=2) if args.autotune in (None, 1) else None
    if args.graph is not None:
        b.set_tuning(pg.TUNE_USE_GRAPH, args.graph)
    for _ in range(args.warmup):
        b.grade(robust)
    t0 = time.perf_counter()
    gpu_ms = []
    for _ in range(args.steps):
        tot, ms = b.grade(robust)
        gpu_ms.append(ms)
    dt = (time.perf_counter() - t0) / args.steps
    assert list(tot) == plain_tot, "batched totals differ from per-circuit grading"
    out = {
        "metric": "vector-pairs graded/s, 256-circuit batch", "value": sum(pairs) / dt, "unit": "pairs/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": dt * 1e3, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "u64", "data": "synthetic (seeded netlists and vectors)",
        "config": {"workload": wl.name, "circuits": wl.count, "pairs": sum(pairs), "gate_pairs": gate_pairs,
                   "gates_min": min(i["n"] - i["num_pi"] for i in infos),
                   "gates_max": max(i["n"] - i["num_pi"] for i in infos), "mode": "robust" if robust else "non-robust",
                   "tuning": tune},
        "gate_evals_per_s": gate_pairs / dt,
        "gpu_ms_per_step": float(np.median(gpu_ms)),
        "per_circuit_plain_ms": plain_ms,
        "speedup_vs_per_circuit": plain_ms / (dt * 1e3),
        # the dense planes of every pass, (10 n + I)/8 bytes per pair, over the
        # grade time: the reverse pass here reads only the (gate, 64-word
        # tile) items that hold a mark, so this passes the HBM peak and is a
        # rate in dense-plane terms, not a measured bandwidth
        "dense_plane_equivalent": {"GBps": algo_bytes / dt / 1e9, "over_hbm_peak": algo_bytes / dt / 1e9 / HBM_PEAK_GBS},
        "forward_algorithmic": {"GBps": fwd_bytes / dt / 1e9, "frac": fwd_bytes / dt / 1e9 / HBM_PEAK_GBS},
        "roofline": batch_roofline(wl.name, fwd_bytes, max(i["levels"] for i in infos) + 1),
        "cpu_baseline": batch_cpu_baseline(nets, vecs, pairs, robust, tot) if args.cpu_sample > 0 else None,
    }
    b.close()
    return out


def batch_roofline(wl_name, fwd_bytes, launches):
    """The batch's dominant kernel, k_fwd_batch (one launch per union level:
    level k of every circuit deep enough): its algorithmic bytes, (5n + I)/8
    per pair of every circuit, over the average duration in the committed
    trace of this workload, and the committed counter traffic over it (the
    batch has no per-kernel event timing)."""
    per = fwd_bytes / launches
    rp_ns = trace_avg_ns(wl_name, "k_fwd_batch")
    traffic = measured_traffic(wl_name, "k_fwd_batch")
    achieved = per / (rp_ns * 1e-9) / 1e9 if rp_ns else None
    return {"bound": "hbm", "kernel": "k_fwd_batch", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS if achieved else None, "traffic": traffic,
            "algorithmic_bytes_per_launch": per, "launches_per_step": launches,
            "rocprof_avg_us": rp_ns / 1e3 if rp_ns else None,
            "frac_rocprof": achieved / HBM_PEAK_GBS if achieved else None,
            "traffic_frac": traffic / (rp_ns * 1e-9) / 1e9 / HBM_PEAK_GBS if traffic and rp_ns else None}


def batch_cpu_baseline(nets, vecs, pairs, robust, gpu_totals, stride=8):
    """The bit-sliced C grader on every `stride`-th circuit of the batch (each
    whole, 16,384-pair chunks) on every host thread: pairs/s over that sample,
    and whether its totals equal the batched GPU grade's."""
    from oracle import oracle as orc

    threads = host_threads()
    idx = list(range(0, len(nets), stride))
    onets = {i: orc.OracleNet(nets[i].arrays()) for i in idx}
    t0 = time.perf_counter()
    totals = [orc.grade_bitsliced(onets[i], vecs[i], pairs[i], 16384, robust, threads=threads)["total"] for i in idx]
    dt = time.perf_counter() - t0
    return {"value": sum(pairs[i] for i in idx) / dt, "unit": "pairs/s", "cores": threads, "kind": "port",
            "sample": f"circuits {idx[0]}, {idx[1]}, ... (every {stride}th, {len(idx)} of {len(nets)}), all their "
                      f"pairs, bit-sliced C grader (oracle/bitsliced.c), {threads} OpenMP threads",
            "matches_gpu": all(int(gpu_totals[i]) == t for i, t in zip(idx, totals)), "host": host_cpu()}


def learned_vs_plain(robust: bool):
    """Held-out learned ranking against plain order: `python -m
    pdfgrade.learned` in a child process (it shares torch's HIP runtime,
    which this process does not load)."""
    import subprocess
    import tempfile

    with tempfile.TemporaryDirectory() as tmp:
        out = os.path.join(tmp, "learned.json")
        cmd = [sys.executable, "-m", "pdfgrade.learned", "--out", out] + ([] if robust else ["--non-robust"])
        # the child's progress lines go to this process's stderr (stdout holds
        # only the JSON line)
        print(f"learned experiment ({'robust' if robust else 'non-robust'}) ...", file=sys.stderr, flush=True)
        r = subprocess.run(cmd, cwd=os.path.join(ROOT, "gpu-path-delay-coverage_amd"), stdout=sys.stderr)
        if r.returncode != 0:
            raise RuntimeError(f"learned experiment failed ({r.returncode}); its output is on stderr above")
        with open(out) as f:
            res = json.load(f)
    s95, g95 = res["summary"]["paths_95"], res["summary"]["segments_95"]
    live = [c for c in res["circuits"] if c["full_paths"] > 0]
    per = [c["learned"] for c in live] or [c["learned"] for c in res["circuits"]]

    def mean(xs):
        xs = [x for x in xs if x is not None]
        return sum(xs) / len(xs) if xs else None

    return {
        "mode": "robust" if robust else "non-robust",
        # the metric: learned time to 95% of the paths over plain in-order
        # time to 95%, both graded in the same chunks (circuits with paths)
        "time_to_95pct_paths_learned_over_plain": s95["learned_over_plain_same_fraction"],
        "time_to_95pct_paths_learned_over_plain_full_grading": s95["learned_over_plain_full"],
        "time_to_95pct_paths_ms": {"learned": mean([c["learned"]["t95_paths_ms"] for c in live]),
                                   "plain": mean([c["plain_curve"]["t95_paths_ms"] for c in live])},
        "pairs_to_95pct_paths": {"learned": s95["pairs_learned_mean"], "plain": s95["pairs_plain_mean"]},
        "pairs_to_95pct_segments": {"learned": g95["pairs_learned_mean"], "plain": g95["pairs_plain_mean"]},
        "learned_stage_ms": {k: mean([p[k] for p in per])
                             for k in ("features_score_ms", "rank_ms", "features_ms", "score_ms")},
        "plain_full_ms": mean([c["plain_ms"] for c in res["circuits"]]),
        "heldout_circuits": len(res["circuits"]), "circuits_without_paths": s95["circuits_without_any"],
        "candidates": res["config"]["candidates"], "records": res["records"],
    }


def trace_avg_ns(workload: str, kernel: str):
    """Average duration of a kernel in the committed rocprofv3 --kernel-trace
    --stats summary of a workload (profiles/k_<kernel>_<workload>.md, the
    "avg_ns" map of its json block), or None."""
    import glob

    for path in sorted(glob.glob(os.path.join(ROOT, "profiles", f"k_*_{workload}.md"))):
        try:
            with open(path) as f:
                text = f.read()
            avg = json.loads(text.split("```json", 1)[1].split("```", 1)[0]).get("avg_ns", {})
        except (OSError, IndexError, ValueError, AttributeError):
            continue
        if kernel in avg:
            return float(avg[kernel])
    return None


def walk_model_bytes(nl, chunk: int, tw: int, seq: bool = True) -> dict:
    """Bytes one level-walk launch must move in its design (the chunk's
    records do not fit on chip): the record of every node that a later
    level reads from outside the workgroup's LDS window written once per
    tile (its reach, the end of its last reader's level minus its id,
    exceeds the window), each such fanin read back once per tile, one
    mark-present byte per node and tile written and read back, and the PI
    words (plus one halo word per tile) read.  The mark words are
    data-dependent and left out (the counters include them)."""
    import numpy as np

    a = nl.arrays()
    n, I = a["n"], a["num_pi"]
    words = (chunk + 63) // 64
    tiles = -(-words // tw)
    nf = tw + 1 if seq else 2 * tw
    nv = (nf + 2) & ~1
    sv = nv + ((tw + 1) & ~1)  # an LDS slot: the record, then the mark words
    cap = (160 * 1024) // (8 * sv)
    rec = 8 * nv
    lo = a["level_off"]
    types, fin_off, fin = a["type"], a["fin_off"], a["fin"]
    nfi = np.diff(fin_off)
    lvl = a["level"]
    hi = lo[lvl + 1].astype(np.int64)  # end of each node's level
    # fanin j of gate g is read when it is fanin 0, or any fanin of a
    # multi-input AND/NAND/OR/NOR
    multi = (types >= 2) & (types <= 5)  # PDF_AND .. PDF_NOR
    gate = np.repeat(np.arange(n), nfi)
    j = np.arange(fin.shape[0]) - np.repeat(fin_off[:-1], nfi)
    read = (j == 0) | multi[gate]
    dist = hi[gate] - fin.astype(np.int64)
    far = int(np.count_nonzero(read & (dist > cap)))
    reach = np.zeros(n, np.int64)
    np.maximum.at(reach, fin[read].astype(np.int64), dist[read])
    kept = int(np.count_nonzero(reach > cap))
    writes = kept * tiles * rec
    bits = 2 * n * tiles
    reads = far * tiles * rec + I * (words + (tiles if seq else 0)) * 8
    return {"bytes": writes + reads + bits, "record_writes": writes, "far_fanin_reads": far * tiles * rec,
            "pi_reads": I * (words + (tiles if seq else 0)) * 8, "bit_bytes": bits, "far_fanins": far,
            "kept_records": kept, "tiles": tiles, "record_bytes": rec, "window_nodes": cap}


def roofline_fields(wl_name, fam, kernel, algo_bytes_launch, avg_ms, launches, model=None):
    """The roofline object of the dominant kernel: algorithmic bytes per
    launch over its HIP-event average (frac, frac_events), over the average
    of the committed rocprofv3 trace of this workload (frac_rocprof), the
    committed counter traffic over the event average (traffic_frac), and the
    implementation's own byte model when given."""
    achieved = algo_bytes_launch / (avg_ms * 1e-3) / 1e9 if avg_ms > 0 else 0.0
    traffic = measured_traffic(wl_name, kernel)
    rp_ns = trace_avg_ns(wl_name, kernel)
    out = {"bound": "hbm", "kernel": kernel, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
           "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
           "algorithmic_bytes_per_launch": algo_bytes_launch, "avg_launch_us": avg_ms * 1e3,
           "launches_per_step": launches, "frac_events": achieved / HBM_PEAK_GBS,
           "rocprof_avg_us": rp_ns / 1e3 if rp_ns else None,
           "frac_rocprof": algo_bytes_launch / (rp_ns * 1e-9) / 1e9 / HBM_PEAK_GBS if rp_ns else None,
           "traffic_frac": traffic / (avg_ms * 1e-3) / 1e9 / HBM_PEAK_GBS if traffic and avg_ms > 0 else None}
    if model:
        out["model_bytes_per_launch"] = model["bytes"]
        out["model_frac"] = model["bytes"] / (avg_ms * 1e-3) / 1e9 / HBM_PEAK_GBS if avg_ms > 0 else None
        out["model"] = model
    return out


def setup_grader(pg, nl, bits, wl, robust, autotune):
    """A grader with the workload's recorded tuning (what pdf_grader_autotune
    chose on an MI355X), or the autotuner's choice, untimed."""
    g = pg.Grader(nl, device=0)
    g.set_vectors(bits, wl.pairs)
    # chunk first: the lanes' buffers are sized for the current chunk
    g.set_chunk(wl.chunk)
    g.set_tuning(pg.TUNE_LANES, wl.lanes)
    g.set_tuning(pg.TUNE_WORDS_PER_LANE, wl.words_per_lane)
    g.set_tuning(pg.TUNE_USE_GRAPH, wl.use_graph)
    g.set_tuning(pg.TUNE_WALK, wl.walk)
    g.set_tuning(pg.TUNE_REV_COMPACT, wl.rev_compact)
    if autotune:
        tune = g.autotune(robust, reps=5)
    else:
        tune = {"lanes": wl.lanes, "chunk_pairs": wl.chunk, "use_graph": wl.use_graph,
                "words_per_lane": wl.words_per_lane, "walk": wl.walk, "rev_compact": wl.rev_compact,
                "source": "recorded"}
    return g, tune


def timed_grades(grade, warmup, steps):
    total = None
    for _ in range(warmup):
        total = grade()
    t0 = time.perf_counter()
    for _ in range(steps):
        t = grade()
        assert total is None or t == total
        total = t
    return total, (time.perf_counter() - t0) / steps


def dominant(kt):
    return max(("fwd", "reverse", "cover", "events", "merge"), key=lambda k: kt[f"{k}_ms"])


def fam_bytes_per_pair(fam, n, I, K, num_segs, walk):
    """Algorithmic bytes per pair of a kernel family (the per-stage figures of
    the measurement plan): sim (2n + I)/8 and forward mark 3n/8 for the fused
    forward kernel; removal 2n/8 and merge 2n/8 for the fused reverse kernel;
    the level walk fuses all four; for segments of length K, each of the S
    segments reads K mark bits and node 0's final-value bit per pair."""
    if fam == "fwd":
        return (9 * n + I) / 8.0 if walk else (5 * n + I) / 8.0
    if fam == "reverse":
        return 4 * n / 8.0
    if fam == "merge":
        return num_segs * (K + 1) / 8.0
    if fam == "cover":
        return n / 8.0
    return 0.0


def side_line(pg, nl, bits, wl, robust, steps, warmup, label):
    """A secondary configuration of the 50k workload (same recorded tuning):
    throughput, the grade's per-kernel HIP-event times and the paths found."""
    g, _ = setup_grader(pg, nl, bits, wl, robust, False)
    total, dt = timed_grades(lambda: g.grade(robust), warmup, steps)
    t_total, kt = g.grade_timed(robust)
    assert t_total == total
    g.close()
    fam = dominant(kt)
    out = {"what": label, "mode": "robust" if robust else "non-robust", "value": wl.pairs / dt, "unit": "pairs/s",
           "ms_per_step": dt * 1e3, "steps": steps, "total_paths": total,
           "kernel_ms": {k: kt[k] for k in ("fwd_ms", "reverse_ms", "events_ms", "cover_ms", "fill_ms", "merge_ms",
                                            "wall_ms")},
           "kernel_launches": {k: kt[k] for k in ("fwd_launches", "reverse_launches", "events_launches",
                                                  "cover_launches")},
           "dominant": f"k_{'walk' if fam == 'fwd' and wl.walk else fam}"}
    return out


def cover_fields(wl_name, kt):
    """k_cover and k_events of a grade: HIP-event average per launch and the
    committed counter traffic of this workload over it."""
    out = {}
    for fam in ("cover", "events"):
        launches = kt[f"{fam}_launches"]
        if not launches:
            continue
        avg_ms = kt[f"{fam}_ms"] / launches
        traffic = measured_traffic(wl_name, f"k_{fam}")
        rp_ns = trace_avg_ns(wl_name, f"k_{fam}")
        out[f"k_{fam}"] = {"avg_launch_us": avg_ms * 1e3, "launches_per_step": launches, "traffic": traffic,
                           "traffic_GBps": traffic / (avg_ms * 1e-3) / 1e9 if traffic else None,
                           "traffic_frac": traffic / (avg_ms * 1e-3) / 1e9 / HBM_PEAK_GBS if traffic else None,
                           "rocprof_avg_us": rp_ns / 1e3 if rp_ns else None}
    return out


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "grading runs on one GPU; --gpus must be 1 (sharding over GPUs is designed, "
                                   "not built)"}))
        sys.exit(2)
    import pdfgrade as pg
    from pdfgrade.configs import BATCH_WORKLOADS, COVER_HEAVY, WORKLOADS

    if args.workload in BATCH_WORKLOADS:
        return print(json.dumps(bench_batch(args, BATCH_WORKLOADS[args.workload])))

    wl = WORKLOADS[args.workload]
    nl = pg.Netlist.generate(wl.netlist)
    info = nl.info()
    n, I = info["n"], info["num_pi"]
    bits = pg.gen_vectors(wl.vector_seed, I, wl.pairs)
    prof_name = wl.name
    robust = bool(args.robust)
    g, tune = setup_grader(pg, nl, bits, wl, robust, args.autotune)
    for kv in args.tune:
        knob, value = kv.split("=", 1)
        g.set_tuning(getattr(pg, "TUNE_" + knob.upper()), int(value))
        tune[knob] = int(value)
        tune["source"] = "recorded + --tune"
    K = args.segs
    grade = (lambda: g.grade(robust)) if K == 1 else (lambda: g.grade_segs(robust, K))
    total, dt = timed_grades(grade, args.warmup, args.steps)
    pairs_per_s = wl.pairs / dt
    ms_per_step = dt * 1e3

    # live per-launch timing of the dominant kernel family: the same chunk
    # shape with the launches serialized (lanes overlap kernels, which would
    # blur one kernel's duration)
    t_total, kt = g.grade_timed(robust) if K == 1 else g.grade_segs(robust, K, timed=True)
    assert t_total == total
    fam = dominant(kt)
    walk = bool(wl.walk) and K == 1
    kernel = "k_walk" if fam == "fwd" and walk else f"k_{fam}"
    num_segs = nl.segments(K).shape[0] if K > 1 else n
    launches = max(1, kt[f"{fam}_launches"])
    algo_launch = fam_bytes_per_pair(fam, n, I, K, num_segs, walk) * wl.pairs / launches
    avg_ms = kt[f"{fam}_ms"] / launches
    model = walk_model_bytes(nl, min(wl.chunk, wl.pairs), wl.walk) if kernel == "k_walk" else None
    roof = roofline_fields(prof_name, fam, kernel, algo_launch, avg_ms, launches, model)
    # end-to-end algorithmic rate (10n + I bits per pair: sim, mark, remove,
    # merge, cover each reading/writing its dense planes once)
    e2e_bytes = (10 * n + I) / 8.0 * wl.pairs
    e2e_gbs = e2e_bytes / dt / 1e9

    cpu = None
    if args.cpu_sample > 0 and K == 1:
        gpu = {"total": total, "first": g.first(), "gate_counts": g.gate_counts()}
        cpu = cpu_baseline(nl, bits, wl, min(args.cpu_sample, wl.pairs), robust, gpu)
    elif args.cpu_sample > 0:
        cpu = segs_cpu_baseline(nl, bits, wl, min(args.cpu_sample, wl.pairs), robust, K)
    g.close()

    headline = wl.name == "single-50k-gates" and K == 1
    side = None
    if args.extra and headline:
        side = {"non_robust": side_line(pg, nl, bits, wl, not robust, args.steps, args.warmup,
                                        "the same workload in the other marking mode"),
                "cover_heavy": None, "batch": None}
        ch = COVER_HEAVY
        ch_nl = pg.Netlist.generate(ch.netlist)
        ch_bits = pg.gen_vectors(ch.vector_seed, ch_nl.info()["num_pi"], ch.pairs)
        line = side_line(pg, ch_nl, ch_bits, ch, robust, args.steps, args.warmup,
                         f"{ch.name}: 50,000 gates in 16 levels, 70% NOT/BUFF, 4,096 POs; many more marked "
                         f"paths, so the merge and the cover DP carry the load")
        ch_kt = dict(line["kernel_ms"], **line["kernel_launches"])
        ch_n = ch_nl.info()["n"]
        line["cover_kernels"] = cover_fields(ch.name, ch_kt)
        # the cover reads the mark plane once: n/8 bytes per pair
        if ch_kt["cover_launches"]:
            line["cover_roofline"] = roofline_fields(ch.name, "cover", "k_cover", ch_n / 8.0 * ch.pairs /
                                                     ch_kt["cover_launches"],
                                                     ch_kt["cover_ms"] / ch_kt["cover_launches"],
                                                     ch_kt["cover_launches"])
        side["cover_heavy"] = line
        if args.batch:
            from argparse import Namespace
            bargs = Namespace(**vars(args))
            bargs.steps, bargs.warmup = 5, 2
            side["batch"] = bench_batch(bargs, BATCH_WORKLOADS["batch-256-circuits"])

    learned_part = None
    if args.learned and headline:
        learned_part = [learned_vs_plain(robust), learned_vs_plain(not robust)]

    out = {
        "metric": ("vector-pairs graded/s at 50k gates; time-to-95%-coverage learned vs plain"
                   if headline else
                   f"vector-pairs graded/s ({prof_name}{f', segments of length {K}' if K > 1 else ''})"),
        "value": pairs_per_s,
        "unit": "pairs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u64",
        "data": "synthetic (seeded netlist and vectors)",
        "config": {"workload": wl.name, "gates": info["n"] - I, "levels": info["levels"], "pis": I,
                   "pos": info["num_po"], "pairs": wl.pairs, "chunk": wl.chunk,
                   "mode": "robust" if robust else "non-robust", "total_paths": total, "tuning": tune,
                   "segs": K, "segments": num_segs},
        "gate_evals_per_s": pairs_per_s * (n - I),
        "roofline": roof,
        "end_to_end": {"algorithmic_GBps": e2e_gbs, "frac": e2e_gbs / HBM_PEAK_GBS},
        "kernel_ms": {k: kt[k] for k in ("fwd_ms", "reverse_ms", "events_ms", "cover_ms", "fill_ms", "merge_ms",
                                         "wall_ms")},
        "cover_kernels": cover_fields(prof_name, kt),
        "cpu_baseline": cpu,
        "also": side,
        "learned_vs_plain": learned_part,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
