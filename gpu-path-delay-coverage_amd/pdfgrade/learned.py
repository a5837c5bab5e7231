"""Learned pair ranking: grader-logged training records, MLP training on the
library's MFMA kernels, ranking of candidate pairs and exact evaluation.

No reference counterpart: this sits on top of the grading path.  The grader
produces every label (each pair's exact solo gain, pdf_pair_gains) and
scores every order (exact in-order grading, pdf_grade_curve); the network
only proposes the order.  Unranked in-order grading stays the fallback.

Candidate pairs are explicit (A, B): A uniform, B = A with each PI flipped
with a per-pair probability q log-uniform in [q_lo, q_hi] (pdf_gen_toggle_
pairs), so pairs range from single-input launches to half the inputs
toggling.

Coverage is reported two ways for each order of the candidate pairs, both
exact and per pair:
  paths     the grader's running path total (coverkernel.cuh:171-230
            semantics: per pair, the marked PI->PO paths through a node
            first detected by that pair), against the total of plain
            in-order grading of every candidate;
  segments  (node, transition) segments detected so far, read from the
            first-detection indices (mergekernel.cuh:141-197), against
            those detected by all candidates (the same set in every order).
Time to a coverage fraction is the event time at the end of the chunk that
holds the pair reaching it, plus, for the learned order, the wall time of
features, scoring and ranking.

Run as `python -m pdfgrade.learned` for the full experiment (32 training
and 8 held-out 50k-gate circuits); it prints one JSON object.
"""
from __future__ import annotations

import argparse
import json
import time
from dataclasses import asdict, dataclass, field

import numpy as np
import torch  # before libpdfgrade.so is loaded: the library then shares torch's HIP runtime

from . import (FEATURES, NONE, TUNE_LANES, TUNE_USE_GRAPH, TUNE_WORDS_PER_LANE, GenConfig, Grader, Mlp, Netlist,
               gen_toggle_pairs)

WIDE_FANIN = (0.1, 0.6, 0.2, 0.1)


def circuit(seed: int, num_gates: int = 50_000, levels: int = 60, num_pi: int = 256,
            num_po: int = 128) -> GenConfig:
    """A 50k-gate-style synthetic netlist (fan-in 1-4, power-law fanout)."""
    return GenConfig(seed=seed, num_pi=num_pi, num_po=num_po, num_gates=num_gates, levels=levels,
                     fanin_p=WIDE_FANIN, prev_level_frac=0.7, fanout_alpha=2.2, fanout_cap=64)


def candidate_bits(num_pi: int, count: int, seed: int, q_lo: float, q_hi: float):
    """Explicit candidate pairs of varied toggle density (see module doc)."""
    return gen_toggle_pairs(seed, num_pi, count, q_lo, q_hi)


@dataclass
class LearnedConfig:
    train_seeds: tuple = tuple(range(5000, 5032))
    heldout_seeds: tuple = tuple(range(5100, 5108))
    train_pair_seed: int = 16000     # training circuit i: pairs from seed 16000 + i
    candidate_seed: int = 15000      # held-out circuit j: candidates from seed 15000 + j
    q_lo: float = 1.0 / 256          # per-pair PI flip probability range
    q_hi: float = 0.5
    train_pairs: int = 65536         # per training circuit: 32 x 65,536 = 2,097,152 records
    candidates: int = 1 << 20        # per held-out circuit
    num_gates: int = 50_000
    levels: int = 60
    num_pi: int = 256
    num_po: int = 128
    max_level: int = 4               # feature simulation depth (8 and 16 ranked no better)
    batch: int = 8192
    epochs: int = 4
    lr: float = 1e-3
    robust: bool = True
    chunk: int = 65536
    curve_chunk: int = 16384         # chunk of the ranked (learned) grading; 0 = chunk
    lanes: int = 2                   # chunk lanes of the candidate grader
    fractions: tuple = (0.9, 0.95, 0.99)
    seed: int = 1


def _round64(x: int) -> int:
    return (x + 63) // 64 * 64


def _torch():
    return torch


# ------------------------------------------------------------------ records

def log_records(cfg: LearnedConfig, device: str = "cuda:0"):
    """Features (512 x records, BF16 bits in int16) and exact solo gains
    (int32 holding uint32) of the training pairs, all on the GPU."""
    torch = _torch()
    total = len(cfg.train_seeds) * cfg.train_pairs
    feat = torch.zeros((FEATURES, _round64(total) + 64), dtype=torch.int16, device=device)
    gains = torch.zeros(total, dtype=torch.int32, device=device)
    for i, s in enumerate(cfg.train_seeds):
        nl = Netlist.generate(circuit(s, cfg.num_gates, cfg.levels, cfg.num_pi, cfg.num_po))
        a, b = candidate_bits(cfg.num_pi, cfg.train_pairs, cfg.train_pair_seed + i, cfg.q_lo, cfg.q_hi)
        with Grader(nl) as g:
            g.set_pairs(a, b, cfg.train_pairs)
            g.set_chunk(cfg.chunk)
            c0 = i * cfg.train_pairs
            g.pair_features(cfg.max_level, cfg.robust, feat[:, c0:c0 + cfg.train_pairs])
            g.pair_gains(cfg.robust, gains[c0:c0 + cfg.train_pairs])
    return feat, gains


def gain_target(gains, per_circuit: int = 0):
    """Regression target: log(1 + gain), standardised over all records, or
    within each circuit's block of `per_circuit` records.  Returns (target,
    mean, std) of the unstandardised values.  (Per-circuit standardising,
    shuffled batches, binary targets and more epochs were measured on
    held-out circuits and did not rank better.)"""
    torch = _torch()
    g = gains.to(torch.int64) & 0xFFFFFFFF
    t = torch.log1p(g.to(torch.float64))
    mean, std = float(t.mean()), float(t.std())
    blocks = t.view(-1, per_circuit) if per_circuit else t.view(1, -1)
    m = blocks.mean(dim=1, keepdim=True)
    sd = blocks.std(dim=1, keepdim=True).clamp_min(1e-6)
    return ((blocks - m) / sd).reshape(-1).to(torch.float32).contiguous(), mean, std



# ----------------------------------------------------------------- training

def train(mlp: Mlp, feat, target, rows: int, cfg: LearnedConfig):
    """Adam over shuffled batches of consecutive records; the learning rate
    decays linearly to a tenth.  Returns the last-batch loss of each epoch."""
    nb = rows // cfg.batch
    rng = np.random.default_rng(cfg.seed)
    steps = nb * cfg.epochs
    losses = []
    k = 0
    for _ in range(cfg.epochs):
        order = rng.permutation(nb)
        loss = None
        for j, bi in enumerate(order):
            c0 = int(bi) * cfg.batch
            lr = cfg.lr * (1.0 - 0.9 * k / max(1, steps - 1))
            loss = mlp.train_step(feat[:, c0:c0 + cfg.batch], target[c0:c0 + cfg.batch], lr,
                                  want_loss=j == nb - 1)
            k += 1
        losses.append(loss)
    mlp.sync()
    return losses


# --------------------------------------------------------------- evaluation

def _segment_pairs(first0, first1):
    """Sorted first-detection pair indices of the detected segments."""
    return np.sort(np.concatenate([first0[first0 != NONE], first1[first1 != NONE]]).astype(np.int64))


def _pairs_to(cum, target):
    """Pairs graded (in order) until the running value reaches target."""
    if target <= 0:
        return 0
    idx = int(np.searchsorted(cum, target, side="left"))
    return None if idx >= len(cum) else idx + 1


class _Order:
    """Exact coverage of one order: per-pair path totals, segment first
    indices and per-chunk event times."""

    def __init__(self, g: Grader, robust: bool, chunk: int, device: str):
        torch = _torch()
        P = g._num_pairs
        pt = torch.zeros(P, dtype=torch.int32, device=device)
        self.cum_chunk, self.ms = g.grade_curve(robust, pt)
        self.cum = np.cumsum(pt.cpu().numpy().view(np.uint32).astype(np.uint64))
        f0, f1 = g.first()
        self.seg = _segment_pairs(f0, f1)
        self.chunk = chunk

    def time_at(self, pairs):
        return None if pairs is None else float(self.ms[max(0, pairs - 1) // self.chunk])


def ranking_quality(scores, gains) -> dict:
    """How well the scores order the exact solo gains: Spearman correlation,
    and the share of positive-gain pairs among the top 1% / 10% of scores
    relative to their share overall (enrichment; 1 = no better than random)."""
    torch = _torch()
    P = scores.numel()
    rs = torch.empty(P, dtype=torch.float64, device=scores.device)
    rg = torch.empty_like(rs)
    ar = torch.arange(P, dtype=torch.float64, device=scores.device)
    rs[torch.argsort(scores, stable=True)] = ar
    rg[torch.argsort(gains, stable=True)] = ar
    rs, rg = rs - rs.mean(), rg - rg.mean()
    spearman = float((rs * rg).sum() / (rs.norm() * rg.norm()).clamp_min(1e-30))
    pos = (gains > 0).to(torch.float64)
    base = float(pos.mean())
    order = torch.argsort(-scores, stable=True)
    out = {"spearman": spearman, "positive_share": base}
    for frac in (0.01, 0.1):
        k = max(1, int(P * frac))
        share = float(pos[order[:k]].mean())
        out[f"enrichment_top{int(frac * 100)}"] = share / base if base > 0 else None
    return out


def evaluate_circuit(mlp: Mlp, nl: Netlist, a, b, cfg: LearnedConfig, device: str = "cuda:0",
                     with_gain_order: bool = True, plain_tuning: dict | None = None) -> dict:
    """Grade the candidates in plain order, in learned order and (reference)
    in exact solo-gain order; return coverage and timing figures.  The plain
    grade runs at plain_tuning (a pdf_grader_autotune result), or at what the
    autotuner picks for this circuit when none is given."""
    torch = _torch()
    P = cfg.candidates
    feat = torch.zeros((FEATURES, _round64(P)), dtype=torch.int16, device=device)
    scores = torch.zeros(P, dtype=torch.float32, device=device)
    g = Grader(nl)
    g.set_pairs(a, b, P)
    # plain in-order grading of every candidate at its fastest configuration:
    # the grader's autotuner (lanes x chunk x graph replay x words per lane),
    # then a warm-up and the timed grade
    if plain_tuning is None:
        plain_tuning = g.autotune(cfg.robust, reps=3)
    else:
        g.set_chunk(plain_tuning["chunk_pairs"])
        g.set_tuning(TUNE_LANES, plain_tuning["lanes"])
        g.set_tuning(TUNE_USE_GRAPH, plain_tuning["use_graph"])
        g.set_tuning(TUNE_WORDS_PER_LANE, plain_tuning["words_per_lane"])
    g.grade(cfg.robust)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    g.grade(cfg.robust)
    grade_ms = (time.perf_counter() - t0) * 1e3
    # the curves and the learned pass use the experiment's chunking
    g.set_tuning(TUNE_LANES, cfg.lanes)
    g.set_tuning(TUNE_USE_GRAPH, 1)
    g.set_tuning(TUNE_WORDS_PER_LANE, 1)
    g.set_chunk(cfg.chunk)
    g.grade(cfg.robust)                      # warm-up at this shape (graph capture)
    out = {"pairs": P, "chunks": g.num_chunks(), "plain_tuning": plain_tuning}

    # untimed warm-up of the learned pass (first launches, scratch), with the
    # separate feature and scoring times for reference; the scores are then
    # reset, so the ranking below is the identity
    g.score_pairs(mlp, cfg.max_level, cfg.robust, feat, scores)
    torch.cuda.synchronize()
    w0 = time.perf_counter()
    g.pair_features(cfg.max_level, cfg.robust, feat)
    w1 = time.perf_counter()
    mlp.forward(feat[:, :P], scores)
    w2 = time.perf_counter()
    scores.zero_()
    g.rank_pairs(scores)                     # equal scores: the identity; allocates the sort scratch
    plain = _Order(g, cfg.robust, cfg.chunk, device)
    full_paths = int(plain.cum_chunk[-1])
    full_segments = len(plain.seg)
    # the plain baseline is the faster of the tuned grade and the curve's own run
    out.update(full_paths=full_paths, full_segments=full_segments, plain_ms=min(grade_ms, float(plain.ms[-1])),
               plain_grade_ms=grade_ms, plain_curve_ms=float(plain.ms[-1]))

    def summarise(name, o: _Order, overhead_ms):
        r = {"overhead_ms": overhead_ms, "final_paths": int(o.cum_chunk[-1]),
             "total_ms": float(o.ms[-1]) + overhead_ms}
        for fr in cfg.fractions:
            key = f"{int(round(fr * 100))}"
            kp = _pairs_to(o.cum, fr * full_paths)
            ks = _pairs_to(np.arange(1, len(o.seg) + 1), fr * full_segments)
            ks = None if ks is None else int(o.seg[ks - 1]) + 1 if ks > 0 else 0
            r[f"pairs{key}_paths"] = kp
            r[f"pairs{key}_segments"] = ks
            r[f"t{key}_paths_ms"] = None if kp is None else o.time_at(kp) + overhead_ms
            r[f"t{key}_segments_ms"] = None if ks is None else o.time_at(ks) + overhead_ms
        out[name] = r

    summarise("plain", plain, 0.0)
    # plain order at the ranked grading's chunk: its time to a fraction on the
    # same footing as the learned order's (smaller chunks finish sooner)
    cc = cfg.curve_chunk or cfg.chunk
    if cc != cfg.chunk:
        g.set_chunk(cc)
        _Order(g, cfg.robust, cc, device)    # warm-up at this shape
        summarise("plain_curve", _Order(g, cfg.robust, cc, device), 0.0)
        g.set_chunk(cfg.chunk)
    else:
        out["plain_curve"] = out["plain"]

    # learned order: features and MLP scores (pipelined), then rank, timed
    # end to end
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    g.score_pairs(mlp, cfg.max_level, cfg.robust, feat, scores)
    t2 = time.perf_counter()
    g.rank_pairs(scores)
    t3 = time.perf_counter()
    # the ranked grading may use smaller chunks than plain full grading (its
    # first chunks are what matters); set up outside the timed region
    if cc != cfg.chunk:
        g.set_chunk(cc)
        _Order(g, cfg.robust, cc, device)    # warm-up at the new shape
    summarise("learned", _Order(g, cfg.robust, cc, device), (t3 - t0) * 1e3)
    out["learned"].update(features_score_ms=(t2 - t0) * 1e3, rank_ms=(t3 - t2) * 1e3,
                          features_ms=(w1 - w0) * 1e3, score_ms=(w2 - w1) * 1e3)

    if with_gain_order:
        # exact solo gains as the score: what a perfect scorer of the target gives
        g.set_pairs(a, b, P)
        g.set_chunk(cfg.chunk)
        gains = torch.zeros(P, dtype=torch.int32, device=device)
        g.pair_gains(cfg.robust, gains)
        gf = (gains.to(torch.int64) & 0xFFFFFFFF).to(torch.float32).contiguous()
        out["ranking_quality"] = ranking_quality(scores, gf)
        g.rank_pairs(gf)
        summarise("gain_order", _Order(g, cfg.robust, cfg.chunk, device), 0.0)
    g.close()
    return out


def run(cfg: LearnedConfig, device: str = "cuda:0", log=print) -> dict:
    torch = _torch()
    t0 = time.perf_counter()
    feat, gains = log_records(cfg, device)
    rows = len(cfg.train_seeds) * cfg.train_pairs
    target, mean, std = gain_target(gains)
    torch.cuda.synchronize()
    t_log = time.perf_counter() - t0
    log(f"logged {rows} records in {t_log:.1f} s; log1p(gain) mean {mean:.3f} std {std:.3f}")
    mlp = Mlp(seed=cfg.seed)
    t0 = time.perf_counter()
    losses = train(mlp, feat, target, rows, cfg)
    t_train = time.perf_counter() - t0
    log(f"trained {cfg.epochs} epochs in {t_train:.1f} s; losses {losses}")
    del feat
    res = {"config": asdict(cfg), "records": rows, "log_s": t_log, "train_s": t_train, "losses": losses,
           "circuits": []}
    # the plain baseline's configuration: autotuned on the first held-out
    # circuit and kept for the others (same gate, level and pair counts)
    plain_tuning = None
    for j, s in enumerate(cfg.heldout_seeds):
        nl = Netlist.generate(circuit(s, cfg.num_gates, cfg.levels, cfg.num_pi, cfg.num_po))
        a, b = candidate_bits(cfg.num_pi, cfg.candidates, cfg.candidate_seed + j, cfg.q_lo, cfg.q_hi)
        r = evaluate_circuit(mlp, nl, a, b, cfg, device, plain_tuning=plain_tuning)
        plain_tuning = r["plain_tuning"]
        r["seed"] = s
        res["circuits"].append(r)
        log(json.dumps({k: r[k] for k in ("seed", "full_paths", "full_segments", "plain_ms", "ranking_quality")})
            + " " +
            json.dumps({o: {k: (round(v, 3) if isinstance(v, float) else v) for k, v in r[o].items()
                            if "95" in k or k == "overhead_ms"}
                        for o in ("plain", "plain_curve", "learned", "gain_order") if o in r}))
    res["summary"] = summarise_runs(res["circuits"], cfg)
    mlp.close()
    return res


def summarise_runs(circuits, cfg: LearnedConfig) -> dict:
    """Per fraction and measure, over the circuits with something to detect
    (a circuit whose candidates detect no path or segment reaches every
    fraction at zero pairs in any order, so it is counted apart):
      learned_over_plain_same_fraction  mean of (learned time to the fraction)
                                        / (plain in-order time to the same
                                        fraction, same chunking) -- the metric
      learned_over_plain_full           mean of (learned time to the fraction)
                                        / (plain grading of every candidate)
    and the pairs each order grades to reach the fraction."""
    out = {}
    for fr in cfg.fractions:
        key = f"{int(round(fr * 100))}"
        for meas in ("paths", "segments"):
            vs_full, vs_same, pairs_l, pairs_p, empty = [], [], [], [], 0
            for r in circuits:
                if r[f"full_{meas}"] == 0:
                    empty += 1
                    continue
                tl, tp = r["learned"][f"t{key}_{meas}_ms"], r["plain_curve"][f"t{key}_{meas}_ms"]
                if tl is None or tp is None:
                    continue
                vs_full.append(tl / r["plain_ms"])
                vs_same.append(tl / tp)
                pairs_l.append(r["learned"][f"pairs{key}_{meas}"])
                pairs_p.append(r["plain_curve"][f"pairs{key}_{meas}"])
            out[f"{meas}_{key}"] = {
                "circuits_reaching": len(vs_full),
                "circuits_without_any": empty,
                "learned_over_plain_same_fraction": float(np.mean(vs_same)) if vs_same else None,
                "learned_over_plain_full": float(np.mean(vs_full)) if vs_full else None,
                "pairs_learned_mean": float(np.mean(pairs_l)) if pairs_l else None,
                "pairs_plain_mean": float(np.mean(pairs_p)) if pairs_p else None,
            }
    return out


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--train-circuits", type=int, default=32)
    ap.add_argument("--heldout-circuits", type=int, default=8)
    ap.add_argument("--train-pairs", type=int, default=65536)
    ap.add_argument("--candidates", type=int, default=1 << 20)
    ap.add_argument("--gates", type=int, default=50_000)
    ap.add_argument("--levels", type=int, default=60)
    ap.add_argument("--max-level", type=int, default=4)
    ap.add_argument("--epochs", type=int, default=4)
    ap.add_argument("--batch", type=int, default=8192)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--chunk", type=int, default=65536)
    ap.add_argument("--curve-chunk", type=int, default=16384, help="chunk of the ranked grading (0 = --chunk)")
    ap.add_argument("--non-robust", action="store_true")
    ap.add_argument("--q-lo", type=float, default=1.0 / 256)
    ap.add_argument("--q-hi", type=float, default=0.5)
    ap.add_argument("--out", default=None)
    args = ap.parse_args(argv)
    cfg = LearnedConfig(train_seeds=tuple(range(5000, 5000 + args.train_circuits)),
                        heldout_seeds=tuple(range(5100, 5100 + args.heldout_circuits)),
                        train_pairs=args.train_pairs, candidates=args.candidates, num_gates=args.gates,
                        levels=args.levels, max_level=args.max_level, epochs=args.epochs, batch=args.batch,
                        lr=args.lr, chunk=args.chunk, curve_chunk=args.curve_chunk, robust=not args.non_robust,
                        q_lo=args.q_lo,
                        q_hi=args.q_hi)
    res = run(cfg, log=lambda s: print(s, flush=True))
    text = json.dumps(res)
    if args.out:
        with open(args.out, "w") as f:
            f.write(text + "\n")
    print(json.dumps(res["summary"]))


if __name__ == "__main__":
    main()
