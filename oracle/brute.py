"""Independent brute-force restatement of the grading semantics, for small
circuits (test infrastructure only).

It shares no code or tables with oracle.c: values are two Boolean frames per
node evaluated with Python operators, sensitization is written from its verbal
definition, and coverage is counted by enumerating every marked PI->PO path
instead of the reverse DP.  The reference behaviours it restates:

  frames    simkernel.cuh:115-148 -- INPT pair (V[p], V[p+1]); BUFF/DFF/FROM copy
            fanin 0; NOT inverts fanin 0; AND/NAND/OR/NOR fold fanins left to
            right with the 2-input function (NAND3 = NAND(NAND(x,y),z)); XOR,
            XNOR and NOT with >= 2 fanins give constant 0 in both frames
            (simkernel.cuh:89-92); a 1-input gate of any type copies fanin 0.
  marks     markkernel.cuh:228-267 -- t(g) and m(fanin 0) and OR over fanin
            positions i of AND over j != i of S(fanin i, fanin j); the reverse
            pass (markkernel.cuh:279-291) keeps a mark only at a PO or below a
            marked fanout.
  first     mergekernel.cuh:141-197 -- earliest pattern marking a node with a
            falling (T0) / rising (T1) transition.
  coverage  coverkernel.cuh:171-230 -- c(g,p) counts marked paths g -> PO that
            contain a node first detected at p; total = sum over PIs.
  segments  K >= 2 (mergekernel.cuh:91-110, coverkernel.cuh:106-154): a
            segment (walk of K nodes along fanouts, shorter only where it
            reaches a node without fanouts) is detected at p when all its
            nodes are marked; its first T0 / T1 pattern follows node 0's
            transition; a node is new at p when some segment holding it was
            first detected at p.
"""
from __future__ import annotations

from functools import lru_cache

UNKN, INPT, AND, NAND, OR, NOR, XOR, XNOR, BUFF, NOT, FROM, DFF = range(12)


def frames(types, fanins, num_pi, v1, v2):
    """v1, v2: PI values of the two vectors.  Returns [(a, b)] per node."""
    out = []
    for g, ty in enumerate(types):
        if ty == INPT:
            out.append((v1[g], v2[g]))
            continue
        ins = [out[f] for f in fanins[g]]
        a, b = ins[0]
        if ty == NOT:
            a, b = 1 - a, 1 - b
        if len(ins) >= 2:
            if ty in (AND, NAND, OR, NOR):
                for xa, xb in ins[1:]:
                    if ty == AND:
                        a, b = a & xa, b & xb
                    elif ty == OR:
                        a, b = a | xa, b | xb
                    elif ty == NAND:
                        a, b = 1 - (a & xa), 1 - (b & xb)
                    else:
                        a, b = 1 - (a | xa), 1 - (b | xb)
            elif ty not in (BUFF, DFF, FROM):
                a, b = 0, 0
        out.append((a, b))
    return out


def sensitizes(robust: bool, ty: int, on, off) -> bool:
    """Off-path input `off` lets a transition on `on` through gate type ty.

    Robust: the on-path input transitions and the off-path input is stable at
    the non-controlling value (1 for AND/NAND, 0 for OR/NOR).  Non-robust, as
    the reference tabulates it: for AND/NAND a falling on-path input needs an
    off-path input whose final value is 1, a rising one needs a stable 1; for
    OR/NOR a rising on-path input needs an off-path final value 0, a falling
    one a stable 0.  XOR/XNOR never sensitize."""
    (a_on, b_on), (a_off, b_off) = on, off
    rising = a_on == 0 and b_on == 1
    falling = a_on == 1 and b_on == 0
    if ty in (AND, NAND):
        if robust:
            return (rising or falling) and a_off == 1 and b_off == 1
        return (falling and b_off == 1) or (rising and a_off == 1 and b_off == 1)
    if ty in (OR, NOR):
        if robust:
            return (rising or falling) and a_off == 0 and b_off == 0
        return (rising and b_off == 0) or (falling and a_off == 0 and b_off == 0)
    return False


def marks_for_pattern(types, fanins, fanouts, po, fr, robust, remove=True):
    """Marks of one pattern; remove=False stops after the forward pass."""
    n = len(types)
    m = [0] * n
    for g in range(n):
        a, b = fr[g]
        t = a ^ b
        ty = types[g]
        if ty == INPT:
            m[g] = t
        elif ty in (BUFF, DFF, FROM, NOT):
            m[g] = t & m[fanins[g][0]]
        elif ty in (AND, NAND, OR, NOR, XOR, XNOR):
            fi = fanins[g]
            ok = any(all(sensitizes(robust, ty, fr[fi[i]], fr[fi[j]]) for j in range(len(fi)) if j != i)
                     for i in range(len(fi)))
            m[g] = int(t and m[fi[0]] and ok)
        else:
            m[g] = 0
    for g in reversed(range(n) if remove else ()):
        if m[g] and not po[g] and not any(m[f] for f in fanouts[g]):
            m[g] = 0
    return m


def suffix_paths(fanouts, po):
    """All paths from each node to a PO (lists of node ids)."""

    @lru_cache(maxsize=None)
    def paths(g):
        res = [(g,)] if po[g] else []
        for f in fanouts[g]:
            res.extend((g,) + rest for rest in paths(f))
        return tuple(res)

    return [paths(g) for g in range(len(po))]


def segment_walks(fanouts, K):
    """Every walk of K nodes along fanouts from each node in id order, fanouts
    in list order, depth first; a walk ending early at a node without fanouts
    is padded with None."""
    out = []

    def walk(path):
        g = path[-1]
        if len(path) == K or not fanouts[g]:
            out.append(tuple(path) + (None,) * (K - len(path)))
            return
        for f in fanouts[g]:
            walk(path + [f])

    for g in range(len(fanouts)):
        walk([g])
    return out


def grade(types, fanins, po, num_pi, vectors_a, vectors_b, robust=True, segs=1):
    """vectors_a/b: list of per-pattern PI value lists (pair p = (a[p], b[p])).
    segs: segment length K.  Returns dict(first_t0, first_t1, gate_counts,
    total, per_pattern) and, for K >= 2, seg_t0 / seg_t1 per segment of
    segment_walks."""
    n = len(types)
    fanouts = [[] for _ in range(n)]
    for g in range(n):
        for f in fanins[g]:
            fanouts[f].append(g)
    P = len(vectors_a)
    fr_all, m_all = [], []
    first0 = [None] * n
    first1 = [None] * n
    for p in range(P):
        fr = frames(types, fanins, num_pi, vectors_a[p], vectors_b[p])
        m = marks_for_pattern(types, fanins, fanouts, po, fr, robust)
        fr_all.append(fr)
        m_all.append(m)
        for g in range(n):
            if m[g]:
                a, b = fr[g]
                if a == 1 and b == 0 and first0[g] is None:
                    first0[g] = p
                if a == 0 and b == 1 and first1[g] is None:
                    first1[g] = p
    walks = segment_walks(fanouts, segs) if segs > 1 else None
    if walks is not None:
        sfirst0 = [None] * len(walks)
        sfirst1 = [None] * len(walks)
        for p in range(P):
            m = m_all[p]
            for s, wk in enumerate(walks):
                if all(m[v] for v in wk if v is not None):
                    a, b = fr_all[p][wk[0]]
                    if b == 0 and sfirst0[s] is None:
                        sfirst0[s] = p
                    if b == 1 and sfirst1[s] is None:
                        sfirst1[s] = p
    sp = suffix_paths(fanouts, po)
    gate_counts = [0] * n
    per_pattern = []
    node_counts = []  # per pattern: c(g, p) of every node
    for p in range(P):
        m = m_all[p]
        if walks is None:
            new = [first0[g] == p or first1[g] == p for g in range(n)]
        else:
            new = [False] * n
            for s, wk in enumerate(walks):
                if sfirst0[s] == p or sfirst1[s] == p:
                    for v in wk:
                        if v is not None:
                            new[v] = True
        tot = 0
        counts = [0] * n
        for g in range(n):
            if not m[g]:
                continue
            cnt = sum(1 for path in sp[g] if all(m[v] for v in path) and any(new[v] for v in path))
            counts[g] = cnt
            gate_counts[g] += cnt
            if types[g] == INPT:
                tot += cnt
        per_pattern.append(tot)
        node_counts.append(counts)
    none = 0xFFFFFFFF
    out = dict(first_t0=[none if x is None else x for x in first0],
               first_t1=[none if x is None else x for x in first1],
               gate_counts=gate_counts, total=sum(per_pattern), per_pattern=per_pattern,
               node_counts=node_counts, frames=fr_all, marks=m_all)
    if walks is not None:
        out["segments"] = [[none if v is None else v for v in wk] for wk in walks]
        out["seg_t0"] = [none if x is None else x for x in sfirst0]
        out["seg_t1"] = [none if x is None else x for x in sfirst1]
    return out


def unpack_sequence(bits, num_pi, P):
    """PI-major packed vectors -> (vectors_a, vectors_b) per pattern, pairing
    V[p] with V[(p+1) mod P]."""
    vals = [[(int(bits[i][p >> 6]) >> (p & 63)) & 1 for i in range(num_pi)] for p in range(P)]
    return vals, [vals[(p + 1) % P] for p in range(P)]
