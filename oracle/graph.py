"""Independent check of the oracle's DP: generic HMM algorithms over an explicit state graph.

TEST INFRASTRUCTURE ONLY (pure Python, small M and L).

The Plan7 local/multihit architecture (and the MSV filter's ungapped
architecture) is written down as an explicit list of states and weighted
edges, following the model description in sentinel-x src/hmmer/model.rs:1-18
and 73-113 (t_xx[k] is the edge out of node k, free local exits, entry B->M_k,
N/J/C emit on their self loops) and dp.rs:67-68 (MSV: uniform entry, M_k ->
M_k+1 free, nu = 2).  A generic algorithm that knows nothing of Plan7 then
computes Viterbi (max), Forward (logsumexp) and Backward over that graph.
It shares no recurrence code with sx_oracle.c, so a dropped term, a wrong
index or a transposed parameter in the oracle shows up as a mismatch.
``enumerate_paths`` brute-forces every state path for the tiniest cases and
pins the generic algorithm itself.
"""
from __future__ import annotations

import math
from typing import Dict, List, Tuple

NEG = -math.inf


def _lse(vals):
    vals = [v for v in vals if v != NEG]
    if not vals:
        return NEG
    mx = max(vals)
    return mx + math.log(sum(math.exp(v - mx) for v in vals))


class Graph:
    """States are either emitting (consume one residue when entered) or silent."""

    def __init__(self):
        self.names: List[str] = []
        self.emit: Dict[int, object] = {}     # state -> function(x) -> log score
        self.edges: List[Tuple[int, int, float]] = []
        self.index: Dict[str, int] = {}

    def add(self, name, emit=None):
        i = len(self.names)
        self.names.append(name)
        self.index[name] = i
        if emit is not None:
            self.emit[i] = emit
        return i

    def edge(self, a, b, w):
        if w != NEG:
            self.edges.append((self.index[a], self.index[b], float(w)))

    def _silent_order(self):
        silent = [i for i in range(len(self.names)) if i not in self.emit]
        deps = {i: set() for i in silent}
        for a, b, _ in self.edges:
            if b in deps and a in deps:
                deps[b].add(a)
        order, done = [], set()
        while len(order) < len(silent):
            progressed = False
            for s in silent:
                if s not in done and deps[s] <= done:
                    order.append(s)
                    done.add(s)
                    progressed = True
            if not progressed:
                raise ValueError("silent cycle")
        return order

    def run(self, dsq, start, end, combine):
        """Forward-type pass; returns (score at `end` after the last residue, table)."""
        n = len(self.names)
        order = self._silent_order()
        incoming = {i: [] for i in range(n)}
        for a, b, w in self.edges:
            incoming[b].append((a, w))
        L = len(dsq)
        tab = [[NEG] * n for _ in range(L + 1)]
        s0 = self.index[start]
        tab[0][s0] = 0.0
        for s in order:
            if s == s0:
                continue
            tab[0][s] = combine([tab[0][a] + w for a, w in incoming[s]])
        for i in range(1, L + 1):
            x = dsq[i - 1]
            for s, e in self.emit.items():
                es = e(x)
                tab[i][s] = combine([tab[i - 1][a] + w for a, w in incoming[s]]) + es if es != NEG else NEG
            for s in order:
                if s == s0:
                    continue
                tab[i][s] = combine([tab[i][a] + w for a, w in incoming[s]])
        return tab[L][self.index[end]], tab

    def backward(self, dsq, end):
        """Generic backward: tab[i][s] = log-sum over completions from state s at position i
        (s already entered; an emitting s has consumed residue i) to `end` after residue L."""
        n = len(self.names)
        order = self._silent_order()
        outgoing = {i: [] for i in range(n)}
        for a, b, w in self.edges:
            outgoing[a].append((b, w))
        L = len(dsq)
        e_idx = self.index[end]
        emitting = [s for s in range(n) if s in self.emit]
        tab = [[NEG] * n for _ in range(L + 1)]
        for i in range(L, -1, -1):
            for s in list(reversed(order)) + emitting:
                vals = [0.0] if (i == L and s == e_idx) else []
                for b, w in outgoing[s]:
                    if b in self.emit:
                        if i < L:
                            es = self.emit[b](dsq[i])
                            if es != NEG:
                                vals.append(w + es + tab[i + 1][b])
                    else:
                        vals.append(w + tab[i][b])
                tab[i][s] = _lse(vals)
        return tab

    def enumerate_paths(self, dsq, start, end, combine):
        """Brute force over every state path (tiny cases only)."""
        outgoing = {i: [] for i in range(len(self.names))}
        for a, b, w in self.edges:
            outgoing[a].append((b, w))
        L = len(dsq)
        s0, e0 = self.index[start], self.index[end]
        scores = []

        def dfs(s, i, acc, depth):
            if depth > 4 * (L + 2) * (len(self.names) + 2):
                return
            if i == L and s == e0:
                scores.append(acc)
            for b, w in outgoing[s]:
                if b in self.emit:
                    if i < L:
                        es = self.emit[b](dsq[i])
                        if es != NEG:
                            dfs(b, i + 1, acc + w + es, depth + 1)
                else:
                    dfs(b, i, acc + w, depth + 1)

        dfs(s0, 0, 0.0, 0)
        return combine(scores) if scores else NEG


def plan7_graph(m, msc, tsc, xs, star=24):
    """Explicit Plan7 local multihit graph.  msc (25, M+1), tsc rows MM MI MD IM II DM DD BM,
    xs: n_loop n_move j_loop j_move c_loop c_move e_move e_loop (log space)."""
    g = Graph()
    g.add("S")
    g.add("N")
    g.add("N'", emit=lambda x: 0.0)
    g.add("B")
    for k in range(1, m + 1):
        g.add(f"M{k}", emit=(lambda k: lambda x: float(msc[x][k]))(k))
    for k in range(1, m):
        g.add(f"I{k}", emit=lambda x: NEG if x == star else 0.0)
    for k in range(2, m + 1):
        g.add(f"D{k}")
    g.add("E")
    g.add("J")
    g.add("J'", emit=lambda x: 0.0)
    g.add("C")
    g.add("C'", emit=lambda x: 0.0)
    g.add("T")
    MM, MI, MD, IM, II, DM, DD, BM = range(8)
    n_loop, n_move, j_loop, j_move, c_loop, c_move, e_move, e_loop = [float(v) for v in xs[:8]]
    g.edge("S", "N", 0.0)
    g.edge("N", "N'", n_loop)
    g.edge("N'", "N", 0.0)
    g.edge("N", "B", n_move)
    g.edge("J", "B", j_move)
    for k in range(1, m + 1):
        g.edge("B", f"M{k}", float(tsc[BM][k]))
        g.edge(f"M{k}", "E", 0.0)
        if k >= 2:
            g.edge(f"D{k}", "E", 0.0)
    for k in range(1, m):
        # edges out of node k to node k+1
        g.edge(f"M{k}", f"M{k+1}", float(tsc[MM][k]))
        g.edge(f"I{k}", f"M{k+1}", float(tsc[IM][k]))
        if k >= 2:
            g.edge(f"D{k}", f"M{k+1}", float(tsc[DM][k]))
            g.edge(f"D{k}", f"D{k+1}", float(tsc[DD][k]))
        g.edge(f"M{k}", f"D{k+1}", float(tsc[MD][k]))
        g.edge(f"M{k}", f"I{k}", float(tsc[MI][k]))
        g.edge(f"I{k}", f"I{k}", float(tsc[II][k]))
    g.edge("E", "J", e_loop)
    g.edge("E", "C", e_move)
    g.edge("J", "J'", j_loop)
    g.edge("J'", "J", 0.0)
    g.edge("C", "C'", c_loop)
    g.edge("C'", "C", 0.0)
    g.edge("C", "T", c_move)
    return g


def msv_graph(m, msc, msv_tbm, L):
    """Explicit MSV filter graph (ungapped, uniform entry, multihit, nu = 2)."""
    lf = float(L)
    tloop = math.log(lf / (lf + 3.0))
    tmove = math.log(3.0 / (lf + 3.0))
    g = Graph()
    for n in ("S", "N"):
        g.add(n)
    g.add("N'", emit=lambda x: 0.0)
    g.add("B")
    for k in range(1, m + 1):
        g.add(f"M{k}", emit=(lambda k: lambda x: float(msc[x][k]))(k))
    g.add("E")
    g.add("J")
    g.add("J'", emit=lambda x: 0.0)
    g.add("C")
    g.add("C'", emit=lambda x: 0.0)
    g.add("T")
    g.edge("S", "N", 0.0)
    g.edge("N", "N'", tloop)
    g.edge("N'", "N", 0.0)
    g.edge("N", "B", tmove)
    g.edge("J", "B", tmove)
    for k in range(1, m + 1):
        g.edge("B", f"M{k}", float(msv_tbm))
        g.edge(f"M{k}", "E", 0.0)
        if k < m:
            g.edge(f"M{k}", f"M{k+1}", 0.0)
    g.edge("E", "J", -math.log(2.0))
    g.edge("E", "C", -math.log(2.0))
    g.edge("J", "J'", tloop)
    g.edge("J'", "J", 0.0)
    g.edge("C", "C'", tloop)
    g.edge("C'", "C", 0.0)
    g.edge("C", "T", tmove)
    return g


def two_state_bias(compo, bg, m, dsq):
    """Log-space forward of the two-state composition HMM (p7_bg_FilterScore model):
    state 0 emits background, state 1 emits the model composition; transitions
    from L0 = 400 and L1 = M/8; start (0.999, 0.001); end from either state.
    Returns the filter score minus nothing (log odds of the two-state model vs
    background, plus the null length term), in nats."""
    l0, l1 = 400.0, m / 8.0
    lt = [[math.log(l0 / (l0 + 1)), math.log(1 / (l0 + 1))], [math.log(1 / (l1 + 1)), math.log(l1 / (l1 + 1))]]
    members = {20: [2, 11], 21: [7, 9], 22: [3, 13], 23: list(range(20))}

    def lemit(x, s):
        if x == 24:
            return 0.0
        if x < 20:
            return 0.0 if s == 0 else math.log(compo[x] / bg[x])
        mem = members[x]
        z = sum(bg[y] for y in mem)
        if s == 0:
            return math.log(sum(bg[y] for y in mem) / z)
        return math.log(sum(compo[y] for y in mem) / z)

    a = [math.log(0.999) + lemit(dsq[0], 0), math.log(0.001) + lemit(dsq[0], 1)]
    for x in dsq[1:]:
        a = [_lse([a[0] + lt[0][s], a[1] + lt[1][s]]) + lemit(x, s) for s in (0, 1)]
    L = len(dsq)
    p1 = L / (L + 1.0)
    return _lse(a) + L * math.log(p1) + math.log(1 - p1)
