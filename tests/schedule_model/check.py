"""Checker for the schedule model's trace (fake_hip.cpp writes it).

The trace lists, in the order the host issued them, the engine's device
work (kernels, copies, event records and waits, each on its queue) and
what the host learnt about it (synchronisations, answered queries, frees).
From it this builds the happens-before order of the device work:

  * a queue runs its nodes in order;
  * a wait holds its queue until the record it names has finished;
  * work enqueued after the host learnt that node X finished runs after X.

Every edge goes from an earlier trace line to a later one (a wait names a
record already issued), so the order has no cycles and a node's ancestors
are known when its line is read. The checks:

  races        two nodes that touch one allocation, at least one writing,
               with neither ordered before the other;
  host reuse   a copy whose source is pinned host memory the host has just
               rewritten, while an earlier node that read or wrote that
               memory may still run;
  frees        an allocation freed while a node that touches it may run;
  residency    persistent launches (workgroups that wait on each other, so
               all must be resident at once) that can run at the same time
               and together need more than the whole device: the largest
               total share over any set of pairwise unordered persistent
               launches (a set of pairwise unordered nodes can all run at
               once: run their ancestors first) must stay at most 1.
"""
from dataclasses import dataclass, field


@dataclass
class Report:
    nodes: int = 0
    races: list = field(default_factory=list)
    host_reuse: list = field(default_factory=list)
    frees: list = field(default_factory=list)
    persistent: int = 0
    max_concurrent_share: float = 0.0
    max_concurrent: list = field(default_factory=list)
    dangling_waits: int = 0

    def ok(self):
        return not (self.races or self.host_reuse or self.frees) and self.max_concurrent_share <= 1.0 + 1e-9


def _max_share_antichain(pers, anc):
    """Largest total share over sets of pairwise unordered persistent nodes.
    pers: list of (id, share). Exhaustive over the comparability structure
    with pruning; the sets that matter are small (a few launches run at once)."""
    ids = [i for i, _ in pers]
    share = dict(pers)
    unordered = {a: {b for b in ids if b != a and not (anc[a] >> b) & 1 and not (anc[b] >> a) & 1}
                 for a in ids}
    best = (0.0, [])

    def grow(chosen, total, cands):
        nonlocal best
        if total > best[0]:
            best = (total, list(chosen))
        for i, c in enumerate(cands):
            rest = [d for d in cands[i + 1:] if d in unordered[c]]
            if total + share[c] + sum(share[d] for d in rest) <= best[0]:
                continue
            chosen.append(c)
            grow(chosen, total + share[c], rest)
            chosen.pop()

    grow([], 0.0, ids)
    return best


def analyse(path):
    rep = Report()
    anc = {}                 # node -> bitset of its ancestors
    tail = {}                # queue -> last node
    host = 0                 # bitset: nodes the host knows have finished
    acc = {}                 # allocation -> list of (node, mode) since its last free
    info = {}                # node -> "name@queue" for messages
    pers = []
    note = ""
    with open(path) as f:
        for line in f:
            t = line.split()
            if not t:
                continue
            kind = t[0]
            if kind == "M":
                note = line[2:].strip()
                continue
            if kind == "H":
                for d in map(int, t[2:]):
                    host |= anc[d] | (1 << d)
                continue
            if kind == "F":
                key = t[1]
                for n, _ in acc.get(key, []):
                    if not (host >> n) & 1:
                        rep.frees.append(f"{key} freed while {info[n]} may run ({note})")
                acc.pop(key, None)
                continue
            nid, q = int(t[1]), int(t[2])
            a = host
            if q in tail:
                a |= anc[tail[q]] | (1 << tail[q])
            rest = t[3:]
            name = {"K": None, "C": "copy", "R": "record", "W": "wait"}[kind]
            share = 0.0
            if kind == "K":
                name, share, rest = rest[0], float(rest[1]), rest[2:]
            elif kind == "W":
                rec = int(rest[0])
                if rec >= 0:
                    a |= anc[rec] | (1 << rec)
                else:
                    rep.dangling_waits += 1
                rest = []
            elif kind == "R":
                rest = []
            anc[nid] = a
            tail[q] = nid
            info[nid] = f"{name}#{nid}@q{q} [{note}]"
            rep.nodes += 1
            if share > 0:
                pers.append((nid, share))
            for tok in rest:
                mode, key = tok.split(":")
                lst = acc.setdefault(key, [])
                if kind == "C" and mode == "r":
                    # the host filled this pinned buffer just before the copy
                    for n, _ in lst:
                        if not (host >> n) & 1:
                            rep.host_reuse.append(f"host rewrote {key} for {info[nid]} while {info[n]} may run")
                for n, m in lst:
                    if (m == "w" or mode == "w") and not (a >> n) & 1:
                        rep.races.append(f"{key}: {info[n]} {m} / {info[nid]} {mode}")
                lst.append((nid, mode))
    rep.persistent = len(pers)
    if pers:
        rep.max_concurrent_share, ids = _max_share_antichain(pers, anc)
        rep.max_concurrent = [info[i] for i in ids]
    return rep
