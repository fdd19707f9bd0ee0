"""Vertex groups for group builds, restated in plain Python (test
infrastructure only: the checker of hsdl_graph_groups).

The reference's generators (Hybrid_Generation_Clean/include/vgroup/):
  * CDLP_group.cuh:20-107: label propagation with a community size cap
    (CDLP/GPU_Community_Detection.cuh:104-210), then a greedy merge of the
    communities up to MAX_GROUP_SIZE (:54-80) and an order by total degree,
    largest first (:82-97);
  * Random_group.cuh:10-33: a shuffle cut every MAX_GROUP_SIZE.
The reference's size test races between GPU threads; the library's rule,
restated here, is deterministic: each round every vertex takes the label
most frequent among its neighbours' labels and its own (one vote; ties: the
smallest) among labels whose community held fewer than max_size vertices
after the previous round, or its own; rounds stop once none changes.
Communities are cut to max_size in id order before the merge.
"""


def propagate(off, col, max_size, rounds):
    n = len(off) - 1
    label = list(range(n))
    size = [1] * n
    if off[n] == 0:
        return label
    for _ in range(rounds):
        nxt = []
        changed = 0
        for v in range(n):
            labs = sorted(label[col[e]] for e in range(off[v], off[v + 1]))
            own = label[v]
            votes = {}
            for l in labs:
                votes[l] = votes.get(l, 0) + 1
            votes[own] = votes.get(own, 0) + 1
            best, best_cnt = None, 0
            for l in sorted(votes):
                if votes[l] > best_cnt and (l == own or size[l] < max_size):
                    best, best_cnt = l, votes[l]
            nxt.append(best)
            changed += best != own
        label = nxt
        size = [0] * n
        for l in label:
            size[l] += 1
        if changed == 0:
            break
    return label


def merge(groups, max_size):
    """CDLP_group.cuh:54-80: take the largest (size, id); join it with the
    largest (size, id) of size <= max_size - its size (the one ordered just
    below it when that is itself); a group with no partner is final."""
    groups = [list(g) for g in groups]
    order = sorted((len(g), i) for i, g in enumerate(groups))
    out = []
    while order:
        s1, a = order[-1]
        room = max_size - s1
        k = max((idx for idx, (s, _) in enumerate(order) if s <= room), default=None)
        if k is not None and k == len(order) - 1:
            k = k - 1 if k > 0 else None
        if k is not None and s1 + order[k][0] <= max_size:
            b = order[k][1]
            groups[a] += groups[b]
            del order[-1]
            del order[k]
            order.append((len(groups[a]), a))
            order.sort()
        else:
            out.append(groups[a])
            del order[-1]
    return out


def cdlp_groups(off, col, max_size, rounds):
    n = len(off) - 1
    label = propagate(off, col, max_size, rounds)
    first, comm = {}, []
    for v in range(n):
        if label[v] not in first:
            first[label[v]] = len(comm)
            comm.append([])
        comm[first[label[v]]].append(v)
    cut = [c[i:i + max_size] for c in comm for i in range(0, len(c), max_size)]
    merged = merge(cut, max_size)
    key = sorted((-sum(off[v + 1] - off[v] for v in g), i) for i, g in enumerate(merged))
    return [sorted(merged[i]) for _, i in key]


def group_of(groups, n):
    out = [None] * n
    for i, g in enumerate(groups):
        for v in g:
            out[v] = i
    return out
