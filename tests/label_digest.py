"""Compact fingerprints of large label sets (test infrastructure).

A full label set at R-MAT scale 18 holds about 176 million labels: too big to
commit, so the fixtures hold, per block of consecutive vertices, the label
count and a 64-bit digest of the block's sorted lists.  The digest of a block
is the wrapping sum over its labels of mix(v, i, hub, hop, dist), where i is
the label's position in L(v); a different label, a missing one or one out of
place changes the sum unless mix values collide (probability about 2^-64).

Both sides (the oracle when the fixture is written, the HIP library's export
in the GPU test) go through this same function, so it pins nothing by itself:
it only compares two label sets that were computed independently.
"""
import numpy as np

_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)
_G = np.uint64(0x9E3779B97F4A7C15)


def _mix(x):
    """splitmix64 finaliser, elementwise on uint64 (wrapping)."""
    with np.errstate(over="ignore"):
        x = x + _G
        x = (x ^ (x >> np.uint64(30))) * _M1
        x = (x ^ (x >> np.uint64(27))) * _M2
        return x ^ (x >> np.uint64(31))


def block_digests(off, hub, hop, dist, block=1024, chunk=1 << 24):
    """Per block of `block` vertices: (label counts u64[nb], digests u64[nb]).

    off: int64/uint64 [n+1]; hub u32, hop u8, dist i32 [total], each vertex's
    labels sorted by (hub, hop, dist)."""
    off = np.asarray(off, dtype=np.int64)
    n = len(off) - 1
    nb = (n + block - 1) // block
    bstart = off[np.minimum(np.arange(nb + 1) * block, n)]
    counts = np.diff(bstart).astype(np.uint64)
    dig = np.zeros(nb, dtype=np.uint64)
    total = int(off[-1])
    # labels in chunks so peak memory stays a few GB at any size
    for c0 in range(0, total, chunk):
        c1 = min(total, c0 + chunk)
        pos = np.arange(c0, c1, dtype=np.int64)
        v = np.searchsorted(off, pos, side="right") - 1
        i = pos - off[v]
        key = (np.asarray(hub[c0:c1], dtype=np.uint64)
               | (np.asarray(hop[c0:c1], dtype=np.uint64) << np.uint64(32)))
        d = np.asarray(dist[c0:c1], dtype=np.int64).astype(np.uint64) & np.uint64(0xFFFFFFFF)
        a = _mix((v.astype(np.uint64) << np.uint64(32)) | i.astype(np.uint64))
        with np.errstate(over="ignore"):
            h = _mix(_mix(a ^ key) + d)
        b = (v // block).astype(np.int64)
        dig = _wrap_add(dig, _exact_block_sums(b, h, nb))
    return counts, dig


def _exact_block_sums(b, h, nb):
    """Wrapping uint64 sum of h per block b (b non-decreasing)."""
    # b is non-decreasing within a chunk (labels are vertex-ordered), so a
    # cumulative wrapping sum and differences at block boundaries are exact
    with np.errstate(over="ignore"):
        cs = np.cumsum(h, dtype=np.uint64)
    out = np.zeros(nb, dtype=np.uint64)
    if len(b) == 0:
        return out
    ends = np.flatnonzero(np.diff(b)) if len(b) > 1 else np.zeros(0, np.int64)
    last = np.append(ends, len(b) - 1)
    blocks = b[last]
    tot = cs[last]
    prev = np.concatenate([[np.uint64(0)], tot[:-1]]).astype(np.uint64)
    with np.errstate(over="ignore"):
        out[blocks] = tot - prev
    return out


def _wrap_add(a, b):
    with np.errstate(over="ignore"):
        return (a + b).astype(np.uint64)
