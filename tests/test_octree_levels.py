"""CPU pin of the all-levels octree build's node characterisation.

With L_i the number of leading octal digits sorted key i shares with key
i-1 (L_0 = L_n = -1) and D_i = min over a <= i < b, b - a <= leaf_cap of
max(L_a, L_b) + 1 (clamped to 1..21), the nodes are exactly the pairs
(i, d) with L_i < d <= D_i, level d ordered by first particle i, internal
when d < D_i with first child (i, d + 1).  Restated here in plain Python and
compared with the oracle's level-by-level tree (the reference's algorithm,
src/linear_octree.c:342-823) on small inputs, repeated positions included."""
import numpy as np
import pytest

from grav_sim_amd import ic
from oracle import oracle as O

MAXL = 21


def _shared(a, b):
    x = int(a) ^ int(b)
    if x == 0:
        return MAXL
    return MAXL - 1 - (x.bit_length() - 1) // 3


def _nodes(sk, cap):
    n = len(sk)
    L = [-1] + [_shared(sk[i - 1], sk[i]) for i in range(1, n)] + [-1]
    D = []
    for i in range(n):
        best = min(max(L[a], L[b]) + 1 for a in range(max(0, i - cap + 1), i + 1)
                   for b in range(i + 1, min(n, a + cap) + 1))
        D.append(min(max(best, 1), MAXL))
    order = [(d, i) for d in range(MAXL + 1) for i in range(n) if L[i] < d <= D[i]]
    index = {nd: k for k, nd in enumerate(order)}
    out = {k: [] for k in ("depth", "first", "count", "nchild", "child0")}
    for d, i in order:
        sh = 3 * (MAXL - d)
        j = i
        while j < n and (int(sk[j]) >> sh) == (int(sk[i]) >> sh):
            j += 1
        internal = d < D[i]
        nch = len({int(sk[t]) >> (sh - 3) for t in range(i, j)}) if internal else 0
        for k, v in zip(out, (d, i, j - i, nch, index[(d + 1, i)] if internal else -1)):
            out[k].append(v)
    return {k: np.array(v) for k, v in out.items()}


def _inputs():
    rng = np.random.default_rng(3)
    yield rng.standard_normal((1, 3))
    yield rng.standard_normal((2, 3))
    yield rng.standard_normal((37, 3))
    x = rng.standard_normal((160, 3))
    x[20:45] = x[20]  # equal keys down to depth 21
    x[90:92] = x[3]
    yield x
    yield ic.plummer(400, 5)[0]


@pytest.mark.parametrize("cap", [1, 2, 5, 16])
def test_node_set_from_shared_digits_matches_oracle(cap):
    for x in _inputs():
        m = np.full(len(x), 1.0 / len(x))
        t = O.build_tree(x, m, cap)
        f = _nodes(t["sorted_keys"], cap)
        for k in f:
            assert np.array_equal(f[k], t[k]), (len(x), cap, k)
