"""Closed-form test graphs (explicit adjacency, int32 [N, d], -1 padding) and
their known distance totals / diameters.  Pure numpy; test infrastructure."""
from __future__ import annotations

import itertools

import numpy as np


def cycle(n):
    v = np.arange(n)
    return np.stack([(v - 1) % n, (v + 1) % n], axis=1).astype(np.int32)


def cycle_total(n):  # sum over ordered pairs: n * floor(n^2/4)
    return n * (n * n // 4)


def complete(n):
    return np.array([[u for u in range(n) if u != v] for v in range(n)], dtype=np.int32)


def hypercube(k):
    n = 1 << k
    return np.array([[v ^ (1 << b) for b in range(k)] for v in range(n)], dtype=np.int32)


def hypercube_total(k):  # 2^k * k * 2^(k-1)
    return (1 << k) * k * (1 << (k - 1))


def petersen():
    edges = [(i, (i + 1) % 5) for i in range(5)] + [(i, i + 5) for i in range(5)] + \
            [(5 + i, 5 + (i + 2) % 5) for i in range(5)]
    adj = [[] for _ in range(10)]
    for a, b in edges:
        adj[a].append(b)
        adj[b].append(a)
    return np.array(adj, dtype=np.int32)


def kdd(d):  # complete bipartite K_{d,d}: total 2d(3d-2), diameter 2
    n = 2 * d
    return np.array([[u for u in range(n) if (u < d) != (v < d)] for v in range(n)],
                    dtype=np.int32)


def circulant(n, jumps):
    v = np.arange(n)
    cols = []
    for j in jumps:
        cols.append((v + j) % n)
        cols.append((v - j) % n)
    return np.stack(cols, axis=1).astype(np.int32)


def part_of_adj(adj):
    """Offset table (M = N) of an explicit adjacency: off = u - v."""
    n = adj.shape[0]
    return (adj - np.arange(n)[:, None]).astype(np.int32)


def disjoint_cycles(n1, n2):
    a = cycle(n1)
    b = cycle(n2) + n1
    return np.concatenate([a, b]).astype(np.int32)


def bridged(A, B):
    """Disjoint union of two regular graphs joined by one 2-opt swap across
    them: edges (a,b) of A and (c,d) of B become the bridges (a,c), (b,d).
    Returns the adjacency and the two bridges."""
    A = np.asarray(A, dtype=np.int32)
    B = np.asarray(B, dtype=np.int32) + A.shape[0]
    adj = np.concatenate([A, B]).copy()
    a, c = 0, A.shape[0]
    b, d = int(adj[a, 0]), int(adj[c, 0])
    for v, old, new in ((a, b, c), (b, a, d), (c, d, a), (d, c, b)):
        adj[v, int(np.nonzero(adj[v] == old)[0][0])] = new
    return adj, ((a, c), (b, d))
