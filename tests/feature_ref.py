"""Set-based restatement of the 32 swap features computed by k_features
(graphgolf-cuda_amd/csrc/gg_learn.hip), from single-source BFS distances and
the adjacency rows.  Test infrastructure."""
import numpy as np


def bfs_dist(adj, x):
    n = adj.shape[0]
    dist = np.full(n, -1, dtype=np.int64)
    dist[x] = 0
    front = np.array([x])
    l = 0
    while front.size:
        l += 1
        nb = np.unique(adj[front].ravel())
        nb = nb[dist[nb] < 0]
        dist[nb] = l
        front = nb
    return dist


def uniq_second_layer(adj, u, v):
    """U(u,v): entries z of v's row with z != u, z not adjacent to u, and
    exactly one entry of z's row adjacent to u (v itself) -- the second-layer
    vertices of u reached only through v."""
    nu = set(int(x) for x in adj[u])
    cnt = 0
    for z in adj[v]:
        z = int(z)
        if z == u or z in nu:
            continue
        common = sum(1 for t in adj[z] if int(t) in nu or int(t) == u)
        cnt += common == 1
    return cnt


def features(adj, moves, diam):
    n, d = adj.shape
    dist_cache = {}

    def D(x):
        if x not in dist_cache:
            dist_cache[x] = bfs_dist(adj, x)
        return dist_cache[x]

    def R(x, l):
        return D(x) <= l

    def L(x, l):
        return D(x) == l

    out = []
    for a, b, c, dd, kind in moves:
        q1 = dd if kind else c
        q2 = c if kind else dd
        f = np.zeros(32)
        for t, (p, q) in enumerate([(a, q1), (b, q2)]):
            f[t * 8 + 0] = D(p)[q] / diam
            f[t * 8 + 1] = np.sum(L(p, 3) & L(q, 3)) / n
            f[t * 8 + 2] = np.sum(~R(p, 3) & ~R(q, 3)) / n
            f[t * 8 + 3] = np.sum(L(p, 2) & L(q, 2)) / d ** 2
            f[t * 8 + 4] = np.sum(L(q, 2) & ~R(p, 3)) / d ** 2
            f[t * 8 + 5] = np.sum(L(p, 2) & ~R(q, 3)) / d ** 2
            f[t * 8 + 6] = (np.sum(L(q, 1) & ~R(p, 3)) + np.sum(L(p, 1) & ~R(q, 3))) / d
            f[t * 8 + 7] = (np.sum(L(q, 1) & ~R(p, 2)) + np.sum(L(p, 1) & ~R(q, 2))) / d
        for t, (x, y) in enumerate([(a, b), (c, dd)]):
            f[16 + t * 6 + 0] = uniq_second_layer(adj, x, y) / d
            f[16 + t * 6 + 1] = uniq_second_layer(adj, y, x) / d
            f[16 + t * 6 + 2] = np.sum(L(x, 2) & L(y, 2)) / d ** 2
            f[16 + t * 6 + 3] = np.sum(R(x, 2) & R(y, 2)) / d ** 2
            f[16 + t * 6 + 4] = (np.sum(L(x, 1) & L(y, 2)) + np.sum(L(x, 2) & L(y, 1))) / d ** 2
            f[16 + t * 6 + 5] = (D(x).max() + D(y).max() - 2 * diam) * 0.5
        for e, x in enumerate([a, b, c, dd]):
            f[28 + e] = np.sum(~R(x, 3)) / n
        out.append(f)
    return np.array(out)
