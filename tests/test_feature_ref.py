"""CPU checks of the feature restatement's loss count U(u,v) (second-layer
vertices of u reached only through v) on graphs where it has a closed form."""
import numpy as np
import pytest

import feature_ref as fr
import graphs as G


def _edges(adj):
    return [(u, int(v)) for u in range(adj.shape[0]) for v in adj[u]]


@pytest.mark.parametrize("n", [5, 8, 31])
def test_cycle_every_edge_has_one_private_second_neighbour(n):
    adj = G.cycle(n)
    assert all(fr.uniq_second_layer(adj, u, v) == 1 for u, v in _edges(adj))


@pytest.mark.parametrize("k", [2, 3, 5])
def test_hypercube_has_none(k):
    # z in N(v) \ {u} differs from u in two bits: two common neighbours
    adj = G.hypercube(k)
    assert all(fr.uniq_second_layer(adj, u, v) == 0 for u, v in _edges(adj))


def test_petersen_girth_five_gives_degree_minus_one():
    adj = G.petersen()
    assert all(fr.uniq_second_layer(adj, u, v) == 2 for u, v in _edges(adj))


@pytest.mark.parametrize("d", [2, 4])
def test_complete_bipartite_has_none(d):
    adj = G.kdd(d)
    assert all(fr.uniq_second_layer(adj, u, v) == 0 for u, v in _edges(adj))


def test_complete_graph_has_none():
    adj = G.complete(6)  # every z is adjacent to u: no second layer at all
    assert all(fr.uniq_second_layer(adj, u, v) == 0 for u, v in _edges(adj))


def test_circulant_matches_brute_force():
    # brute force from the distance matrix: z in N(v), d(u,z) = 2 and exactly
    # one common neighbour of u and z
    rng = np.random.default_rng(3)
    adj = G.circulant(40, [1, 7])
    n = adj.shape[0]
    D = np.array([fr.bfs_dist(adj, x) for x in range(n)])
    for _ in range(30):
        u = int(rng.integers(0, n))
        v = int(adj[u][rng.integers(0, adj.shape[1])])
        want = 0
        for z in adj[v]:
            z = int(z)
            if D[u, z] != 2:
                continue
            common = int(np.sum((D[u] == 1) & (D[z] == 1)))
            want += common == 1
        assert fr.uniq_second_layer(adj, u, v) == want
