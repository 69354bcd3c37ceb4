"""Small seeded test graphs (inputs only).

The reference's algorithm tests use Graphs.jl's barabasi_albert(30, 3) and
dorogovtsev_mendes(30) under Random.seed!(1234) (test/bfs.jl:10-12,
test/runtests.jl:19).  Julia's RNG streams cannot be reproduced here, so the
same generator families are drawn with numpy seeds instead.
"""
import numpy as np
import scipy.sparse as sp


def barabasi_albert(n, k, seed):
    """Preferential attachment: each new vertex links to k distinct existing ones."""
    import networkx as nx
    g = nx.barabasi_albert_graph(n, k, seed=seed)
    return sorted((min(u, v), max(u, v)) for u, v in g.edges())


def dorogovtsev_mendes(n, seed):
    """Dorogovtsev-Mendes process: start from a triangle; each new vertex picks
    a uniformly random existing edge and links to both of its ends."""
    rng = np.random.default_rng(seed)
    edges = [(0, 1), (1, 2), (0, 2)]
    for v in range(3, n):
        a, b = edges[rng.integers(len(edges))]
        edges.append((a, v))
        edges.append((b, v))
    return sorted(set((min(u, v), max(u, v)) for u, v in edges))


def undirected_adj(n, edges, dtype=np.bool_, weights=None):
    """Symmetric adjacency (scipy CSR).  weights: per undirected edge, or None."""
    u = np.array([e[0] for e in edges], np.int64)
    v = np.array([e[1] for e in edges], np.int64)
    if weights is None:
        w = np.ones(len(edges))
    else:
        w = np.asarray(weights, np.float64)
    S = sp.coo_matrix((np.concatenate([w, w]), (np.concatenate([u, v]), np.concatenate([v, u]))),
                      shape=(n, n)).tocsr()
    S.sum_duplicates()
    S.sort_indices()
    return S.astype(dtype)


def random_digraph(rng, n, p, wmax=None, dtype=np.float32):
    """Directed G(n, p) without self-loops; integer weights in 1..wmax if given."""
    M = rng.random((n, n)) < p
    np.fill_diagonal(M, False)
    S = sp.csr_matrix(M.astype(np.float64))
    if wmax is not None:
        S.data = rng.integers(1, wmax + 1, S.nnz).astype(np.float64)
    S.sort_indices()
    return S.astype(dtype)
