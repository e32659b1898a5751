"""Small graph builders with known answers, shared by CPU and GPU tests."""
import itertools

import numpy as np

import gpm_amd as G


def from_edges(n, edges):
    e = np.asarray(list(edges), dtype=np.uint32).reshape(-1, 2)
    return G.csr_from_edges(n, e[:, 0], e[:, 1], G.CSR_CLEAN)


def complete(n):
    return from_edges(n, itertools.combinations(range(n), 2))


def complete_bipartite(a, b):
    return from_edges(a + b, [(i, a + j) for i in range(a) for j in range(b)])


def wheel(n):
    """Hub 0 joined to a cycle on 1..n-1."""
    rim = n - 1
    e = [(0, i) for i in range(1, n)]
    e += [(1 + i, 1 + (i + 1) % rim) for i in range(rim)]
    return from_edges(n, e)


def turan(n, r):
    """Complete r-partite graph with parts as equal as possible."""
    part = [i % r for i in range(n)]
    return from_edges(n, [(i, j) for i, j in itertools.combinations(range(n), 2)
                          if part[i] != part[j]])


def turan_cliques(n, r, k):
    """Number of k-cliques of T(n, r): e_k of the part sizes."""
    sizes = [len(range(p, n, r)) for p in range(r)]
    e = [1] + [0] * k
    for s in sizes:
        for j in range(k, 0, -1):
            e[j] += e[j - 1] * s
    return e[k]


def disjoint_union(a, b):
    """CSR of a + b with b's ids shifted by a.n."""
    off = np.concatenate([a.offsets, b.offsets[1:] + a.offsets[-1]]).astype(np.uint32)
    adj = np.concatenate([a.adj, b.adj + a.n]).astype(np.uint32)
    return G.HostCSR(off, adj)


def er(n, m, seed):
    return G.er_graph(n, m, seed)


def to_nx(g):
    import networkx as nx
    x = nx.Graph()
    x.add_nodes_from(range(g.n))
    for u in range(g.n):
        for v in g.adj[g.offsets[u]:g.offsets[u + 1]]:
            if u < v:
                x.add_edge(u, int(v))
    return x


def adjacency_matrix(g):
    A = np.zeros((g.n, g.n), dtype=np.uint8)
    for u in range(g.n):
        A[u, g.adj[g.offsets[u]:g.offsets[u + 1]]] = 1
    return A


def write_labeled(path, g, labels):
    """Writes g (symmetric CSR) in the sub_ceci text format ("t", "v", "e")."""
    deg = g.degrees()
    with open(path, "w") as f:
        f.write(f"t {g.n} {g.m // 2}\n")
        for v in range(g.n):
            f.write(f"v {v} {int(labels[v])} {int(deg[v])}\n")
        for v in range(g.n):
            for w in g.adj[g.offsets[v]:g.offsets[v + 1]]:
                if v < w:
                    f.write(f"e {v} {int(w)}\n")
