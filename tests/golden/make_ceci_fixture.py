"""Packs the reference's CECI known-answer data into tests/golden/ceci_hprd.npz.

Inputs (data files held by the reference's own test, read as text):
  sub_ceci/test/data_graph/HPRD.graph      labeled data graph ("t/v/e" format)
  sub_ceci/test/query_graph/*.graph        200 dense 16-vertex queries
  sub_ceci/test/expected_output.res        "#Embeddings" per query
Run in a container that has the reference checkout:
  python tests/golden/make_ceci_fixture.py /root/reference
"""
import os
import sys

import numpy as np


def read_tve(path):
    """Parses a 't n m / v id label deg / e u v' file into (labels, edges)."""
    labels, edges = {}, []
    with open(path) as f:
        for line in f:
            p = line.split()
            if not p:
                continue
            if p[0] == "v":
                labels[int(p[1])] = int(p[2])
            elif p[0] == "e":
                edges.append((int(p[1]), int(p[2])))
    n = max(labels) + 1 if labels else 0
    lab = np.zeros(n, dtype=np.uint32)
    for v, l in labels.items():
        lab[v] = l
    return lab, np.asarray(edges, dtype=np.uint32).reshape(-1, 2)


def main(ref):
    base = os.path.join(ref, "sub_ceci", "test")
    dlab, dedges = read_tve(os.path.join(base, "data_graph", "HPRD.graph"))
    expected = {}
    with open(os.path.join(base, "expected_output.res")) as f:
        for line in f:
            if ":" in line:
                k, v = line.split(":")
                expected[k.strip()] = int(v.strip())
    names = sorted(expected, key=lambda s: int(s.rsplit("_", 1)[1]))
    q_n, q_m, q_edges, q_labels, q_count = [], [], [], [], []
    for name in names:
        lab, e = read_tve(os.path.join(base, "query_graph", name + ".graph"))
        q_n.append(len(lab))
        q_m.append(len(e))
        q_edges.append(e.reshape(-1))
        q_labels.append(lab)
        q_count.append(expected[name])
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "ceci_hprd.npz")
    np.savez_compressed(
        out,
        data_n=np.uint32(len(dlab)), data_labels=dlab, data_edges=dedges,
        names=np.asarray(names), q_n=np.asarray(q_n, dtype=np.uint32),
        q_m=np.asarray(q_m, dtype=np.uint32),
        q_edges=np.concatenate(q_edges).astype(np.uint32),
        q_labels=np.concatenate(q_labels).astype(np.uint32),
        expected=np.asarray(q_count, dtype=np.uint64))
    print("wrote", out, len(names), "queries;", len(dlab), "data vertices,",
          len(dedges), "edges")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
