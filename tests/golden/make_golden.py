"""Writes tests/golden/induced_counts.json: brute-force induced subgraph-isomorphism counts.

Small random labelled graphs; counts come from networkx's VF2 GraphMatcher
(subgraph_isomorphisms_iter enumerates node-induced embeddings) with a label-equality
node_match.  For refinement depth K <= 2 every filter of the reference path is sound, so the
pipeline's count-all result must equal these counts exactly; for K >= 3 it may only be smaller
(the distance-exactly-k shell test is unsound there).

Run:  python tests/golden/make_golden.py   (needs networkx; the fixture is committed)
"""
import json
import random
from pathlib import Path

import networkx as nx
from networkx.algorithms import isomorphism as iso

OUT = Path(__file__).resolve().parent / "induced_counts.json"


def random_graph(rng, n, p_extra, labels, max_deg=None):
    """Random connected graph: a random tree plus extra edges with probability p_extra."""
    g = nx.Graph()
    for i in range(n):
        g.add_node(i, label=rng.choice(labels))
    for i in range(1, n):
        cands = [j for j in range(i) if max_deg is None or g.degree(j) < max_deg]
        if not cands:
            cands = list(range(i))
        g.add_edge(i, rng.choice(cands))
    for u in range(n):
        for v in range(u + 1, n):
            if not g.has_edge(u, v) and rng.random() < p_extra:
                if max_deg is None or (g.degree(u) < max_deg and g.degree(v) < max_deg):
                    g.add_edge(u, v)
    return g


def count_induced(data: nx.Graph, query: nx.Graph) -> int:
    gm = iso.GraphMatcher(data, query, node_match=lambda a, b: a["label"] == b["label"])
    return sum(1 for _ in gm.subgraph_isomorphisms_iter())


def as_json(g: nx.Graph):
    return {"labels": [int(g.nodes[i]["label"]) for i in range(g.number_of_nodes())],
            "edges": [[int(u), int(v)] for u, v in sorted(tuple(sorted(e)) for e in g.edges())]}


def main():
    rng = random.Random(20260101)
    cases = []
    # (data sizes, query sizes, edge density, label alphabet, max degree)
    specs = [
        ((6, 12), (2, 5), 0.10, [8], None),            # one label: structure only
        ((6, 12), (2, 5), 0.15, [8, 7], None),         # two labels
        ((8, 16), (3, 6), 0.08, [8, 8, 8, 7, 0], 4),   # molecule-like, degree <= 4
        ((10, 20), (3, 7), 0.05, [8, 7, 0, 9], 4),
        ((5, 9), (2, 4), 0.35, [8, 7], None),          # dense: many rings of 5/6
    ]
    for si, (dsz, qsz, dens, labels, maxdeg) in enumerate(specs):
        data = [random_graph(rng, rng.randint(*dsz), dens, labels, maxdeg) for _ in range(12)]
        queries = []
        for _ in range(10):
            if rng.random() < 0.5:
                # connected induced subgraph of a data graph (guarantees some matches)
                src = rng.choice(data)
                k = rng.randint(*qsz)
                start = rng.randrange(src.number_of_nodes())
                nodes = [start]
                frontier = set(src.neighbors(start))
                while len(nodes) < k and frontier:
                    v = rng.choice(sorted(frontier))
                    frontier.discard(v)
                    nodes.append(v)
                    frontier |= set(w for w in src.neighbors(v) if w not in nodes)
                # relabel in discovery order
                mapping = {old: i for i, old in enumerate(nodes)}
                q = nx.Graph()
                for old, i in mapping.items():
                    q.add_node(i, label=src.nodes[old]["label"])
                for u, v in src.subgraph(nodes).edges():
                    q.add_edge(mapping[u], mapping[v])
            else:
                q = random_graph(rng, rng.randint(*qsz), dens, labels, maxdeg)
            queries.append(q)
        counts = [[count_induced(d, q) for d in data] for q in queries]
        cases.append({"name": f"spec{si}", "data": [as_json(d) for d in data],
                      "queries": [as_json(q) for q in queries], "counts": counts})
    meta = {"generator": "tests/golden/make_golden.py", "networkx": nx.__version__,
            "method": "GraphMatcher(data, query, node_match=label equality)"
                      ".subgraph_isomorphisms_iter(): node-induced embeddings",
            "cases": cases}
    OUT.write_text(json.dumps(meta, indent=None, separators=(",", ":")))
    total = sum(sum(map(sum, c["counts"])) for c in cases)
    print(f"wrote {OUT} ({len(cases)} cases, total embeddings {total})")


if __name__ == "__main__":
    main()
