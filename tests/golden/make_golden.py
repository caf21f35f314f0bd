"""Writes the golden vectors under tests/golden/ (run in the build container).

nx_golden.json   networkx 3.x betweenness (unnormalised, x2 = ordered source/target
                 pairs, the reference's convention) and farness (sum of BFS
                 distances inside the component) of each small vendored graph.
ref_golden.json  outputs of the reference's OWN computeBC (FP32, per-predecessor
                 formula) and computeCC / compute_D1_AP_CC_shareBased_DegreeOrder,
                 produced by the drivers built from its sources by oracle/Makefile
                 (`make -C oracle ref`); skipped when those are not built.

Vectors are indexed by the file's own ids startNodeID..endNodeID.

RAW_GRAPHS are multigraph fixtures (repeated edges, self-loops) that the
reference loads as they are (Lib/AdjList/AdjList.c:76-92); they get
reference vectors only (networkx graphs are simple).  `python make_golden.py
ref` rewrites ref_golden.json alone.
"""
import sys
import json
import os
import subprocess
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
DS = os.path.join(HERE, "datasets")
ROOT = os.path.dirname(os.path.dirname(HERE))
GRAPHS = ["karate.txt", "dolphins.txt", "football.txt", "polbooks.txt", "testdata.txt",
          "testdata1.txt", "testdata2.txt", "ia-fb-messages.mtx"]
RAW_GRAPHS = ["multi.txt"]


def read_edges(path):
    with open(path) as f:
        n, m = map(int, f.readline().split()[:2])
        edges = [tuple(map(int, f.readline().split()[:2])) for _ in range(m)]
    zero = any(a == 0 or b == 0 for a, b in edges)
    start = 0 if zero else 1
    return n, edges, start


def nx_golden():
    import networkx as nx
    out = {}
    for name in GRAPHS:
        n, edges, start = read_edges(os.path.join(DS, name))
        G = nx.Graph()
        G.add_nodes_from(range(start, start + n))
        G.add_edges_from((a, b) for a, b in edges if a != b)
        bc = nx.betweenness_centrality(G, normalized=False)
        far = {}
        for s in G.nodes:
            far[s] = int(sum(nx.single_source_shortest_path_length(G, s).values()))
        ids = list(range(start, start + n))
        out[name] = {"start": start, "bc": [2.0 * bc[i] for i in ids], "farness": [far[i] for i in ids],
                     "generator": "networkx " + nx.__version__}
    with open(os.path.join(HERE, "nx_golden.json"), "w") as f:
        json.dump(out, f)


def ref_golden():
    ref_cc = os.path.join(ROOT, "oracle", "_ref", "ref_cc")
    ref_bc = os.path.join(ROOT, "oracle", "_ref", "ref_bc")
    if not (os.path.exists(ref_cc) and os.path.exists(ref_bc)):
        print("oracle/_ref not built; skipping ref_golden.json")
        return
    out = {}
    with tempfile.TemporaryDirectory() as td:
        for name in GRAPHS + RAW_GRAPHS:
            path = os.path.join(DS, name)
            entry = {"raw": True} if name in RAW_GRAPHS else {}
            for mode in ("cc", "d1ap"):
                o = os.path.join(td, mode)
                subprocess.run([ref_cc, mode, path, o], check=True, stdout=subprocess.DEVNULL,
                               timeout=600)
                rows = [tuple(map(int, l.split())) for l in open(o)]
                entry["start"] = rows[0][0]
                entry[mode] = [r[1] for r in rows]
            o = os.path.join(td, "bc")
            subprocess.run([ref_bc, path, o], check=True, stdout=subprocess.DEVNULL, timeout=600)
            entry["bc_f32"] = [float(l) for l in open(o)]
            out[name] = entry
    with open(os.path.join(HERE, "ref_golden.json"), "w") as f:
        json.dump(out, f)


if __name__ == "__main__":
    if sys.argv[1:] != ["ref"]:
        nx_golden()
    ref_golden()
