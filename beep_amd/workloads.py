"""Synthetic workloads of the enumeration path (deterministic generators).

Each graph is symmetrised with self loops and duplicates removed and rows
sorted, i.e. what the reference reads from a (src, dst)-sorted .bel file.

  rmat16       R-MAT scale 16, 16 * 2^16 samples, (.57,.19,.19,.05), seed 1
               -- triangle, sgc
  chunglu1m    Chung-Lu n = 1,134,890, gamma 2.2, i0 3.5, 3,070,000 samples,
               seed 2 (2,983,214 undirected edges, max degree 29,277)
               -- 4-clique then diamond, sgm, one query at a time
  rmat22cut    R-MAT scale 22 with ids >= 3,774,768 rejected, (.57,.19,.19,.05),
               seed 3 -- the 8 named central-node patterns, batched
  chunglu3m    Chung-Lu n = 3,072,441, gamma 2.0, i0 700, 121,350,000 samples, seed 4
               (about 117M undirected edges, max degree about 34k) -- serving stream
  rmat26       R-MAT scale 26, 27 * 2^26 samples, seed 5 -- large single GPU
"""
from . import Graph

WORKLOADS = {
    "rmat16": dict(kind="rmat", scale=16, samples=16 << 16, seed=1, n_keep=0,
                   queries=[("triangle", "sgc")]),
    "chunglu1m": dict(kind="chung_lu", n=1134890, samples=3070000, gamma=2.2, i0=3.5, seed=2,
                      queries=[("4clique", "sgm"), ("diamond", "sgm")]),
    "rmat22cut": dict(kind="rmat", scale=22, samples=17_030_000, seed=3, n_keep=3774768,
                      queries=[("triangle", "sgc"), ("tailed_triangle", "sgc"), ("diamond", "sgc"),
                               ("4clique", "sgc"), ("4star", "sgc"), ("w5", "sgc"),
                               ("5clique", "sgc"), ("6clique", "sgc")]),
    "chunglu3m": dict(kind="chung_lu", n=3072441, samples=121_350_000, gamma=2.0, i0=700.0, seed=4,
                      queries=[]),
    "rmat26": dict(kind="rmat", scale=26, samples=27 << 26, seed=5, n_keep=0,
                   queries=[("4clique", "sgc"), ("w5", "sgc"), ("5clique", "sgc")]),
}


def make_graph(name):
    w = WORKLOADS[name]
    if w["kind"] == "rmat":
        return Graph.rmat(w["scale"], w["samples"], 0.57, 0.19, 0.19, w["seed"], w["n_keep"])
    return Graph.chung_lu(w["n"], w["samples"], w["gamma"], w["i0"], w["seed"])
