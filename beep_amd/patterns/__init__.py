"""Pattern library: named .mtx queries and the 207 central-node patterns."""
import os

HERE = os.path.dirname(os.path.abspath(__file__))
NAMES = ["triangle", "wedge", "tailed_triangle", "diamond", "4clique", "4star", "w5",
         "5clique", "6clique"]


def mtx_path(name):
    p = os.path.join(HERE, f"{name}.mtx")
    if not os.path.exists(p):
        raise FileNotFoundError(p)
    return p


def read_mtx_edges(name):
    """Undirected edge list of a named query (each pair once, a < b)."""
    with open(mtx_path(name)) as f:
        lines = [ln.split() for ln in f if ln.strip() and not ln.startswith("%")]
    return sorted({(min(int(a), int(b)), max(int(a), int(b))) for a, b in lines[1:]})


def central207():
    """[(k, [(a, b), ...]), ...] for the 207 central-node patterns on 3..7 vertices."""
    out = []
    with open(os.path.join(HERE, "central207.txt")) as f:
        for ln in f:
            parts = ln.split()
            k = int(parts[0])
            edges = [tuple(int(x) for x in p.split("-")) for p in parts[2:]]
            out.append((k, edges))
    return out


def directed(edges):
    """Both directions, grouped by ascending source: (src list, dst list)."""
    d = sorted({(a, b) for a, b in edges} | {(b, a) for a, b in edges})
    return [a for a, _ in d], [b for _, b in d]
