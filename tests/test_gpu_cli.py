"""The drop-in CLI end to end on the GPU: `beep -g G.bel -t Q.mtx -m sgc|sgm`
(src/main.cu:27-344) must end with the reference's line
`------------- Counter = N` (subgraph_matching.cuh:1277), exit 0
(main.cu:343), with N equal to the oracle's count for the same files."""
import os
import subprocess

import numpy as np
import pytest

import beep_amd as B
import oracle_py as O
from beep_amd import patterns as P

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "beep_amd", "bin", "beep")


def er_bel(path, n, p, seed):
    rng = np.random.default_rng(seed)
    a = np.triu(rng.random((n, n)) < p, 1)
    u, v = np.nonzero(a)
    pairs = sorted(set(zip(u.tolist(), v.tolist())) | set(zip(v.tolist(), u.tolist())))
    g = B.Graph.from_edges([x for x, _ in pairs], [y for _, y in pairs])
    g.write_bel(path)


def oracle_count(gpath, qname, task):
    src, dst = O.read_edges(gpath)
    og = O.Graph(src, dst)
    s, d = P.directed(P.read_mtx_edges(qname))
    return O.Prepared(og, O.plan(s, d, B.mode_of(task))).count()["count"]


def run_cli(args):
    r = subprocess.run([CLI] + args, capture_output=True, text=True, timeout=120)
    return r


@pytest.mark.parametrize("qname,task", [("triangle", "sgc"), ("triangle", "sgm"), ("4clique", "sgm"),
                                        ("diamond", "sgm"), ("diamond", "sgc"), ("w5", "sgc")])
def test_cli_counter_line(tmp_path, qname, task):
    gpath = str(tmp_path / "g.bel")
    er_bel(gpath, 120, 0.2, 3)
    print("[case] cli", qname, task, flush=True)
    r = run_cli(["-g", gpath, "-t", P.mtx_path(qname), "-m", task, "-d", "0"])
    assert r.returncode == 0, r.stderr
    last = r.stdout.strip().splitlines()[-1]
    assert last.startswith("------------- Counter = "), last
    n = int(last.split("=")[1])
    assert n == oracle_count(gpath, qname, task)


def test_cli_rmat16_triangles_with_sort(tmp_path):
    """Config C1 through the CLI: R-MAT scale 16 written as .bel, read back
    with -s 1 (main.cu:88-91), triangle sgc."""
    gpath = str(tmp_path / "rmat16.bel")
    B.Graph.rmat(16, 16 << 16, 0.57, 0.19, 0.19, 1).write_bel(gpath)
    print("[case] cli rmat16 triangle sgc -s 1", flush=True)
    r = run_cli(["-g", gpath, "-t", P.mtx_path("triangle"), "-m", "sgc", "-s", "1"])
    assert r.returncode == 0, r.stderr
    last = r.stdout.strip().splitlines()[-1]
    assert last == f"------------- Counter = {oracle_count(gpath, 'triangle', 'sgc')}"
