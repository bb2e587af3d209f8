"""The C-ABI library loads and exports every entry point include/beep.h declares;
the CLI keeps the reference's flag surface.  No compute call needs a GPU here."""
import os
import re
import subprocess

import pytest

import beep_amd as B
from beep_amd import _ffi
from beep_amd import patterns as P

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "beep.h")
CLI = os.path.join(ROOT, "beep_amd", "bin", "beep")


def declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(beep_[a-z_]+)\s*\(", text)))


def test_every_declared_symbol_is_exported():
    lib = _ffi.lib()
    names = declared()
    assert "beep_count" in names and len(names) >= 20
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_ctypes_signatures_cover_header():
    assert set(declared()) == set(_ffi.SIGNATURES)


def test_version_and_last_error():
    assert B.lib().beep_version().startswith(b"beep")
    with pytest.raises(B.BeepError) as e:
        B.Graph.load("/nonexistent.bel")
    assert "unable to open" in str(e.value)


def test_null_arguments_are_errors():
    lib = _ffi.lib()
    assert lib.beep_graph_load(None, 0, None) == -22
    assert lib.beep_count(None, None, None, None) == -22
    assert lib.beep_prep_upload(None, 0) == -22


def test_cli_help_and_refusals(tmp_path):
    r = subprocess.run([CLI, "-h"], capture_output=True, text=True)
    assert r.returncode == 0 and "-g <Src graph FileName>" in r.stderr
    r = subprocess.run([CLI, "-m", "bogus"], capture_output=True, text=True)
    assert r.returncode == 0 and "Unrecognized -mt option" in r.stderr  # reference: exit(0)
    r = subprocess.run([CLI, "-Z"], capture_output=True, text=True)
    assert r.returncode == 0 and "Unrecognized option!" in r.stderr
    r = subprocess.run([CLI, "-n", "2", "-g", "x.bel"], capture_output=True, text=True)
    assert r.returncode == 2 and "exactly one GPU" in r.stderr


def test_cli_loads_graph_for_other_tasks(tmp_path):
    g = B.Graph.rmat(8, 2000, 0.57, 0.19, 0.19, 1)
    path = str(tmp_path / "g.bel")
    g.write_bel(path)
    r = subprocess.run([CLI, "-g", path, "-m", "tc"], capture_output=True, text=True)
    assert r.returncode == 0
    assert f"{g.m},{g.n},{g.n}" in r.stdout


def test_cli_reports_missing_query(tmp_path):
    g = B.Graph.rmat(8, 2000, 0.57, 0.19, 0.19, 1)
    path = str(tmp_path / "g.bel")
    g.write_bel(path)
    r = subprocess.run([CLI, "-g", path, "-t", str(tmp_path / "none.mtx"), "-m", "sgm"],
                       capture_output=True, text=True)
    assert r.returncode == 1 and "unable to open" in r.stderr
    assert os.path.exists(P.mtx_path("triangle"))


def test_bad_options_are_refused_before_any_device_call():
    """Option checks run before anything touches a GPU, so they hold without one."""
    import ctypes as C

    import beep_amd as B
    from beep_amd import _ffi
    g = B.Graph.from_edges([0, 0, 1, 1, 2, 2], [1, 2, 0, 2, 0, 1])
    plan = B.Plan.from_undirected([(0, 1), (1, 2), (0, 2)], "sgc")
    pp = B.Prepared(g, plan)
    lib = _ffi.lib()
    for field, value in (("build_mode", 7), ("deadline_ms", -1.0), ("lanes", 65), ("emit", 9)):
        o = _ffi.Opts()
        lib.beep_opts_default(C.byref(o))
        setattr(o, field, value)
        r = _ffi.Result()
        assert lib.beep_count(pp._h, plan._h, C.byref(o), C.byref(r)) == -22, field
        assert lib.beep_last_error()
    o = _ffi.Opts()
    lib.beep_opts_default(C.byref(o))
    o.emit, o.match_cap = B.EMIT_BUFFER, 10
    assert lib.beep_count(pp._h, plan._h, C.byref(o), C.byref(_ffi.Result())) == -22
    assert b"match_buf" in lib.beep_last_error()
