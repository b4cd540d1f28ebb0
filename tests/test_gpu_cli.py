"""The `bfs` command line as the reference's harness drives it
(plot.py:30-32: `bin <data> 8` through a shell, time = the next-to-last
whitespace token of stdout minus its "ms"), and its report lines
(design/shared-q/atomic-sa/our-sep/upper-bound/bfs.cu:206-208,297-300,311)
checked against the oracle."""
import os
import subprocess

import numpy as np
import pytest

from conftest import PKG, UNREACHED

CLI = os.path.join(PKG, "bin", "bfs")


def report(stdout):
    """The reference's report fields from the text output."""
    f = {}
    for line in stdout.splitlines():
        line = line.replace("\033[0m", "").replace("\033[0;1m", "").replace("\033[31m", "")
        if line.startswith("- Dist  ="):
            f["dist"] = int(line.split("=")[1])
        elif line.startswith("- Touch ="):
            f["touch"] = int(line.split("(")[1].split("/")[0])
        elif line.startswith("- Error ="):
            f["error"] = int(line.split("=")[1])
        elif line.startswith("# Nodes :"):
            f["nodes"] = int(line.split(":")[1])
        elif line.startswith("# Edges :"):
            f["edges"] = int(line.split(":")[1])
        elif line.startswith("Source  :"):
            f["source"] = int(line.split(":")[1])
    return f


@pytest.mark.gpu
@pytest.mark.parametrize("kind", ["dat", "snap"])
def test_cli_as_plot_py_runs_it(dfq, oracle, tmp_path, kind):
    g = dfq.Csr.rmat(12, 8, seed=12)
    src = int(g.pick_sources(1, seed=3)[0]) if kind == "dat" else 0
    if kind == "dat":
        path = str(tmp_path / "rmat12.dat")
        g.save_dat(path, source=src)
    else:
        path = str(tmp_path / "rmat12.txt")
        g.save_snap(path)
    print("cli", kind, path, "source", src, flush=True)
    # plot.py:31: Popen("{}/{} {}/{} 8".format(...), shell=True, stdout=PIPE)
    p = subprocess.Popen("{} {} 8".format(CLI, path), shell=True, stdout=subprocess.PIPE,
                         stderr=subprocess.PIPE)
    stdout, stderr = p.communicate(timeout=120)
    assert p.returncode == 0, stderr.decode()
    # plot.py:32, verbatim: the token before the trailing reset, minus "ms"
    t = float(stdout.split()[-2][:-2].decode("UTF-8"))
    assert 0.0 < t < 1000.0
    assert stdout.endswith(b"\033[0m")
    f = report(stdout.decode())
    lv = oracle.bfs(g.row_offsets, g.col_idx, src)
    reached = lv != UNREACHED
    assert f["nodes"] == g.n and f["edges"] == g.m and f["source"] == src
    assert f["dist"] == int(lv[reached].max())
    assert f["touch"] == int(reached.sum())
    assert f["error"] == 0


@pytest.mark.gpu
def test_cli_interactive_sources(dfq, oracle, tmp_path):
    """`bfs <file> $`: after each report "Exit? [y/n]"; on "n" a new source
    is read (bfs.cu:202-214,302-308)."""
    g = dfq.Csr.amazon_like(4096, seed=5)
    path = str(tmp_path / "a.dat")
    g.save_dat(path, source=0)
    p = subprocess.run([CLI, path, "$"], input=b"n\n17\nn\n4095\ny\n", capture_output=True, timeout=120)
    assert p.returncode == 0, p.stderr.decode()
    out = p.stdout.decode()
    blocks = out.split("- Dist  =")[1:]
    assert len(blocks) == 3
    dists = [int(b.split("\n")[0]) for b in blocks]
    for s, d in zip((0, 17, 4095), dists):
        lv = oracle.bfs(g.row_offsets, g.col_idx, s)
        assert d == int(lv[lv != UNREACHED].max())
    assert "- Error = 0" in out and out.endswith("\033[0m")


def test_cli_refuses_bad_input_without_a_gpu(tmp_path):
    """Errors before any device work come back as a message and exit 1."""
    p = subprocess.run([CLI], capture_output=True, timeout=60)
    assert p.returncode == 2 and b"usage" in p.stderr
    bad = tmp_path / "bad.dat"
    bad.write_text("1 1 0\n18446744073709551615 2\n0 1\n")
    p = subprocess.run([CLI, str(bad), "8"], capture_output=True, timeout=60)
    assert p.returncode == 1 and b"out of range" in p.stderr
