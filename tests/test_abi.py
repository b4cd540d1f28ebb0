"""The C-ABI library loads and exports every function that include/*.h
declares; error paths that need no GPU return codes instead of crashing."""
import os
import re
import subprocess

import numpy as np
import pytest

from conftest import ROOT, gpu_available


def declared_functions():
    names = set()
    for h in os.listdir(os.path.join(ROOT, "include")):
        if not h.endswith(".h"):
            continue
        text = open(os.path.join(ROOT, "include", h)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"\b(dfq_[a-z0-9_]+)\s*\(", text):
            names.add(m.group(1))
    return sorted(names)


def test_headers_declare_functions():
    names = declared_functions()
    assert "dfq_bfs" in names and "dfq_graph_create" in names and "dfq_load_snap" in names


def test_every_declared_symbol_is_exported(dfq):
    out = subprocess.run(["nm", "-D", "--defined-only", dfq.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if line.strip()}
    missing = [n for n in declared_functions() if n not in exported]
    assert not missing, missing


def test_headers_compile_as_c(tmp_path):
    src = tmp_path / "t.c"
    src.write_text('#include "dfq_bfs.h"\n#include "dfq_graph.h"\nint main(void){return 0;}\n')
    subprocess.run(["gcc", "-std=c99", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"),
                    "-c", str(src), "-o", str(tmp_path / "t.o")], check=True)


def test_strerror_and_config(dfq):
    assert dfq.lib.dfq_strerror(0) == b"ok"
    assert b"range" in dfq.lib.dfq_strerror(dfq.ERR_RANGE)
    c = dfq.default_config()
    assert c.block_size in (256, 512, 1024) and c.small_max <= 64
    assert c.small_max <= c.medium_max <= c.large_max


@pytest.mark.skipif(gpu_available(), reason="checks the no-GPU error path")
def test_no_device_is_an_error_code(dfq):
    with pytest.raises(dfq.DfqError) as e:
        dfq.device_info()
    assert e.value.code == dfq.ERR_NO_DEVICE
    g = dfq.Csr.rmat(6, 4, seed=1)
    with pytest.raises(dfq.DfqError) as e:
        dfq.DeviceGraph(g)
    assert e.value.code == dfq.ERR_NO_DEVICE


def test_graph_create_validates_input_before_device(dfq):
    # offsets not ending at m / column out of range are rejected on the host
    with pytest.raises(dfq.DfqError) as e:
        dfq.DeviceGraph(row_offsets=np.array([0, 1, 3]), col_idx=np.array([1, 0], np.uint32))
    assert e.value.code == dfq.ERR_RANGE
    with pytest.raises(dfq.DfqError) as e:
        dfq.DeviceGraph(row_offsets=np.array([0, 1, 2]), col_idx=np.array([1, 2], np.uint32))
    assert e.value.code == dfq.ERR_RANGE
