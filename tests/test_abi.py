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
    src.write_text('#include "dfq_bfs.h"\n#include "dfq_graph.h"\n#include "dfq_batch.h"\n'
                   '#include "dfq_serve.h"\nint main(void){return 0;}\n')
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


def test_serve_config_and_footprint(dfq):
    """Serving defaults and the admission footprint (no GPU needed)."""
    c = dfq.default_serve_config()
    assert c.block_size == 1024 and c.xcds_per_team == 1 and c.ring_slots == 1024 and c.arena_bytes == 0
    n, m = 1 << 20, 16 << 20

    def al(x):
        return (x + 255) // 256 * 256
    words = (n + 31) // 32
    # item capacity: m / 256 + min(n, m / (large_max + 1)) + 1
    cap = m // 256 + min(n, m // 257) + 1
    for xpt, segs in ((1, 1), (2, 2), (8, 8), (0, 8)):
        expect = (al(4 * n) + 2 * al(4 * words) + al(8 * (words + 1)) + 2 * al(4 * segs * n)
                  + al(8 * segs * cap))
        assert dfq.serve_footprint(n, m, 256, xpt, 0) == expect
        assert dfq.serve_footprint(n, m, 256, xpt, dfq.SERVE_PARENTS) == expect + al(4 * n)
    # a larger row cut-off can only shrink the work-item bound
    assert dfq.serve_footprint(n, m, 4096, 1, 0) < dfq.serve_footprint(n, m, 256, 1, 0)
    # tiny graph: no CSR entries still has room for one item per segment
    assert dfq.serve_footprint(1, 0, 256, 1, 0) == al(4) * 3 + al(8) + 2 * al(4) + al(8)


def test_server_create_rejects_bad_input(dfq):
    ro = np.array([0, 1, 2], np.int64)
    ci = np.array([1, 0], np.uint32)
    n = np.array([2], np.uint32)
    m = np.array([2], np.uint64)
    h = dfq._P()
    mp = m.ctypes.data_as(dfq.C.POINTER(dfq.C.c_uint64))
    cfg = dfq.default_serve_config()
    cfg.ring_slots = 1000  # not a power of two
    assert dfq.lib.dfq_server_create(1, dfq._u32(n), mp, dfq._i64(ro), dfq._u32(ci), dfq.C.byref(cfg),
                                     dfq.C.byref(h)) == dfq.ERR_INVALID_ARG
    cfg = dfq.default_serve_config()
    cfg.xcds_per_team = 3
    assert dfq.lib.dfq_server_create(1, dfq._u32(n), mp, dfq._i64(ro), dfq._u32(ci), dfq.C.byref(cfg),
                                     dfq.C.byref(h)) == dfq.ERR_INVALID_ARG
    bad = np.array([0, 1, 3], np.int64)  # last offset != m
    assert dfq.lib.dfq_server_create(1, dfq._u32(n), mp, dfq._i64(bad), dfq._u32(ci), None,
                                     dfq.C.byref(h)) == dfq.ERR_RANGE
    # the rest needs a device; without one the create call reports it
    if not gpu_available():
        assert dfq.lib.dfq_server_create(1, dfq._u32(n), mp, dfq._i64(ro), dfq._u32(ci), None,
                                         dfq.C.byref(h)) == dfq.ERR_NO_DEVICE
    # calls on a null server are argument errors, not crashes
    t = dfq.C.c_uint64()
    assert dfq.lib.dfq_server_submit(None, 0, 0, 0, dfq.C.byref(t)) == dfq.ERR_INVALID_ARG
    assert dfq.lib.dfq_server_start(None) == dfq.ERR_INVALID_ARG
    assert dfq.lib.dfq_server_release(None, 0) == dfq.ERR_INVALID_ARG
    assert dfq.lib.dfq_strerror(dfq.ERR_PENDING) == b"query still running"


def test_config_default_and_launch_modes(dfq):
    """Defaults need no GPU; the launch-mode constants of the header and the
    binding agree."""
    c = dfq.Config()
    assert dfq.lib.dfq_config_default(dfq.C.byref(c)) == 0
    assert c.launch_mode == dfq.LAUNCH_PLAIN and c.bu_enter_div == 40 and c.bu_exit_div == 96
    text = open(os.path.join(ROOT, "include", "dfq_bfs.h")).read()
    consts = dict(re.findall(r"#define (DFQ_LAUNCH_[A-Z_]+) (\d+)", text))
    assert consts == {"DFQ_LAUNCH_PLAIN": "0", "DFQ_LAUNCH_COOPERATIVE": "1", "DFQ_LAUNCH_PLAIN_ABANDON": "2"}
    assert (dfq.LAUNCH_PLAIN, dfq.LAUNCH_COOPERATIVE, dfq.LAUNCH_PLAIN_ABANDON) == (0, 1, 2)
