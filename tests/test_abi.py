"""The C ABI library and the host logic (CPU only: no compute calls)."""
import ctypes
import subprocess
import os
import re

import numpy as np
import pytest

import diverg_amd as dv
from diverg_amd import _native as N

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for h in os.listdir(os.path.join(ROOT, "include")):
        if not h.endswith(".h"):
            continue
        text = open(os.path.join(ROOT, "include", h)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^\s*(?:const\s+)?[\w\s\*]+?\b(dv_\w+)\s*\(", text, re.M):
            names.add(m.group(1))
    return sorted(names)


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(N.LIB_PATH)
    names = declared_functions()
    assert len(names) >= 15
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    # the Python binding declares every one of them
    bound = {n for n, _, _ in N.SIGNATURES}
    assert set(names) <= bound, set(names) - bound


def test_abi_version_and_status_strings():
    assert N.lib().dv_abi_version() == 1
    assert N.lib().dv_status_string(0) == b"ok"
    assert N.lib().dv_status_string(6) == b"index out of range"


def test_struct_sizes_match_header():
    assert ctypes.sizeof(N.dv_rcrs) == 8 * 6 + 4 * 2 + 8
    assert ctypes.sizeof(N.dv_build_stats) == 8 * 5 + 4 * 2 + 8 * 2 + 4 * 4
    assert ctypes.sizeof(N.dv_tuning) == 4 * 12
    assert ctypes.sizeof(N.dv_kernel_time) == 8 * 4


def _canonical(m):
    rm = m.rowmap.astype(np.int64)
    en = m.entries.astype(np.int64)
    assert rm[0] == 0 and rm[-1] == len(en) and np.all(np.diff(rm) % 2 == 0)
    s, e = en[0::2], en[1::2]
    assert np.all(s <= e) and np.all(e < m.ncols)
    # within a row: sorted, disjoint, non-adjacent
    row_of = np.repeat(np.arange(m.nrows), np.diff(rm) // 2)
    same_row = row_of[1:] == row_of[:-1]
    assert np.all(s[1:][same_row] > e[:-1][same_row] + 1)


@pytest.mark.parametrize("preset", [1, 2])
def test_generator_canonical_and_deterministic(preset):
    kw = {} if preset == 1 else {"backbone_bp": 300000}
    a = dv.generate_graph(preset, **kw)
    b = dv.generate_graph(preset, **kw)
    assert np.array_equal(a.entries, b.entries) and np.array_equal(a.rowmap, b.rowmap)
    _canonical(a)
    assert a.nrows == a.ncols
    assert 1.0 < a.nnz / a.nrows < 1.05
    c = dv.generate_graph(preset, seed=12345, **kw)
    assert not (len(c.entries) == len(a.entries) and np.array_equal(c.entries, a.entries))


def test_generator_preset_sizes():
    assert dv.graph_preset(1).backbone_bp == 100000
    assert dv.graph_preset(2).backbone_bp == 5000000
    assert dv.graph_preset(3).backbone_bp == 50000000
    assert dv.graph_preset(3).sv_rate == 1e-5
    lens = [dv.graph_preset(4, c).backbone_bp for c in range(24)]
    assert all(100000 <= x <= 120000000 for x in lens)
    assert len(set(lens)) == 24
    with pytest.raises(dv.DiVerGError):
        dv.graph_preset(9)


def test_generator_structural_variants():
    a, vpos = dv.generate_graph(3, backbone_bp=200000, with_vpos=True)
    _canonical(a)
    assert a.nrows > 200000          # bubble paths add vertices
    assert np.all(np.diff(vpos.astype(np.int64)) >= 0)
    assert vpos[-1] == 199999
    # edges go forward in the vertex order except none (acyclic by construction)
    rows = np.repeat(np.arange(a.nrows), np.diff(a.rowmap.astype(np.int64)) // 2)
    assert np.all(a.entries[0::2] > rows)


def test_generator_u32_rowmap():
    a = dv.generate_graph(1, size_bytes=4)
    assert a.rowmap.dtype == np.uint32
    b = dv.generate_graph(1, size_bytes=8)
    assert np.array_equal(a.rowmap.astype(np.uint64), b.rowmap)


def test_rangecrs_nnz_wraps_for_u32():
    big = dv.RangeCRS(2 ** 32 - 2, np.array([0, 2 ** 32 - 3] * 3, np.uint32),
                      np.array([0, 2, 4, 6], np.uint32))
    assert big.nnz == (3 * (2 ** 32 - 2)) % (1 << 32)


def test_missing_library_fails_loudly(monkeypatch):
    monkeypatch.setattr(N, "_lib", None)
    monkeypatch.setattr(N, "LIB_PATH", "/nonexistent/libdiverg_amd.so")
    with pytest.raises(dv.NativeMissing):
        N.lib()


def test_init_without_device_reports_device_error():
    hip = ctypes.CDLL("libamdhip64.so")
    n = ctypes.c_int(0)
    if hip.hipGetDeviceCount(ctypes.byref(n)) == 0 and n.value > 0:
        pytest.skip("a GPU is present")
    with pytest.raises(dv.DiVerGError) as e:
        dv.Context(0)
    assert e.value.status == 4


def test_cpp_wrapper_headers_compile(tmp_path):
    """The drop-in headers compile as the reference's call sites use them."""
    src = tmp_path / "use.cpp"
    src.write_text(
        "#include <diverg/dindex.hpp>\n"
        "using M = diverg::CRSMatrix<diverg::crs_matrix::RangeDynamic, bool, uint32_t, uint64_t>;\n"
        "using M32 = diverg::CRSMatrix<diverg::crs_matrix::RangeDynamic, bool, uint32_t, uint32_t>;\n"
        "M f(M const& a) { return diverg::util::create_distance_index(a, 150, 450); }\n"
        "M32 g(M32 const& a) { return diverg::range_power(diverg::range_spadd(a, a), 3); }\n"
        "M h(M const& a) { return diverg::range_spgemm(a, a, diverg::DefaultSparseConfiguration{}); }\n")
    res = subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-Wall", "-Werror",
                          "-I" + os.path.join(ROOT, "include"),
                          "-I" + os.path.join(ROOT, "diverg_amd", "include"), str(src)],
                         capture_output=True, text=True)
    assert res.returncode == 0, res.stderr
