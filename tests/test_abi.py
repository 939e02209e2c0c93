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
    assert ctypes.sizeof(N.dv_tuning) == 4 * 13
    assert ctypes.sizeof(N.dv_kernel_time) == 8 * 4
    assert ctypes.sizeof(N.dv_server_opts) == 8 + 4 * 4
    assert ctypes.sizeof(N.dv_request_info) == 4 * 4 + 8 * 4


STRUCTS = ["dv_rcrs", "dv_build_stats", "dv_graph_info", "dv_tuning", "dv_kernel_time",
           "dv_graph_params", "dv_server_opts", "dv_request_info"]


def test_struct_layouts_match_header(tmp_path):
    """Every ctypes struct has the size, field names and field offsets the C
    compiler gives the header's struct (a C probe prints them)."""
    lines = ["#include <stddef.h>", "#include <stdio.h>", "#include <diverg_amd.h>",
             "#include <diverg_amd_gen.h>", "#include <diverg_amd_graph.h>", "int main(void) {"]
    for name in STRUCTS:
        cls = getattr(N, name)
        lines.append(f'  printf("{name}:size %zu\\n", sizeof({name}));')
        for f in cls._fields_:
            lines.append(f'  printf("{name}.{f[0]} %zu %zu\\n", offsetof({name}, {f[0]}), '
                         f'sizeof((({name} *)0)->{f[0]}));')
    lines += ["  return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    res = subprocess.run(["gcc", "-std=c11", "-Wall", "-Werror", "-I" + os.path.join(ROOT, "include"),
                          "-o", str(exe), str(src)], capture_output=True, text=True)
    assert res.returncode == 0, res.stderr
    got = {}
    for line in subprocess.run([str(exe)], capture_output=True, text=True,
                               check=True).stdout.splitlines():
        key, *vals = line.split()
        got[key] = tuple(int(v) for v in vals)
    for name in STRUCTS:
        cls = getattr(N, name)
        assert got[f"{name}:size"] == (ctypes.sizeof(cls),), name
        for f in cls._fields_:
            fd = getattr(cls, f[0])
            assert got[f"{name}.{f[0]}"] == (fd.offset, fd.size), f"{name}.{f[0]}"


def test_server_rejects_cu_masked_query_stream():
    """A CU-masked query stream is refused before any HIP call (so this runs
    without a GPU): the mask the earlier code built placed every chosen CU on
    one XCD."""
    o = N.dv_server_opts(0, 0, 0, 16, 0)
    h = ctypes.c_void_p()
    rc = N.lib().dv_server_start(0, ctypes.byref(o), ctypes.byref(h))
    assert rc == 1 and not h.value   # DV_ERR_INVALID, no server
    with pytest.raises(dv.DiVerGError) as e:
        dv.Server(query_cus=4)
    assert e.value.status == 1


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
        "M h(M const& a) { return diverg::range_spgemm(a, a, diverg::DefaultSparseConfiguration{}); }\n"
        # rcrs_benchmark.cpp:449-450, 685-720
        "using RT = diverg::SparseConfig<diverg::grid::RunTime, diverg::HBitVectorAccumulator<>,\n"
        "                                diverg::TeamSequentialPartition>;\n"
        "M k(M const& a, diverg::amd::Timer* t1, diverg::amd::Timer* t2) {\n"
        "  RT config; config.grid.m_team_size = 4; config.grid.m_vector_size = 8;\n"
        "  return diverg::util::create_distance_index(a, 150, 450, config, t1, t2); }\n"
        # test_range_sparse.cpp:522-545 (int32_t ordinals, every tag pair)
        "using MI = diverg::CRSMatrix<diverg::crs_matrix::RangeDynamic, bool, int32_t, uint64_t>;\n"
        "template <typename A, typename P> MI r(MI const& m) {\n"
        "  typedef diverg::SparseConfig<diverg::grid::Suggested, A, P> config_type;\n"
        "  typedef typename config_type::execution_space execution_space;\n"
        "  auto rc = diverg::range_spgemm(m, m, config_type{});\n"
        "  execution_space space; space.fence(); return rc; }\n"
        "template MI r<diverg::BTreeAccumulator, diverg::ThreadRangePolicyPartition>(MI const&);\n"
        "template MI r<diverg::HBitVectorAccumulator<>, diverg::ThreadParallelPartition>(MI const&);\n"
        # range_sparse.hpp:430-468, 2283-2368 (the ETI operator classes)
        "M32 s(M32 const& a) { using S = diverg::amd::DeviceSpace;\n"
        "  auto i = diverg::create_range_identity_matrix<M32>(a.numRows(), S{});\n"
        "  auto c = diverg::RangeSpAdd<M32, S>::spadd(i, a, S{});\n"
        "  c = diverg::RangeSpGEMM<M32, diverg::DefaultSparseConfiguration>::spgemm(c, a, {});\n"
        "  return diverg::RangePower<M32, diverg::DefaultSparseConfiguration>::power(c, 5, {}); }\n")
    res = subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-Wall", "-Werror",
                          "-I" + os.path.join(ROOT, "include"),
                          "-I" + os.path.join(ROOT, "diverg_amd", "include"), str(src)],
                         capture_output=True, text=True)
    assert res.returncode == 0, res.stderr
