"""Write the golden fixtures under tests/golden/.

    python tests/golden/make_golden.py

Fixtures (all small, JSON or npz without pickles):
  hbitvector_kat.json   the known-answer test of test_hbitvector.cpp:569-661:
                        5 rows of ranges over 11431 columns and the compacted
                        entries/rowmap the reference expects
  crsmatrix_tiny.json   the 10x10, 25-cell matrix of test_crsmatrix.cpp:398-422
  example26.json        manuscript Example 26 (p.26-27): the word
                        0011110011001110 (bit 0 first) -> [2,5],[8,9],[12,14]
  micro_graphs.json     hand-built graphs (chain, SNP bubble, deletion, self
                        loop, two components) with D for several [d1, d2],
                        computed by dense walk-length layering and checked
                        equal to the oracle before writing
  preset1_2000bp.npz       A of the preset-1 generator at 2000 bp and its D for
                        [15, 45], computed by sparse walk-length layering
                        (scipy) and checked equal to the oracle
Each expected output is produced by a method independent of the oracle (the
reference's literal arrays, dense or scipy products) and compared with the
oracle before it is written.
"""
import json
import os
import sys

import numpy as np
import scipy.sparse as sp

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import diverg_amd as dv  # noqa: E402
from oracle import oracle as orc  # noqa: E402
from tests.rcrs_util import from_dense, walk_index_dense  # noqa: E402

# test_hbitvector.cpp:569-608 (input) and :617-661 (expected)
KAT_IN_ROWMAP = [0, 10, 18, 18, 20, 34]
KAT_IN_ENTRIES = [0, 63, 500, 511, 512, 639, 1472, 1535, 4091, 4200,
                  1, 64, 500, 639, 1471, 1555, 11300, 11430,
                  0, 11430,
                  32, 32, 65, 130, 140, 514, 543, 1034, 1036, 1036, 1038, 2080, 8911, 8911]
KAT_OUT_ROWMAP = [0, 8, 16, 16, 18, 32]
KAT_OUT_ENTRIES = [0, 63, 500, 639, 1472, 1535, 4091, 4200,
                   1, 64, 500, 639, 1471, 1555, 11300, 11430,
                   0, 11430,
                   32, 32, 65, 130, 140, 514, 543, 1034, 1036, 1036, 1038, 2080, 8911, 8911]
KAT_NCOLS = 11431

# test_crsmatrix.cpp:398-422
TINY_CELLS = [(0, 3), (0, 4), (0, 5), (0, 6), (1, 7), (1, 8), (1, 9), (2, 0), (2, 1), (2, 6),
              (4, 1), (4, 2), (5, 0), (5, 1), (5, 2), (5, 3), (5, 4), (5, 5), (5, 6), (5, 7),
              (5, 8), (5, 9), (8, 9), (9, 8), (9, 9)]


def adj_from_edges(n, edges):
    a = np.zeros((n, n), dtype=bool)
    for u, v in edges:
        a[u, v] = True
    return a


def micro_graphs():
    g = {}
    g["chain12"] = adj_from_edges(12, [(i, i + 1) for i in range(11)])
    # 0..5 backbone, SNP at 3: 2 -> 6 -> 4 (alt vertex numbered after 2 would
    # reorder; here it is appended, orientation does not matter)
    g["snp"] = adj_from_edges(7, [(0, 1), (1, 2), (2, 3), (3, 4), (4, 5), (2, 6), (6, 4)])
    g["deletion"] = adj_from_edges(10, [(i, i + 1) for i in range(9)] + [(2, 6)])
    g["self_loop"] = adj_from_edges(6, [(i, i + 1) for i in range(5)] + [(3, 3)])
    g["two_components"] = adj_from_edges(
        9, [(0, 1), (1, 2), (2, 3), (0, 2), (4, 5), (5, 6), (6, 7), (7, 8), (5, 7)])
    g["cycle"] = adj_from_edges(5, [(0, 1), (1, 2), (2, 3), (3, 4), (4, 0)])
    return g


BOUNDS = [(0, 1), (1, 1), (1, 3), (2, 5), (0, 6), (4, 4), (3, 9)]


def mat_json(m):
    return {"ncols": int(m.ncols), "rowmap": [int(x) for x in m.rowmap],
            "entries": [int(x) for x in m.entries], "nnz": int(m.nnz)}


def same(a, b):
    return (np.array_equal(np.asarray(a.rowmap, np.uint64), np.asarray(b.rowmap, np.uint64))
            and np.array_equal(np.asarray(a.entries, np.uint64), np.asarray(b.entries, np.uint64))
            and int(a.nnz) == int(b.nnz))


def sparse_walk_index(a, dlo, dup):
    """D by sparse layering: cur = I; D |= cur for lengths in [dlo, dup]."""
    n = a.nrows
    rows = np.repeat(np.arange(n), np.diff(a.rowmap.astype(np.int64)) // 2)
    cols_s = a.entries[0::2].astype(np.int64)
    cols_e = a.entries[1::2].astype(np.int64)
    r, c = [], []
    for i, s, e in zip(rows, cols_s, cols_e):
        r.extend([i] * (e - s + 1))
        c.extend(range(s, e + 1))
    A = sp.csr_matrix((np.ones(len(r), np.int32), (r, c)), shape=(n, n))
    cur = sp.identity(n, dtype=np.int32, format="csr")
    D = sp.csr_matrix((n, n), dtype=np.int32)
    for length in range(dup + 1):
        if length >= dlo:
            D = ((D + cur) > 0).astype(np.int32)
        cur = ((cur @ A) > 0).astype(np.int32)
    return from_dense(D.toarray() > 0)


def main():
    # KAT: identity(5) . M canonicalises M's rows through the bit window
    kat = dv.RangeCRS(KAT_NCOLS, np.array(KAT_IN_ENTRIES, np.uint32),
                      np.array(KAT_IN_ROWMAP, np.uint64))
    got = orc.spgemm(orc.identity(5), kat)
    assert got.rowmap.tolist() == KAT_OUT_ROWMAP and got.entries.tolist() == KAT_OUT_ENTRIES
    with open(os.path.join(HERE, "hbitvector_kat.json"), "w") as f:
        json.dump({"source": "test_hbitvector.cpp:569-661", "ncols": KAT_NCOLS,
                   "rowmap": KAT_IN_ROWMAP, "entries": KAT_IN_ENTRIES,
                   "expected_rowmap": KAT_OUT_ROWMAP,
                   "expected_entries": KAT_OUT_ENTRIES}, f, indent=1)

    with open(os.path.join(HERE, "crsmatrix_tiny.json"), "w") as f:
        json.dump({"source": "test_crsmatrix.cpp:398-422", "nrows": 10, "ncols": 10,
                   "cells": TINY_CELLS}, f, indent=1)

    word = "0011110011001110"  # bit 0 first
    with open(os.path.join(HERE, "example26.json"), "w") as f:
        json.dump({"source": "manuscript Example 26, p.26-27", "bits": word,
                   "expected_entries": [2, 5, 8, 9, 12, 14]}, f, indent=1)

    out = {}
    for name, a in micro_graphs().items():
        ar = from_dense(a)
        cases = []
        for dlo, dup in BOUNDS:
            if dup == 0:
                continue
            exp = from_dense(walk_index_dense(a, dlo, dup))
            o = orc.dindex(ar, dlo, dup)
            assert same(o, exp), (name, dlo, dup)
            cases.append({"dlo": dlo, "dup": dup, "D": mat_json(exp)})
        out[name] = {"A": mat_json(ar), "cases": cases}
    with open(os.path.join(HERE, "micro_graphs.json"), "w") as f:
        json.dump(out, f)

    a = dv.generate_graph(1, backbone_bp=2000)
    exp = sparse_walk_index(a, 15, 45)
    o = orc.dindex(a, 15, 45)
    assert same(o, exp)
    np.savez(os.path.join(HERE, "preset1_2000bp.npz"), a_rowmap=a.rowmap, a_entries=a.entries,
             a_ncols=a.ncols, a_nnz=a.nnz, d_rowmap=np.asarray(exp.rowmap, np.uint64),
             d_entries=exp.entries, d_nnz=exp.nnz, dlo=15, dup=45)
    print("golden fixtures written")


if __name__ == "__main__":
    main()
