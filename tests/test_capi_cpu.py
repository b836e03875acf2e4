"""C ABI checks that need no GPU: the library loads, exports every function
declared in include/spmv_amd.h, builds exchange plans equal to the ones real
MPI runs of the reference algorithm produce (tests/golden/l2gmap), and raises
the reference's errors."""
import ctypes
import glob
import json
import os
import re

import numpy as np
import pytest

import spmv_amd as S
from spmv_amd import problems as P

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = os.path.join(ROOT, "tests", "golden", "l2gmap")


def declared_functions():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        src = open(h).read()
        names |= set(re.findall(r"^\s*(?:int|const char\*)\s+(spmvx_\w+)\s*\(", src, re.M))
    return sorted(names)


def test_exports_every_declared_symbol():
    names = declared_functions()
    assert len(names) >= 20
    L = ctypes.CDLL(S.library_path)
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing


def test_abi_version():
    """The loaded library reports the version the header declares."""
    import re
    hdr = open(os.path.join(os.path.dirname(GOLD), "..", "..", "include", "spmv_amd.h")).read()
    want = int(re.search(r"#define SPMVX_ABI_VERSION (\d+)", hdr).group(1))
    assert S.lib().spmvx_abi_version() == want == 2


CASES = sorted({tuple(os.path.basename(p)[:-5].split("_")[:2])
                for p in glob.glob(os.path.join(GOLD, "*.json"))})


@pytest.mark.parametrize("case,size", CASES)
def test_plan_matches_mpi_golden(case, size):
    size = int(size)
    gold = [json.load(open(os.path.join(GOLD, f"{case}_{size}_{r}.json"))) for r in range(size)]
    sizes = [g["local_size"] for g in gold]
    for r, g in enumerate(gold):
        lo, hi = g["ranges"][r], g["ranges"][r + 1]
        remote = {q: [x for x in gold[q]["ghosts"] if lo <= x < hi] for q in range(size) if q != r}
        m = S.L2GMap(g["local_size"], g["ghosts"], local_sizes=sizes, rank=r, remote_ghosts=remote)
        plan = m.plan()
        for k in ["neighbours", "send_count", "recv_count", "send_offset", "recv_offset",
                  "indexbuf"]:
            want = g[k]
            if k in ("send_count", "recv_count") and not g["neighbours"]:
                want = [0]
            assert plan[k] == want, (k, r)
        assert m.local_size() == g["local_size"]
        assert m.num_ghosts() == len(g["ghosts"])
        assert m.global_size() == g["ranges"][-1]
        assert m.global_offset() == lo
        assert list(m.ghosts()) == g["ghosts"]
        for i, gi in enumerate(g["ghosts"]):
            assert m.global_to_local(gi) == g["local_size"] + i
        if g["local_size"]:
            assert m.global_to_local(lo) == 0


def test_plan_errors_match_reference_messages():
    with pytest.raises(S.SpmvError, match="Ghosts must be sorted"):
        S.L2GMap(3, [5, 4], local_sizes=[3, 3], rank=0)
    with pytest.raises(S.SpmvError, match="Ghost index in local range"):
        S.L2GMap(3, [4], local_sizes=[3, 3], rank=1)
    with pytest.raises(S.SpmvError, match="outside the global range"):
        S.L2GMap(3, [7], local_sizes=[3, 3], rank=0)
    m = S.L2GMap(3, [4], local_sizes=[3, 3], rank=0)
    with pytest.raises(S.SpmvError, match="neither owned nor a ghost"):
        m.global_to_local(5)


def test_single_rank_map():
    m = S.L2GMap(10)
    assert (m.local_size(), m.num_ghosts(), m.global_size(), m.global_offset()) == (10, 0, 10, 0)
    assert m.plan()["send_count"] == [0] and m.plan()["indexbuf"] == []


def test_slab_decomposition_plan():
    kw = P.slab_decomposition(4, 3, 5)
    m = S.L2GMap(**kw)
    plan = m.plan()
    assert plan["neighbours"] == [1]
    assert plan["send_count"] == [12] and plan["recv_count"] == [12]
    assert plan["send_offset"] == [60, 72]
    assert plan["indexbuf"] == list(range(48, 60))


def test_matrix_create_validates_before_device():
    rp = np.array([0, 1, 2], np.int32)
    with pytest.raises(S.SpmvError, match="column index out of range"):
        S.Matrix(rp, np.array([0, 5], np.int32), np.ones(2), ncols=2)
    with pytest.raises(S.SpmvError, match="rowptr\\[nrows\\] must equal nnz"):
        S.Matrix(rp, np.array([0], np.int32), np.ones(1), ncols=2)
    with pytest.raises(S.SpmvError, match="non-decreasing"):
        S.Matrix(np.array([0, 2, 1, 2], np.int32), np.array([0, 1], np.int32), np.ones(2), ncols=2)
