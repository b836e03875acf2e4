"""Pins the L2GMap plan restatement (oracle/l2gmap.py) against real MPI runs
of spmv/L2GMap.cpp:15-170 restated in oracle/l2gmap_mpi.c (fixtures written
by tests/golden/make_l2gmap_golden.py under mpirun -n 1..4).  CPU only."""
import glob
import json
import os

import pytest

from oracle import l2gmap as L

GOLD = os.path.join(os.path.dirname(__file__), "golden", "l2gmap")
KEYS = ["neighbours", "send_count", "recv_count", "send_offset", "recv_offset", "indexbuf"]


def load_case(case, size):
    ranks = [json.load(open(os.path.join(GOLD, f"{case}_{size}_{r}.json"))) for r in range(size)]
    return ranks


CASES = sorted({tuple(os.path.basename(p)[:-5].split("_")[:2])
                for p in glob.glob(os.path.join(GOLD, "*.json"))})


def test_fixtures_present():
    assert len(CASES) == 16


@pytest.mark.parametrize("case,size", CASES)
def test_plan_matches_mpi(case, size):
    gold = load_case(case, int(size))
    plans = L.build_plans([g["local_size"] for g in gold], [g["ghosts"] for g in gold])
    for g, p in zip(gold, plans):
        assert p["ranges"] == g["ranges"]
        for k in KEYS:
            want = g[k]
            if k in ("send_count", "recv_count") and not g["neighbours"]:
                want = [0]
            assert p[k] == want, (k, g["rank"])


@pytest.mark.parametrize("case,size", CASES)
def test_update_and_reverse_match_mpi(case, size):
    gold = load_case(case, int(size))
    plans = L.build_plans([g["local_size"] for g in gold], [g["ghosts"] for g in gold])
    vecs, wecs = [], []
    for r, g in enumerate(gold):
        ls, ng = g["local_size"], len(g["ghosts"])
        r0 = g["ranges"][r]
        vecs.append([float(r0 + i) for i in range(ls)] + [-1.0] * ng)
        wecs.append([0.0] * ls + [float(1000 * (r + 1) + i) for i in range(ng)])
    L.update_all(plans, vecs)
    L.reverse_update_all(plans, wecs)
    for r, g in enumerate(gold):
        assert vecs[r] == g["update_v"]
        assert wecs[r] == g["reverse_w"]
        # update lands each ghost's global index in the ghost region
        assert vecs[r][g["local_size"]:] == [float(x) for x in g["ghosts"]]


def test_plan_errors():
    with pytest.raises(RuntimeError, match="Ghosts must be sorted"):
        L.build_plans([3, 3], [[4, 3], []])
    with pytest.raises(RuntimeError, match="Ghost index in local range"):
        L.build_plans([3, 3], [[1], []])
