"""Failure isolation and admission by state memory on the GPU engines
(search.h's per-query errors: a query that fails must not take its batch
down).  Every failure here is forced from the host; every kernel runs to its
normal end."""
import numpy as np
import pytest

import nigiri_amd as na
import oracle
from test_gpu_parity import image_for, say, with_destinations

pytestmark = pytest.mark.gpu


def w1():
    return image_for("w1", lambda: na.Timetable.generate(1))


def check_ok(tt, qs, res, skip):
    for i in range(len(qs)):
        if i in skip:
            continue
        ref = oracle.search(tt, qs[i])
        assert res[i]["status"] == 0, i
        assert res[i]["journeys"] == ref["journeys"] and res[i]["legs"] == ref["legs"], i


@pytest.mark.parametrize("lanes", [64, 1])
def test_gpu_poisoned_query_fails_alone(lanes):
    tt, img = w1()
    eng = na.Engine(img, lanes=lanes, groups=1 if lanes == 64 else 2, max_k=16)
    n = 48 if lanes == 64 else 4
    qs = na.Queries.generate(tt, 1, n, seed=0x7172 + 61)
    bad = np.array([tt.n_locations + 5], np.uint32)  # a location that does not exist
    qs[1].start_loc = bad.ctypes.data_as(qs[1].start_loc.__class__)
    qs._keep.append(bad)
    say(f"poisoned query, lanes={lanes} n={n}")
    res = eng.search(qs)
    assert res[1]["status"] == na.EINVAL and res[1]["journeys"] == []
    check_ok(tt, qs, res, {1})
    if lanes == 64:
        say("poisoned query in the serving loop")
        served, _ = eng.serve(qs, np.linspace(0.0, 20.0, n))
        assert served[1]["status"] == na.EINVAL
        for i in range(n):
            if i != 1:
                assert served[i]["status"] == 0 and served[i]["journeys"] == res[i]["journeys"], i


def test_gpu_query_beyond_the_state_rows_fails_alone():
    # an engine of max_k 9 holds rows for max_transfers <= 7; a query asking
    # for 14 transfers (end_k 16) is refused with ECAPACITY, the rest run
    tt, img = w1()
    eng = na.Engine(img, lanes=64, groups=1, max_k=9)
    qs = na.Queries.generate(tt, 1, 32, seed=0x7172 + 67)
    qs[5].max_transfers = 14
    say("end_k beyond max_k")
    res = eng.search(qs)
    assert res[5]["status"] == na.ECAPACITY and res[5]["journeys"] == []
    check_ok(tt, qs, res, {5})


def test_gpu_admission_waits_for_state_memory():
    # a budget for three queries' state: 40 queries still all complete, at
    # most three at a time, with the oracle's results; a budget below one
    # query's state refuses every query alone
    tt, img = w1()
    eng = na.Engine(img, lanes=64, groups=2, max_k=9)
    qs = na.Queries.generate(tt, 1, 40, seed=0x7172 + 71)
    per_query = tt.n_locations * (2 * 9 + 9)  # end_k = 9 for max_transfers 7
    eng.configure(state_budget_bytes=3 * per_query)
    assert eng.get_config()["state_budget_bytes"] == 3 * per_query
    say("budget of three queries, n=40")
    eng.stats(reset=True)
    res = eng.search(qs)
    check_ok(tt, qs, res, set())
    served, _ = eng.serve(qs, np.zeros(len(qs)))
    assert [r["journeys"] for r in served] == [r["journeys"] for r in res]
    eng.configure(state_budget_bytes=per_query - 1)
    say("budget below one query")
    res = eng.search(qs.subset(list(range(4))))
    assert all(r["status"] == na.ECAPACITY for r in res)
    eng.configure(state_budget_bytes=0)
    res = eng.search(qs.subset(list(range(4))))
    check_ok(tt, qs.subset(list(range(4))), res, set())


def test_gpu_query_beyond_the_destination_limits_fails_alone():
    # 65 exact destinations or 257 egress offsets are beyond the engines'
    # per-query limits: refused with EUNSUPPORTED, the rest of the batch runs
    tt, img = w1()
    eng = na.Engine(img, lanes=64, groups=1, max_k=16)
    qs = na.Queries.generate(tt, 1, 24, seed=0x7172 + 89)
    big = na.Queries.generate(tt, 1, 2, seed=0x7172 + 97)
    with_destinations(big, tt.n_locations, np.random.default_rng(97), 65, 257)
    qs.q[3] = big.q[0]
    qs.q[8] = big.q[1]
    qs._keep.append(big)
    say("destinations beyond the limits")
    res = eng.search(qs)
    assert res[3]["status"] == na.EUNSUPPORTED and res[8]["status"] == na.EUNSUPPORTED
    check_ok(tt, qs, res, {3, 8})
