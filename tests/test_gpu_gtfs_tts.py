"""GPU parity for the path's one float operation (non-default transfer-time
settings, transfer_time_settings.h:21-31), query::max_interval_ and the
search timeout: the GTFS fixtures of nigiri's tests and generated workloads
with every settings combination, on the batched and the plain engine,
against the golden journeys and the CPU oracle (journeys, legs, state)."""
import pytest

import nigiri_amd as na
import oracle
from golden import gtfs_routing as G
from test_gpu_parity import compare, image_for, say
from test_oracle_gtfs import (TTS_CASES, day_minutes, hm, max_interval_query, max_interval_tt,
                              named_legs, transfer_time_tt, tts_query)

pytestmark = pytest.mark.gpu

ENGINES = [(64, 1, 16), (1, 2, 16)]  # (lanes, groups, max_k)

# (tts_min_transfer_time, tts_additional_time, tts_factor): each term alone,
# pairs, all three, factors that round down (x * 1.3, x * 2.5) and a factor
# below 1, which search::init raises to 1 (search.h:102-108)
SETTINGS = [(0, 0, 1.5), (10, 0, 1.0), (0, 2, 1.0), (3, 1, 1.0), (3, 5, 2.5), (0, 1, 1.3),
            (4, 0, 2.0), (0, 0, 0.5)]


@pytest.mark.parametrize("lanes,groups,max_k", ENGINES)
def test_gpu_gtfs_transfer_time_settings_golden(lanes, groups, max_k):
    tt = transfer_time_tt()
    img = na.Image(tt)
    eng = na.Engine(img, lanes=lanes, groups=groups, max_k=max_k)
    day0 = day_minutes(G.DATE)
    for i, (fields, fwd, expected) in enumerate(TTS_CASES):
        say(f"gtfs transfer-time case {i} fwd={fwd} {fields} lanes={lanes}")
        qs = na.Queries.build([tts_query(tt, fields, fwd)])
        res = compare(tt, eng, qs)
        assert [named_legs(tt, l, day0) for l in res[0]["legs"]] == expected


@pytest.mark.parametrize("lanes,groups,max_k", ENGINES)
def test_gpu_gtfs_max_interval_and_timeout(lanes, groups, max_k):
    tt = max_interval_tt()
    img = na.Image(tt)
    eng = na.Engine(img, lanes=lanes, groups=groups, max_k=max_k)
    day0 = day_minutes(G.DATE)
    say(f"gtfs max_interval lanes={lanes}")
    res = compare(tt, eng, na.Queries.build([max_interval_query(tt)]))
    assert [named_legs(tt, l, day0) for l in res[0]["legs"]] == G.MAX_INTERVAL_EXPECTED
    base = dict(interval=("08:00", "09:00"), min_connection_count=4, max_interval=None)
    for extra in ({}, {"timeout_s": 0}, {"timeout_s": 3600}):
        say(f"gtfs timeout {extra} lanes={lanes}")
        q = na.Queries.build([max_interval_query(tt, **base, **extra)])
        got = eng.search(q)[0]
        ref = oracle.search(tt, q[0])
        assert got["journeys"] == ref["journeys"] and got["legs"] == ref["legs"]
        assert (got["interval_from"], got["interval_to"]) == ref["interval"]
        if extra.get("timeout_s") == 0:
            assert ref["interval"] == (day0 + hm("08:00"), day0 + hm("14:00"))


def with_settings(qs, mtt, add, factor):
    for i in range(len(qs)):
        q = qs[i]
        q.tts_default = 0
        q.tts_min_transfer_time = mtt
        q.tts_additional_time = add
        q.tts_factor = factor
    return qs


@pytest.mark.parametrize("lanes,groups,max_k", ENGINES)
def test_gpu_transfer_time_settings_workload1(lanes, groups, max_k):
    tt, img = image_for("w1", lambda: na.Timetable.generate(1))
    eng = na.Engine(img, lanes=lanes, groups=groups, max_k=max_k)
    n = 64 if lanes == 64 else 2
    for j, (mtt, add, factor) in enumerate(SETTINGS):
        say(f"workload1 settings {(mtt, add, factor)} lanes={lanes} n={n}")
        qs = with_settings(na.Queries.generate(tt, 1, n, seed=0x7172 + 100 + j), mtt, add, factor)
        compare(tt, eng, qs)


@pytest.mark.parametrize("lanes,groups,max_k", [(64, 2, 10), (1, 2, 10)])
def test_gpu_transfer_time_settings_workload3(lanes, groups, max_k):
    tt, img = image_for("w3s", lambda: na.Timetable.generate(3, scale=0.03))
    eng = na.Engine(img, lanes=lanes, groups=groups, max_k=max_k)
    n = 96 if lanes == 64 else 4
    for j, (mtt, add, factor) in enumerate(SETTINGS[:5]):
        say(f"workload3 settings {(mtt, add, factor)} lanes={lanes} n={n}")
        qs = with_settings(na.Queries.generate(tt, 3, n, seed=0x7172 + 200 + j), mtt, add, factor)
        compare(tt, eng, qs)
