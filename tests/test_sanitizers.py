"""The oracle and the library's host sources (timetable, generator, GTFS
loader, time zones, contraction hierarchy, query preparation, journey legs,
host C ABI) built with -fsanitize=address,undefined -- host code only, no GPU
code -- and run on generated workloads 1-4 and on the GTFS feed of nigiri's
transfer-time test: the counterpart of nigiri's linux-sanitizer preset
(CMakePresets.json, CI sanitizer job).  The run also checks that the host
contraction-hierarchy lower bounds equal the oracle's Dijkstra."""
import datetime
import os
import subprocess

import nigiri_amd as na
from golden import gtfs_routing as G

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE = os.path.join(ROOT, "oracle")


def _days(ymd):
    return (datetime.date(*ymd) - datetime.date(1970, 1, 1)).days


def test_host_code_under_address_and_undefined_behaviour_sanitizers(tmp_path):
    subprocess.run(["make", "-C", ORACLE, "-j8", "sanitize"], check=True, stdout=subprocess.PIPE,
                   stderr=subprocess.STDOUT)
    for name, text in na.parse_mem_dir(G.TRANSFER_TIME_FEED).items():
        (tmp_path / name).write_text(text)
    env = dict(os.environ)
    try:
        import tzdata
        env.setdefault("TZDIR", os.path.join(os.path.dirname(tzdata.__file__), "zoneinfo"))
    except ImportError:
        pass
    d0, d1 = (_days(x) for x in G.TRANSFER_TIME_RANGE)
    r = subprocess.run([os.path.join(ORACLE, "_build", "sanitize_check"), str(tmp_path), str(d0), str(d1),
                        "A", "C"], stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True, env=env,
                       timeout=900)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-4000:]
    assert "sanitize_check: ok" in r.stdout and "gtfs feed" in r.stdout, r.stdout
