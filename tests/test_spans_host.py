"""The busy time of a call's round launches (eik_last_stats [17] and [18], the bench line's
`achieved_device`) is the length of the union of the launches' spans over one or two streams
(eik_spans.h).  Host build of that function: hand-computed cases (disjoint, overlapping, nested,
touching, unsorted, reversed and empty spans) and random integer intervals against a brute-force
count of the unit cells they cover."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(os.path.dirname(HERE), "eikonalsolvers.jl_amd", "csrc")


def test_span_union_matches_cell_count(tmp_path):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ is not available")
    exe = str(tmp_path / "spans_test")
    subprocess.run([gxx, "-O2", "-std=c++17", "-I", CSRC, "-o", exe, os.path.join(HERE, "host", "spans_test.cpp")],
                   check=True)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0 and r.stdout.strip() == "ok", r.stdout + r.stderr
