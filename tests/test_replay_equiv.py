"""The forward kernel's two local replays agree bit for bit.

csrc/eik_stencil.h replays the heap's acceptance sequence per candidate neighbour (replay_plain)
and, where no init neighbour, cascade key or "odd" axis is present, per axis (replay_axes).  The
header is compiled for the host (tests/host/hip/hip_runtime.h stands in for the HIP header) and
both replays run on random neighbourhoods; value, pop key and derivative codes must be identical.
The second test shows the comparison is sensitive: with the odd neighbourhoods included the two
replays differ.

The mixed precision's arithmetic (eval_mix: FP64 values and comparisons, the quadratic's increment
in FP32) is also checked against the oracle's restatement of it (og_stencil_mixed), bit for bit,
with the header and the oracle compiled together for the host; against the FP64 stencil it must
differ (the comparison would notice an arithmetic that silently fell back to FP64).
"""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(os.path.dirname(HERE), "eikonalsolvers.jl_amd", "csrc")


@pytest.fixture(scope="module")
def exe(tmp_path_factory):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ is not available")
    out = str(tmp_path_factory.mktemp("replay") / "replay_equiv")
    subprocess.run([gxx, "-O2", "-std=c++17", "-ffp-contract=off", "-w", "-I", os.path.join(HERE, "host"),
                    "-I", CSRC, "-o", out, os.path.join(HERE, "host", "replay_equiv.cpp")], check=True)
    return out


def test_axis_replay_equals_candidate_replay(exe):
    r = subprocess.run([exe, "regular", "200000"], capture_output=True, text=True, timeout=300)
    print(r.stdout)
    assert r.returncode == 0 and "mismatches: 0" in r.stdout, r.stdout + r.stderr


def test_odd_neighbourhoods_need_the_candidate_replay(exe):
    r = subprocess.run([exe, "all", "20000"], capture_output=True, text=True, timeout=300)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr


@pytest.fixture(scope="module")
def mix_exe(tmp_path_factory):
    gxx, gcc = shutil.which("g++"), shutil.which("gcc")
    if gxx is None or gcc is None:
        pytest.skip("g++/gcc not available")
    d = tmp_path_factory.mktemp("mix")
    ob = str(d / "eik_oracle.o")
    oracle_dir = os.path.join(os.path.dirname(HERE), "oracle")
    subprocess.run([gcc, "-O2", "-std=gnu11", "-ffp-contract=off", "-fno-fast-math", "-w", "-c",
                    os.path.join(oracle_dir, "eik_oracle.c"), "-o", ob], check=True)
    out = str(d / "mix_equiv")
    subprocess.run([gxx, "-O2", "-std=c++17", "-ffp-contract=off", "-w", "-I", os.path.join(HERE, "host"), "-I", CSRC,
                    "-I", oracle_dir, "-o", out, os.path.join(HERE, "host", "mix_equiv.cpp"), ob, "-lm"], check=True)
    return out


def test_mixed_stencil_equals_oracle_mixed_stencil(mix_exe):
    r = subprocess.run([mix_exe, "200000"], capture_output=True, text=True, timeout=300)
    print(r.stdout)
    assert r.returncode == 0 and "mismatches: 0" in r.stdout, r.stdout + r.stderr


def test_mixed_stencil_differs_from_fp64(mix_exe):
    r = subprocess.run([mix_exe, "20000", "fp64"], capture_output=True, text=True, timeout=300)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
