"""The forward solver's fixed point, emulated on the host against the oracle's heap FMM.

tests/host/fixpoint_emul.cpp applies the device's local heap replay (eik_stencil.h) to every node
of radial-velocity 3D-spherical and smooth Cartesian 2D/3D models until nothing changes, and
compares the result bit for bit with og_fmm_record.

* Under random Gauss-Seidel, Jacobi, partial and torn-read orders the iteration reaches the heap's
  answer every time: the replay has one fixed point on these models, whatever the schedule.
* With the kernel's own wake-up rule (a node is re-evaluated only when a stencil neighbour changed
  with a key, old or new, not above its own) and reads racing writes, two things are needed: key
  pools that only ever hold a node's old or new key (written on every change, started as an
  unreached node's key), and a verification pass at convergence.  Both as in k_fim_round: every run
  reaches the heap's answer.  Without either, runs end off it (the emulator notices what the kernel
  guards against).
"""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(ROOT, "eikonalsolvers.jl_amd", "csrc")
ORACLE = os.path.join(ROOT, "oracle")


@pytest.fixture(scope="module")
def emul(tmp_path_factory):
    gxx, gcc = shutil.which("g++"), shutil.which("gcc")
    if gxx is None or gcc is None:
        pytest.skip("g++/gcc not available")
    d = tmp_path_factory.mktemp("fixpoint")
    ob = str(d / "eik_oracle.o")
    subprocess.run([gcc, "-O2", "-std=gnu11", "-ffp-contract=off", "-fno-fast-math", "-w", "-c",
                    os.path.join(ORACLE, "eik_oracle.c"), "-o", ob], check=True)
    out = str(d / "fixpoint_emul")
    subprocess.run([gxx, "-O2", "-std=c++17", "-ffp-contract=off", "-w", "-I", os.path.join(HERE, "host"), "-I", CSRC,
                    "-I", ORACLE, "-o", out, os.path.join(HERE, "host", "fixpoint_emul.cpp"), ob, "-lm"], check=True)
    return out


def _last(stdout):
    return [ln for ln in stdout.splitlines() if ln.startswith("runs ")][-1]


def test_replay_fixed_point_is_the_heap_under_any_order(emul):
    r = subprocess.run([emul, "2", "1", "orders"], capture_output=True, text=True, timeout=600)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "runs off the heap's answer 0" in _last(r.stdout)


def test_kernel_wakeup_rule_with_racing_reads_reaches_the_heap(emul):
    r = subprocess.run([emul, "3", "2", "filtered"], capture_output=True, text=True, timeout=600)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    last = _last(r.stdout)
    assert "runs off the heap's answer 0" in last
    # the unguarded variants (no verification, or cascade-only key pools) do end off it
    unguarded = int(last.rsplit(" ", 1)[1])
    assert unguarded >= 1
