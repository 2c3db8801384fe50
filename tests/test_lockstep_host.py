"""The lockstep batch (damask_amd/csrc/batch.cpp) on the host: the file
is compiled unchanged against a host stand-in of the HIP runtime
(tests/lockstep_host) and driven by 32 threads over RVEs that leave the
lockstep at different rounds.  Built with ThreadSanitizer it must finish with
every batched launch reading the input of the round it was submitted for, every
RVE stream waiting for its launch, and no data race reported.  The same harness
with the stream waits dropped must fail, which shows that its checks bite."""
import os
import shutil
import subprocess

import pytest

from conftest import ROOT

HERE = os.path.join(ROOT, "tests", "lockstep_host")
SRC = [os.path.join(HERE, "harness.cpp"), os.path.join(ROOT, "damask_amd", "csrc", "batch.cpp")]
INC = ["-I", HERE, "-I", os.path.join(ROOT, "damask_amd", "csrc")]


def _build(out, extra):
    if shutil.which("g++") is None:
        pytest.skip("no g++")
    cmd = ["g++", "-std=c++17", "-O1", "-g", *extra, *INC, *SRC, "-o", out, "-lpthread"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, "host build of batch.cpp failed: " + r.stderr[-2000:]
    return out


@pytest.mark.parametrize("groups,wait_us", [("1", ""), ("2", ""), ("1", "30"), ("2", "0")])
def test_lockstep_batch_host_logic_under_thread_sanitizer(tmp_path, groups, wait_us):
    """one or two lockstep groups; rounds that wait for every RVE, or that
    start after wait_us microseconds without the RVEs not yet submitted"""
    exe = _build(str(tmp_path / "lockstep_tsan"), ["-fsanitize=thread"])
    env = dict(os.environ, TSAN_OPTIONS="halt_on_error=1", DMK_LOCKSTEP_GROUPS=groups, DMK_LOCKSTEP_WAIT_US=wait_us)
    r = subprocess.run([exe, "120", "32", "2"], capture_output=True, text=True, timeout=600, env=env)
    print(r.stdout[-2000:], r.stderr[-4000:])
    assert "ThreadSanitizer" not in r.stderr
    assert r.returncode == 0
    assert r.stdout.strip().endswith("ok")


def test_lockstep_timing_under_thread_sanitizer(tmp_path):
    """DMK_BATCH_TIMING: the per-round timing (events, host functions behind
    each round, the first submitter's return) is recorded without data races
    and printed per group"""
    exe = _build(str(tmp_path / "lockstep_tsan_timing"), ["-fsanitize=thread"])
    env = dict(os.environ, TSAN_OPTIONS="halt_on_error=1", DMK_LOCKSTEP_GROUPS="2", DMK_LOCKSTEP_WAIT_US="30",
               DMK_BATCH_TIMING="1")
    r = subprocess.run([exe, "120", "32", "2"], capture_output=True, text=True, timeout=600, env=env)
    print(r.stdout[-2000:], r.stderr[-4000:])
    assert "ThreadSanitizer" not in r.stderr
    assert r.returncode == 0
    assert r.stdout.strip().endswith("ok")
    assert "lockstep group 0 timing" in r.stderr and "lockstep group 1 timing" in r.stderr


def test_lockstep_harness_detects_missing_stream_waits(tmp_path):
    exe = _build(str(tmp_path / "lockstep_plain"), [])
    r = subprocess.run([exe, "120", "32", "2", "drop-waits"], capture_output=True, text=True, timeout=600)
    print(r.stdout[-2000:])
    assert r.returncode == 1 and "FAILED" in r.stdout
