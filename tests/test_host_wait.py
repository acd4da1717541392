"""The bounded host wait for device-published level signals (mpsboost_amd/csrc/host_wait.hpp),
checked on the CPU with threads standing in for the device (tests/native/host_wait_check.cpp):

* a word that never reaches the sequence number ends the wait at its deadline, unpublished;
* the poll callback (where the library checks the stream for errors) can end the wait;
* 8 fits waiting level by level all finish, and so do 16 waiting fits on a process limited to
  2 cores while the threads that publish need those cores: past the first tens of
  microseconds a waiter yields and then sleeps instead of holding a core."""

import json
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def results(tmp_path_factory):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not available")
    exe = str(tmp_path_factory.mktemp("wait") / "host_wait_check")
    subprocess.run([gxx, "-O2", "-std=c++17", "-pthread", "-I", os.path.join(ROOT, "mpsboost_amd", "csrc"),
                    os.path.join(ROOT, "tests", "native", "host_wait_check.cpp"), "-o", exe], check=True)
    out = subprocess.run([exe], check=True, capture_output=True, text=True, timeout=120).stdout
    lines = [json.loads(line) for line in out.splitlines() if line.strip()]
    return {(r["scenario"], r.get("cores")): r for r in lines}


def test_published_word_returns_at_once(results):
    r = results[("already", None)]
    assert r["published"] == 1 and r["calls"] == 0


def test_deadline_ends_an_unpublished_wait(results):
    r = results[("deadline", None)]
    assert r["published"] == 0 and r["stopped"] == 0 and r["seen"] == 3
    assert 0.2 <= r["seconds"] < 2.0


def test_poll_error_ends_the_wait(results):
    r = results[("poll_stops", None)]
    assert r["published"] == 0 and r["stopped"] == 7 and r["calls"] == 3


def test_concurrent_waiters_finish(results):
    r = results[("many_waiters", 0)]
    assert r["done"] == r["waiters"] == 8 and r["failed"] == 0


def test_waiters_do_not_starve_publishers_on_few_cores(results):
    r = results[("many_waiters", 2)]
    assert r["done"] == r["waiters"] == 16 and r["failed"] == 0
    # 16 producers x 20 levels x 2 ms of busy work on 2 cores needs >= 0.32 s of wall time.
    # Waiters that held their cores would take about as much CPU as the producers (0.64 s,
    # their fair share of the 2 cores); these take well under half of it.
    assert r["waiter_cpu_s"] < 0.5 * (16 * 20 * 0.002), r
