"""CPU: the serving loop's admission rule and footprint estimate
(gpupi_serve.h), called through the C ABI with no device work."""
import numpy as np
import pytest

import gpupi_amd as g
from helpers import run_jobs

GB = 10 ** 9
PI, MUL = g.JOB_PI, g.JOB_MUL


def test_fcfs_within_budget():
    # everything fits: admitted in arrival order
    assert g.serve_admit([1 * GB, 2 * GB, 3 * GB], [MUL, PI, MUL], 200 * GB, 0, 4) == [0, 1, 2]


def test_head_of_line_waits_for_memory():
    # the head does not fit: nothing behind it may pass (no starvation of big jobs)
    assert g.serve_admit([150 * GB, 1 * GB], [MUL, MUL], 200 * GB, 60 * GB, 4) == []
    # it fits once memory is released
    assert g.serve_admit([150 * GB, 1 * GB], [MUL, MUL], 200 * GB, 40 * GB, 4) == [0, 1]
    # budget is consumed in order: the third job waits
    assert g.serve_admit([90 * GB, 90 * GB, 30 * GB, 1 * GB], [MUL] * 4, 200 * GB, 0, 4) == [0, 1]


def test_pi_without_worker_lets_later_jobs_pass():
    assert g.serve_admit([1 * GB, 1 * GB, 1 * GB], [PI, MUL, PI], 200 * GB, 0, 0) == [1]
    assert g.serve_admit([1 * GB, 1 * GB, 1 * GB, 1 * GB], [PI, PI, MUL, PI], 200 * GB, 0, 1) == [0, 2]


def test_empty_queue():
    assert g.serve_admit([], [], 200 * GB, 0, 4) == []


def test_footprint_estimates():
    lib = g.lib()
    # multiply: operands + product + ragged NTT scratch, monotone in size
    small = lib.gpupi_serve_job_bytes(MUL, 4096, 4096)
    big = lib.gpupi_serve_job_bytes(MUL, 1 << 26, 1 << 26)
    assert small < big
    assert big >= 8 * (2 << 26) + 32 * (1 << 27)  # four planes of the 2^27 transform at least
    assert big < 10 * GB
    # pi: the largest job of the serving mix (10^8 digits) fits the budget many times over
    assert lib.gpupi_serve_job_bytes(PI, 10 ** 8, 0) < 5 * GB
    assert lib.gpupi_serve_job_bytes(PI, 10 ** 5, 0) < lib.gpupi_serve_job_bytes(PI, 10 ** 8, 0)


def test_server_threading_simulation(tmp_path):
    """serve.hip's threads, queues and admission run against host stand-ins
    for HIP and the library (tests/native/serve_sim.cpp): random mixes and a
    one-job-at-a-time budget; every job finishes once with its own result,
    reservations never exceed the budget, no stall."""
    import os
    import subprocess
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    exe = tmp_path / "serve_sim"
    subprocess.run(["g++", "-O1", "-std=c++17", "-x", "c++", "-I", os.path.join(root, "tests", "native", "stub"),
                    "-I", os.path.join(root, "include"), os.path.join(root, "tests", "native", "serve_sim.cpp"),
                    "-o", str(exe), "-pthread"], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    assert "serve sim: ok" in out.stdout


class _FakeServer:
    """Returns every submitted job's result at the first results() call."""

    class _Result:
        def __init__(self, jid):
            self.id, self.status = jid, 0

    def __init__(self, lose: int = 0):
        self.next, self.queued, self.lose = 0, [], lose

    def _submit(self):
        self.queued.append(self.next)
        self.next += 1
        return self.next - 1

    def submit_pi(self, digits, keep_digits=False):
        return self._submit()

    def submit_mul(self, A, B):
        return self._submit()

    def results(self, wait=True):
        out = [(self._Result(i), None) for i in self.queued[self.lose:]]
        self.queued = []
        return out


def test_run_jobs_returns_once_every_result_is_in():
    a, b = np.zeros(5, np.uint32), np.zeros(7, np.uint32)
    out = run_jobs(_FakeServer(), [1000, 2000], [(a, b), (b, a)])
    assert sorted(out) == [("mul", (5, 7)), ("mul", (7, 5)), ("pi", 1000), ("pi", 2000)]


def test_run_jobs_refuses_jobs_of_one_shape():
    # jobs of one shape would share one key: the helper refuses them instead
    # of waiting for a count of keys that can never be reached
    a = np.zeros(5, np.uint32)
    with pytest.raises(AssertionError, match="one shape"):
        run_jobs(_FakeServer(), [], [(a, a), (a, a)])


def test_run_jobs_stops_when_the_server_has_nothing_left():
    a, b = np.zeros(5, np.uint32), np.zeros(7, np.uint32)
    with pytest.raises(AssertionError, match="no result"):
        run_jobs(_FakeServer(lose=1), [], [(a, b), (b, a)])
