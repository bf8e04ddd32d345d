"""The MI355X device inside the reference's own runtime, on the CPU.

Chai programs written against the reference's API (tests/frontend/) and the
reference's sample programs run on the reference's scheduler and interpreter,
compiled here from its sources (oracle/frontend.mk).  Three runs of each:

- interpreter only: the reference's own result;
- with the MI355X device over the CPU oracle (tests/frontend/oracle_backend.c):
  every VectorTrace the scheduler hands the device is lowered to a chai_trace
  (chai_amd/device/DeviceMI355X.cpp) and evaluated by the oracle.  Equal
  results pin the lowering, and the oracle's restatement of the interpreter,
  against the interpreter itself;
- with the device but no library: every trace the device is given fails and
  the scheduler re-queues it on the interpreter (runtime/Scheduler.cpp:
  221-254), with the interpreter's results.
"""
import os

import numpy as np
import pytest

from tests import fe_programs as fe

pytestmark = pytest.mark.skipif(not fe.built(), reason="reference-API programs not built (needs the reference tree)")


def _refused(stderr):
    return [l for l in stderr.splitlines() if "refused" in l or "failed on the device" in l]


@pytest.mark.parametrize("case", fe.CASES)
def test_lowering_over_oracle_equals_interpreter(case, tmp_path):
    ref, _ = fe.run_case(case, str(tmp_path / "interp"))
    got, err = fe.run_case(case, str(tmp_path / "dev"), device="oracle")
    assert not _refused(err), err
    for name in ref:
        assert np.array_equal(ref[name].view(np.uint8), got[name].view(np.uint8)), (case, name)


@pytest.mark.parametrize("case", ["multi", "mm_int", "red"])
def test_device_failure_falls_back_to_interpreter(case, tmp_path):
    ref, _ = fe.run_case(case, str(tmp_path / "interp"))
    got, err = fe.run_case(case, str(tmp_path / "dev"), device="missing")
    assert "not available" in err
    for name in ref:
        assert np.array_equal(ref[name].view(np.uint8), got[name].view(np.uint8)), (case, name)


@pytest.mark.parametrize("sample", [s for s in fe.SAMPLES if s not in fe.RNG_SAMPLES])
def test_reference_samples_over_oracle(sample):
    ref, _ = fe.run_sample(sample)
    got, err = fe.run_sample(sample, device="oracle")
    assert not _refused(err), err
    assert ref == got and ref.strip()


def test_rng_samples_over_oracle():
    """pi.cpp and monte.cpp draw Philox streams (RNG_DEFAULT = RNG_PHILOX).
    The interpreter's values change from run to run, so its output is no
    reference; the device's (over the oracle) is the same on every run, and
    pi.cpp's estimate, the mean of 10^6 Bernoulli draws with p = pi / 4, is
    within 5 standard deviations (5 * 4 sqrt(p (1 - p) / 10^6)) of pi."""
    a, err = fe.run_sample("pi", device="oracle")
    b, _ = fe.run_sample("pi", device="oracle")
    assert not _refused(err), err
    assert a == b
    est = float(a.split("=")[1])
    p = np.pi / 4
    assert abs(est - np.pi) <= 5 * 4 * np.sqrt(p * (1 - p) / 1e6)
    m1, err = fe.run_sample("monte", device="oracle")
    m2, _ = fe.run_sample("monte", device="oracle")
    assert not _refused(err), err
    assert m1 == m2 and m1.startswith("mean is ")
