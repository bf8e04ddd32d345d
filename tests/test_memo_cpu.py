"""CPU tests of the autotuning journal (chai_hip_memo_*), restating the
reference's AutoTuneMemo state machine (autotune/AutoTuneMemo.cpp:16-188)."""
import pytest

import chai_amd as chai

PS, CK, SA, ST, SP, GO, BO, BC, AC = -1, -2, -3, -4, -5, -6, -7, -100, -101


@pytest.fixture
def journal(tmp_path):
    chai.memo_close()
    p = tmp_path / "journal0.AMD"
    chai.memo_open(p, 0)
    yield p
    chai.memo_close()


def good_run(k, us):
    for s in (PS, CK, SA, ST, SP):
        chai.memo_record(k, s)
    chai.memo_record(k, GO)
    chai.memo_record(k, us)


def test_good_kernel_journaled(journal):
    good_run("k_good", 1234)
    assert chai.memo_status("k_good") == (1, 1234)
    lines = journal.read_text().split("\n")
    assert lines[0] == "0 k_good -1" and "0 k_good 1234" in lines


def test_bad_transition_blacklists(journal):
    chai.memo_record("k_skip", PS)
    chai.memo_record("k_skip", ST)           # skipped COMPILE / SET_ARGS
    assert chai.memo_status("k_skip")[0] == -1
    chai.memo_record("k_first", CK)          # first state must be PRINT_SOURCE
    assert chai.memo_status("k_first")[0] == -1
    good_run("k_bad", 10)
    chai.memo_record("k_bad", BO)
    assert chai.memo_status("k_bad")[0] == -1


@pytest.mark.parametrize("last", [CK, ST, BC])
def test_crash_states_blacklisted_on_reload(journal, last):
    """A process that died in COMPILE_KERNEL / START_TIMER / BEFORE_CHECK leaves
    that state last in the journal; reloading marks the kernel bad."""
    seq = {CK: [PS, CK], ST: [PS, CK, SA, ST], BC: [PS, CK, SA, ST, SP, GO, BC]}[last]
    for s in seq:
        chai.memo_record("k_crash", s)
    good_run("k_ok", 77)
    chai.memo_close()                        # "crash"
    chai.memo_open(journal, 0)
    assert chai.memo_status("k_crash")[0] == -1
    assert chai.memo_status("k_ok") == (1, 77)


def test_check_states_are_valid(journal):
    good_run("k_chk", 5)
    chai.memo_record("k_chk", BC)
    chai.memo_record("k_chk", AC)
    assert chai.memo_status("k_chk") == (1, 5)


def test_other_device_lines_ignored(journal, tmp_path):
    chai.memo_close()
    p = tmp_path / "j2"
    p.write_text("1 k_dev1 -1\n1 k_dev1 -2\n0 k_dev0 -1\n")
    chai.memo_open(p, 0)
    assert chai.memo_status("k_dev1")[0] == 0
    assert chai.memo_status("k_dev0")[0] == 0


def test_journal_reload_roundtrip(journal):
    good_run("k_a", 11)
    good_run("k_b", 22)
    chai.memo_close()
    chai.memo_open(journal, 0)
    assert chai.memo_status("k_a") == (1, 11) and chai.memo_status("k_b") == (1, 22)


def test_saved_memo_reload_matches_reference_quirk(journal, tmp_path):
    """saveMemoFile writes each kernel's final state then its times
    (AutoTuneMemo.cpp:202-243); loading that file starts a kernel at a state
    other than PRINT_SOURCE, which updateState (AutoTuneMemo.cpp:27-31) treats
    as bad -- kept for parity: a saved memo reloads with every kernel bad."""
    good_run("k_a", 11)
    out = tmp_path / "saved"
    chai.memo_save(out)
    assert out.read_text().split("\n")[:2] == ["0 k_a -6", "0 k_a 11"]
    chai.memo_close()
    chai.memo_open(out, 0)
    assert chai.memo_status("k_a")[0] == -1
