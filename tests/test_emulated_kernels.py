"""The HIP kernels run on the CPU (tests/emu: the kernel sources compiled by g++
against a host shim, one thread per lane, LDS filled with garbage, AddressSanitizer
and UBSan on) and compared with the oracle: counters, 5-set histogram and result
lines.  CPU only: this checks the device code's logic and its memory accesses
where no GPU is present; the GPU parity tests check the same on the MI355X."""
import os
import shutil
import subprocess

import pytest

import ps4

EMU_DIR = os.path.join(ps4.REPO_DIR, "tests", "emu")
EMU_BIN = os.path.join(EMU_DIR, "_build", "ps4_emu")

pytestmark = pytest.mark.skipif(shutil.which("g++") is None or shutil.which("make") is None,
                                reason="needs g++ and make")


@pytest.fixture(scope="module")
def emu():
    r = subprocess.run(["make", "-C", EMU_DIR, "-j8"], capture_output=True, text=True, timeout=900)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    assert os.access(EMU_BIN, os.X_OK)
    return EMU_BIN


def run(emu, *args, timeout=600):
    r = subprocess.run([emu, *map(str, args)], capture_output=True, text=True, timeout=timeout)
    out = r.stdout + r.stderr
    assert r.returncode == 0, out[-3000:]
    assert "MISMATCH" not in out and "ERROR" not in out, out[-3000:]
    return out


def test_emulated_census_small_s(emu):
    """Every 4-set, extension and line of [1, 3*10^5) (2,907 4-sets)."""
    out = run(emu, 1, 300000)
    assert "[1,300000) OK 4-sets=2907" in out


def test_emulated_batch_mixed_intervals(emu):
    """One batched call over ranges whose tiles mix intervals: the first S with a
    4-set, windows near 10^9 and 9*10^11 (the extension kernel's larger lists
    and the late residue masks at high S)."""
    out = run(emu, "--batch", 4205, 4206, 1000000000, 1000020000, 900000000000, 900000005000)
    assert out.count(" OK ") == 3


def test_emulated_triple_query(emu):
    """The triple query on the first 4-set's triple {482, 359, 2}: its fourth
    member 3362 (d <= 10^6)."""
    out = run(emu, "--triple", 2, 359, 482, 1000000)
    assert "triple OK" in out
