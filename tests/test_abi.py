"""C-ABI surface and host-side constants -- CPU only, no device calls."""
import ctypes
import os
import re

import pytest

import ps4

REPO = ps4.REPO_DIR
HEADER = os.path.join(REPO, "include", "ps4.h")
KERNELS = os.path.join(ps4.PKG_DIR, "csrc", "ps4_kernels.hip")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"^[A-Za-z_][\w \*]*?\b(ps4_\w+)\s*\(", text, flags=re.M)
    return sorted(set(names))


def test_header_declares_the_entry_points():
    names = declared_functions()
    for must in ("ps4_open", "ps4_search", "ps4_search_batch", "ps4_extend_triple", "ps4_close",
                 "ps4_strerror", "ps4_last_error"):
        assert must in names


def test_library_loads_and_exports_every_declared_symbol():
    lib = ctypes.CDLL(ps4.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_strerror_texts():
    lib = ps4.load()
    for rc in (0, -1, -2, -3, -4, -5, -6, -7):
        assert lib.ps4_strerror(rc)
    assert b"NOT certified" in lib.ps4_strerror(-2)


def test_cli_binary_present():
    assert os.access(ps4.CLI_PATH, os.X_OK)


def test_struct_layouts_match_header():
    # ps4_counters: 10 + 100 + 3 u64; ps4_config as declared
    assert ctypes.sizeof(ps4.Counters) == 113 * 8
    assert ps4.Config.S_hi_max.offset == 8 and ps4.Config.dump4_limit.offset == 24


@pytest.mark.parametrize("name,m,shift,lim", [("mod63", 63, 20, 4095), ("mod65", 65, 20, 4095),
                                              ("mod11", 11, 19, 46189), ("mod13", 13, 19, 46189),
                                              ("mod17", 17, 19, 46189), ("mod19", 19, 20, 46189)])
def test_kernel_residue_constants_are_exact(name, m, shift, lim):
    """The multiply-shift divisions in the extension kernel equal true division
    over their whole input range (inputs are residues below lim)."""
    src = open(KERNELS).read()
    mt = re.search(name + r"\(u32 s\) \{ return s - (\d+)u \* \(\(s \* (\d+)u\) >> (\d+)\); \}", src)
    assert mt, name
    mm, magic, sh = int(mt.group(1)), int(mt.group(2)), int(mt.group(3))
    assert mm == m and sh == shift and magic < 2**24
    for s in range(lim):
        assert s * magic < 2**32
        assert (s * magic) >> sh == s // m


def test_fold4095_is_exact():
    for r in range(0, 262080):
        s = (r >> 12) + (r & 4095)
        s = s - 4095 if s >= 4095 else s
        assert s == r % 4095


def test_codegen_has_no_signed_shift_of_unsigned_products(tmp_path):
    """Regression guard for a bring-up fault: the gfx950 code of the kernels must
    not shift a multiply result arithmetically (with __umul24 the compiler emitted
    v_mul_u32_u24 + v_ashrrev_i32, turning residues >= 2^31/magic into indices
    ~90000 and reading far past the residue tables)."""
    import shutil
    import subprocess
    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(hipcc):
        pytest.skip("hipcc not available")
    out = tmp_path / "k.s"
    src = os.path.join(ps4.PKG_DIR, "csrc", "ps4_kernels.hip")
    subprocess.run([hipcc, "--offload-arch=gfx950", "-O3", "-std=c++17", "-Xclang", "-target-feature", "-Xclang",
                    "-packed-fp32-ops", "-I" + os.path.join(ps4.PKG_DIR, "csrc"), "-I" + os.path.join(REPO, "include"),
                    "--cuda-device-only", "-S", src, "-o", str(out)], check=True, stdout=subprocess.PIPE,
                   stderr=subprocess.PIPE)
    asm = out.read_text()
    assert "v_ashrrev_i32" not in asm
    assert "__umul24(" not in open(src).read()


def test_division_magic_of_the_candidate_loop_is_exact():
    """t / nL = ((t << 9) * ceil(2^23 / nL)) >> 32 for every t < 8192 and
    1 <= nL < 1024 (the extension kernel's index of a candidate in its lists)."""
    import numpy as np
    t = np.arange(8192, dtype=np.uint64)
    for d in range(1, 1024):
        m = (0x800000 + d - 1) // d
        assert m < (1 << 24)
        assert np.array_equal(((t << np.uint64(9)) * np.uint64(m)) >> np.uint64(32), t // np.uint64(d)), d
