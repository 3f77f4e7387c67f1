"""CPU: the C-ABI library loads, exports every entry point include/*.h
declares, and does not link the oracle or GMP; the CLI rejects bad
arguments before touching a device."""
import ctypes
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "gpupi_amd", "libgpupi.so")
CLI = os.path.join(ROOT, "gpupi_amd", "pi_bs_gpu")
CXX_ONLY = {"compact"}  # BatchMPArray::compact, an inline C++ member


def declared_functions():
    names = set()
    for h in sorted(os.listdir(os.path.join(ROOT, "include"))):
        if not h.endswith(".h"):
            continue
        text = open(os.path.join(ROOT, "include", h)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        text = re.sub(r"//[^\n]*", "", text)
        for m in re.finditer(r"^\s*(?:[A-Za-z_][\w]*[\s\*]+)+\**([A-Za-z_]\w*)\s*\(([^;{}]*)\)\s*;", text, re.M):
            if m.group(1) not in CXX_ONLY and not m.group(0).lstrip().startswith(("typedef", "return")):
                names.add(m.group(1))
    return names


def test_headers_declare_the_reference_surface():
    names = declared_functions()
    for required in ["batch_mp_init", "batch_mp_mul", "batch_mp_shift_add", "batch_mp_exactdiv_small",
                     "batch_mul_ntt", "batch_mul_naive", "batch_mul_small", "convert_decimal_equal_split",
                     "convert_decimal_precompute_powers", "batch_decimal_naive", "batch_shift_bits",
                     "batch_limblength_max", "gpupi_pi_run"]:
        assert required in names


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(LIB)
    missing = [n for n in sorted(declared_functions()) if not hasattr(lib, n)]
    assert not missing, missing


def test_launch_shape_setters_validate_without_a_device():
    # process-wide launch parameters: range checks and read-back (no device call)
    lib = ctypes.CDLL(LIB)
    lib.gpupi_ntt_set_launch.argtypes = [ctypes.c_uint32, ctypes.c_int]
    lib.gpupi_group_stage_set_shape.argtypes = [ctypes.c_uint32]
    lib.gpupi_group_stage_get_shape.restype = ctypes.c_uint32
    m, c = ctypes.c_uint32(), ctypes.c_int()
    lib.gpupi_ntt_get_launch(ctypes.byref(m), ctypes.byref(c))
    assert (m.value, c.value) == (5, 5)
    assert lib.gpupi_group_stage_get_shape() == 4
    for mask, chunk in ((16, 5), (5, 9), (5, -1)):
        assert lib.gpupi_ntt_set_launch(mask, chunk) != 0
    for gpb in (0, 3, 32):
        assert lib.gpupi_group_stage_set_shape(gpb) != 0
    assert lib.gpupi_ntt_set_launch(15, 0) == 0 and lib.gpupi_group_stage_set_shape(16) == 0
    lib.gpupi_ntt_get_launch(ctypes.byref(m), ctypes.byref(c))
    assert (m.value, c.value) == (15, 0) and lib.gpupi_group_stage_get_shape() == 16
    assert lib.gpupi_ntt_set_launch(5, 5) == 0 and lib.gpupi_group_stage_set_shape(4) == 0


def test_library_does_not_link_the_oracle():
    out = subprocess.run(["readelf", "-d", LIB], capture_output=True, text=True, check=True).stdout
    needed = re.findall(r"\(NEEDED\).*\[(.*)\]", out)
    assert needed and not any("gmp" in n or "oracle" in n for n in needed), needed


@pytest.mark.parametrize("args", [[], ["abc"], ["10", "--bogus"], ["-5"], ["10", "--benchmark", "--full-print"],
                                  ["10", "--verify-bbp"], ["10", "--verify-bbp", "0"], ["10", "--verify-bbp", "x"],
                                  ["10", "--verify-bbp", "600000000"], ["10", "--output"],
                                  ["10", "--output", "f", "--benchmark"]])
def test_cli_rejects_bad_arguments(args):
    r = subprocess.run([CLI] + args, capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "Usage" in r.stderr


def test_ntt_split_table_validates_and_restores():
    # host-only calls: the run-time NTT level table (no device work)
    lib = ctypes.CDLL(LIB)
    lib.gpupi_ntt_set_split.argtypes = [ctypes.c_int] * 3
    lib.gpupi_ntt_get_split.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_int)]
    m = (ctypes.c_int * 3)()

    def levels(k):
        n = lib.gpupi_ntt_get_split(k, m)
        return tuple(m[i] for i in range(n))

    assert levels(10) == (10,)
    assert levels(20) == (8, 12)
    assert levels(28) == (8, 8, 12)
    assert levels(31) == () and levels(3) == ()
    # two levels: column 2^5..2^9 and contiguous 2^6..2^12
    assert lib.gpupi_ntt_set_split(20, 9, 0) == 0 and levels(20) == (9, 11)
    for bad in [(20, 7, 0), (20, 4, 0), (20, 10, 0), (20, 8, 5), (10, 5, 0), (28, 10, 6), (28, 5, 5),
                (28, 9, 3), (31, 8, 8), (20, 0, 3), (20, -1, 0)]:
        assert lib.gpupi_ntt_set_split(*bad) != 0, bad
    assert levels(20) == (9, 11)
    # three levels: m1 = 0 leaves a 4096-point contiguous level
    assert lib.gpupi_ntt_set_split(28, 7, 0) == 0 and levels(28) == (7, 9, 12)
    assert lib.gpupi_ntt_set_split(28, 9, 10) != 0 and lib.gpupi_ntt_set_split(28, 8, 5) != 0
    assert lib.gpupi_ntt_set_split(28, 9, 8) == 0 and levels(28) == (9, 8, 11)
    assert lib.gpupi_ntt_set_split(20, 0, 0) == 0 and levels(20) == (8, 12)
    assert lib.gpupi_ntt_set_split(28, 0, 0) == 0 and levels(28) == (8, 8, 12)


def test_tuner_enumerates_exactly_the_accepted_splits():
    import sys
    sys.path.insert(0, ROOT)
    from gpupi_amd import tune
    lib = ctypes.CDLL(LIB)
    lib.gpupi_ntt_set_split.argtypes = [ctypes.c_int] * 3
    for k in range(13, 31):
        accepted = set()
        for m0 in range(1, 13):
            for m1 in ([0] if k <= 21 else range(1, 13)):
                if lib.gpupi_ntt_set_split(k, m0, m1) == 0:
                    accepted.add((m0, m1))
        assert lib.gpupi_ntt_set_split(k, 0, 0) == 0
        assert accepted == set(tune.splits(k)), k
