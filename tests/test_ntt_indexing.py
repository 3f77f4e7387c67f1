"""CPU: the NTT's addressing (gpupi_amd/csrc/ntt_plan.hpp, the same header the
kernels include) keeps every access in bounds and covers each transform level
exactly once; checked for every transform size 2^4..2^30 (exhaustively up to
2^23, sampled tiles above)."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_ntt_index_math(tmp_path):
    exe = tmp_path / "ntt_index_check"
    subprocess.run(["g++", "-O2", "-std=c++17", "-I", os.path.join(ROOT, "gpupi_amd", "csrc"),
                    os.path.join(ROOT, "tests", "native", "ntt_index_check.cpp"), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=600)
    assert out.returncode == 0, out.stdout + out.stderr
    assert "index check: ok" in out.stdout


def test_ntt_arith_bounds(tmp_path):
    """The butterflies' re-centring plans (gpupi_amd/csrc/ntt_arith.hpp) keep
    every value exact and within its bound: each butterfly kind, radix and
    prime on all sign patterns at its input bounds and on random inputs,
    against 128-bit modular arithmetic."""
    exe = tmp_path / "ntt_bounds_check"
    subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-I", os.path.join(ROOT, "gpupi_amd", "csrc"),
                    os.path.join(ROOT, "tests", "native", "ntt_bounds_check.cpp"), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=600)
    assert out.returncode == 0, out.stdout + out.stderr
    assert "bounds check: ok" in out.stdout


def test_group_exact_division_step(tmp_path):
    """The 17-term group stage's exact division by 13^2 17^3 7^6 11^3
    (gpupi_amd/csrc/exact_div.hpp): the short carry step equals the 128-bit
    step on random and edge inputs, and whole divisions of multiples of the
    divisor return the multiplier."""
    exe = tmp_path / "hensel_step_check"
    subprocess.run(["g++", "-O2", "-std=c++17", "-I", os.path.join(ROOT, "gpupi_amd", "csrc"),
                    os.path.join(ROOT, "tests", "native", "hensel_step_check.cpp"), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=600)
    assert out.returncode == 0, out.stdout + out.stderr
    assert "hensel check: ok" in out.stdout


def test_ragged_assembly_rules(tmp_path):
    """Products run in pieces (gpupi_amd/csrc/ragged.hpp, the rules the
    ragged launches and the assembly kernel use): the pieces tile the longer
    operand, every staged word the assembly reads lies inside its piece's
    product, and the assembled words with carries equal the schoolbook
    product, for both kinds of split (two pieces shorter than the other
    operand, many pieces at least as long) on random, all-ones and sparse
    limbs."""
    exe = tmp_path / "ragged_assembly_check"
    subprocess.run(["g++", "-O2", "-std=c++17", "-I", os.path.join(ROOT, "gpupi_amd", "csrc"),
                    os.path.join(ROOT, "tests", "native", "ragged_assembly_check.cpp"), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=600)
    assert out.returncode == 0, out.stdout + out.stderr
    assert "ragged assembly check: ok" in out.stdout
