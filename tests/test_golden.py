"""CPU: the oracle against the committed golden fixtures (tests/golden/).

- pi_digits.txt: pi from an independent implementation (mpmath); the
  oracle's digit string must be its prefix at several sizes, under both
  precision conventions;
- convert_cases.json: the reference's four convert_decimal known answers
  (convert_decimal_test.cu:301-304); the fixture must equal what the test
  generator makes and the oracle's conversion must reproduce it;
- readme_1e9.json: the numbers the reference README publishes for 10^9
  digits; the oracle's sizing must give the published RET length and
  rounded digit count, and bench.py must check the same values."""
import ctypes
import json
import os

import numpy as np
import pytest

import oracle_ffi as of
import refcases

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load_json(name):
    with open(os.path.join(GOLDEN, name)) as f:
        return json.load(f)


def pi_fixture():
    with open(os.path.join(GOLDEN, "pi_digits.txt")) as f:
        return f.read().strip()


def test_pi_fixture_shape():
    s = pi_fixture()
    assert len(s) == 20001 and s.startswith("31415926535897932384626433832795028841971")
    # a known digit string: the six nines at decimal places 762-767
    assert s[762:768] == "999999"


@pytest.mark.parametrize("digits", [1, 9, 100, 575, 576, 2000, 19999])
@pytest.mark.parametrize("readme_prec", [False, True])
def test_oracle_digits_match_fixture(digits, readme_prec):
    _, dec, _ = of.pi_digits(digits, readme_prec)
    assert "3" + dec == pi_fixture()[:digits + 1]


def test_convert_fixture_is_the_reference_generator():
    fx = load_json("convert_cases.json")
    gen = refcases.convert_cases()
    assert [(c["digits"], c["bits"], c["words"], c["expected"]) for c in fx] == [tuple(g) for g in gen]
    # the expansion is exact: floor(X 10^d / 2^bits)
    for c in fx:
        x = sum(w << (32 * i) for i, w in enumerate(c["words"]))
        assert c["expected"] == str((x * 10 ** c["digits"]) >> c["bits"]).zfill(c["digits"])


@pytest.mark.parametrize("case", range(4))
def test_oracle_convert_matches_fixture(case):
    c = load_json("convert_cases.json")[case]
    total, bits, words, expected = c["digits"], c["bits"], c["words"], c["expected"]
    leaf = 477
    levels = (total // leaf).bit_length() - 1
    powers = (of.OArray * max(levels, 1))()
    assert of.lib().o_convert_precompute_powers(leaf, levels, powers) == 0
    h = of.HostArray(np.array([words], dtype=np.uint32))
    out = ctypes.create_string_buffer(total)
    assert of.lib().o_convert_equal_split(h.oarray(), bits, total, leaf, powers, levels, out) == 0
    for i in range(levels):
        of.lib().o_free(powers[i].data)
    assert out.raw[:total].decode() == expected


def test_readme_fixture_sizing_and_bench_check():
    fx = load_json("readme_1e9.json")
    assert fx == refcases.README_1E9
    L = of.lib()
    leaf = ctypes.c_uint32(0)
    assert L.o_round_up_decimal_digits(fx["target_digits"], ctypes.byref(leaf)) == fx["convert_digits"]
    assert L.o_prec_limbs_for_digits(fx["convert_digits"]) + 1 == fx["ret_len"]
    # pi's leading hex digits (RET's top limbs) and decimal head agree with
    # the independent fixture
    assert fx["ret_preview"].startswith("3243f6a88")
    assert fx["dec_head"] == pi_fixture()[:64]
    import importlib.util
    spec = importlib.util.spec_from_file_location(
        "bench", os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "bench.py"))
    bench = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(bench)
    for k in ("ret_len", "ret_hash", "ret_preview", "dec_head", "dec_tail"):
        assert bench.README_1E9[k] == fx[k]
