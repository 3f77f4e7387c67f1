"""Known-answer inputs defined by the reference's own tests.

convert_cases(): the four convert_decimal correctness cases of
convert_decimal_test.cu:301-304 (477 * 2^k digits, k = 0..3, seeds
0x1234..0x4567), with the test's word generator (:58-63), fraction fill
(:65-80) and its rejection of inputs whose expansion has a run of more than
12 equal 0s or 9s (:109-125, :139-145).  The expected digits are the exact
expansion floor(X * 10^d / 2^bits) (:82-107)."""
from __future__ import annotations

import math

MASK64 = (1 << 64) - 1
LOG2_10 = 3.32192809488736234787


def bits_for_digits(d: int) -> int:
    return int(math.ceil(d * LOG2_10) + 34)


def _next_word(state: int):
    state ^= state >> 12
    state ^= (state << 25) & MASK64
    state ^= state >> 27
    return state, ((state * 2685821657736338717) & MASK64) >> 32


def fill_random_fraction(nwords: int, bits: int, seed: int):
    state = seed if seed else 1
    words = []
    for _ in range(nwords):
        state, w = _next_word(state)
        words.append(w)
    extra = bits & 31
    if extra:
        words[-1] &= (1 << extra) - 1
    if words and words[-1] == 0:
        words[-1] = 1
        if extra:
            words[-1] &= (1 << extra) - 1
    return words


def exact_digits(words, bits: int, digits: int) -> str:
    x = sum(w << (32 * i) for i, w in enumerate(words))
    return str((x * 10 ** digits) >> bits).zfill(digits)[-digits:] if digits else ""


def max_zero_nine_run(s: str) -> int:
    best = cur = 0
    prev = ""
    for ch in s:
        if ch in "09" and ch == prev:
            cur += 1
        elif ch in "09":
            cur = 1
        else:
            cur = 0
        prev = ch
        best = max(best, cur)
    return best


def convert_cases():
    out = []
    for total, seed in [(477, 0x1234), (954, 0x2345), (1908, 0x3456), (3816, 0x4567)]:
        bits = bits_for_digits(total) + 64
        nwords = (bits + 31) // 32
        for attempt in range(8):
            words = fill_random_fraction(nwords, bits, (seed + attempt * 0x9E3779B97F4A7C15) & MASK64)
            expected = exact_digits(words, bits, total)
            if max_zero_nine_run(expected) <= 12:
                break
        out.append((total, bits, words, expected))
    return out


# Published in the reference README (README.md:72-74), 10^9 digits, under the
# precision convention that sizes RET from the rounded-up digit count.
README_1E9 = {
    "target_digits": 1000000000,
    "convert_digits": 1000341504,
    "ret_len": 103847955,
    "ret_preview": "3243f6a88...d6e80bbbf2a7f738",
    "ret_hash": 354748112,
    "dec_head": "3141592653589793238462643383279502884197169399375105820974944592",
    "dec_tail": "1990548327874671398682093196353628204612755715171395115275045519",
}
