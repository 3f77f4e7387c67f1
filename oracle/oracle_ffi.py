"""ctypes bindings of oracle/liboracle.so — TEST INFRASTRUCTURE ONLY.

The oracle is a CPU restatement (GMP) of the reference pipeline used to check
the HIP library; only tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg may load it.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
CLI_PATH = os.path.join(HERE, "pi_oracle")


class OArray(ctypes.Structure):
    _fields_ = [
        ("data", ctypes.c_void_p),
        ("length", ctypes.c_uint32),
        ("batch_size", ctypes.c_uint32),
        ("stride", ctypes.c_uint32),
    ]


class OPiInfo(ctypes.Structure):
    _fields_ = [
        ("target_digits", ctypes.c_uint32),
        ("groups", ctypes.c_uint32),
        ("prec_limbs", ctypes.c_uint32),
        ("inverse_bits", ctypes.c_uint64),
    ]


_lib = None


def build() -> None:
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        A = OArray
        u32 = ctypes.c_uint32
        for name, args in {
            "o_mul": [A, A, A], "o_add": [A, A, A], "o_sub": [A, A, A],
            "o_mul_small": [A, u32, A], "o_add_small": [A, u32, A], "o_sub_small": [A, u32, A],
            "o_exactdiv_small": [A, u32], "o_shift_add": [A, A, u32, A], "o_shift_sub": [A, A, u32, A],
            "o_shift_bits": [A, A, ctypes.c_int32],
        }.items():
            getattr(L, name).argtypes = args
            getattr(L, name).restype = ctypes.c_int
        L.o_limblength_max.argtypes = [A]
        L.o_limblength_max.restype = u32
        L.o_bitlength_max.argtypes = [A]
        L.o_bitlength_max.restype = u32
        L.o_decimal_naive.argtypes = [ctypes.c_void_p, ctypes.c_void_p, u32, ctypes.c_int, u32, ctypes.c_int, u32]
        L.o_leaf17.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, u32]
        L.o_binary_split.argtypes = [u32, ctypes.POINTER(A), ctypes.POINTER(A)]
        L.o_leaf_stage.argtypes = [u32, ctypes.POINTER(A), ctypes.POINTER(A), ctypes.POINTER(A)]
        L.o_sqrt_10005.argtypes = [ctypes.c_int, ctypes.POINTER(A), ctypes.POINTER(A)]
        L.o_recursive_inverse.argtypes = [A, ctypes.c_uint64, ctypes.POINTER(A)]
        L.o_pi_fractional.argtypes = [ctypes.c_int, ctypes.POINTER(A), ctypes.POINTER(OPiInfo)]
        L.o_round_up_decimal_digits.argtypes = [u32, ctypes.POINTER(u32)]
        L.o_round_up_decimal_digits.restype = u32
        L.o_prec_limbs_for_digits.argtypes = [u32]
        L.o_prec_limbs_for_digits.restype = u32
        L.o_bits_for_digits.argtypes = [u32]
        L.o_bits_for_digits.restype = u32
        L.o_convert_precompute_powers.argtypes = [u32, u32, ctypes.POINTER(A)]
        L.o_convert_equal_split.argtypes = [A, u32, u32, u32, ctypes.POINTER(A), u32, ctypes.c_void_p]
        L.o_pi_digits.argtypes = [u32, ctypes.c_int, ctypes.POINTER(ctypes.POINTER(u32)), ctypes.POINTER(u32),
                                  ctypes.POINTER(ctypes.c_char_p), ctypes.POINTER(u32)]
        L.o_mod_hash.argtypes = [ctypes.c_void_p, u32]
        L.o_mod_hash.restype = u32
        L.o_decimal_hash.argtypes = [ctypes.c_char_p, u32]
        L.o_decimal_hash.restype = u32
        L.o_free.argtypes = [ctypes.c_void_p]
        L.o_set_threads.argtypes = [ctypes.c_int]
        L.o_set_threads.restype = None
        _lib = L
    return _lib


class HostArray:
    """Rows held in a numpy buffer with an explicit stride (0 = broadcast)."""

    def __init__(self, rows: np.ndarray, stride: int | None = None, pad: int = 0, extra: int = 0):
        rows = np.asarray(rows, dtype=np.uint32)
        if rows.ndim == 1:
            rows = rows[None, :]
        n, length = rows.shape
        self.length = length
        self.batch = n
        self.stride = length if stride is None else stride
        span = max(self.stride, length)
        self.buf = np.full(n * max(span, 1) + extra + 1, pad, dtype=np.uint32)
        for i in range(n):
            s = i * self.stride
            self.buf[s:s + length] = rows[i]

    @classmethod
    def empty(cls, n: int, length: int, stride: int | None = None, pad: int = 0):
        return cls(np.zeros((n, length), dtype=np.uint32), stride, pad)

    def oarray(self, length: int | None = None, batch: int | None = None, stride: int | None = None) -> OArray:
        o = OArray(self.buf.ctypes.data, self.length if length is None else length,
                   self.batch if batch is None else batch, self.stride if stride is None else stride)
        o._owner = self  # the struct points into self.buf: keep it alive as long as the struct
        return o

    def rows(self, length: int | None = None) -> np.ndarray:
        length = self.length if length is None else length
        out = np.zeros((self.batch, length), dtype=np.uint32)
        for i in range(self.batch):
            s = i * self.stride
            out[i] = self.buf[s:s + length]
        return out


def take_array(a: OArray) -> np.ndarray:
    """Copy a malloc'd batch-1 oracle array and free it."""
    n = a.length
    out = np.ctypeslib.as_array(ctypes.cast(a.data, ctypes.POINTER(ctypes.c_uint32)), shape=(max(n, 1),))[:n].copy()
    lib().o_free(a.data)
    return out


def take_rows(a: OArray) -> np.ndarray:
    """Copy a malloc'd oracle array as (batch_size, length) rows and free it."""
    n, length, stride = a.batch_size, a.length, a.stride if a.stride else a.length
    total = (n - 1) * stride + length if n else 0
    flat = np.ctypeslib.as_array(ctypes.cast(a.data, ctypes.POINTER(ctypes.c_uint32)), shape=(max(total, 1),))
    out = np.zeros((n, length), dtype=np.uint32)
    for i in range(n):
        out[i] = flat[i * stride: i * stride + length]
    lib().o_free(a.data)
    return out


def pi_digits(digits: int, readme_prec: bool = False):
    """(RET limbs, digit string after '3', convert_digits)."""
    L = lib()
    ret = ctypes.POINTER(ctypes.c_uint32)()
    ret_len = ctypes.c_uint32(0)
    dig = ctypes.c_char_p()
    cd = ctypes.c_uint32(0)
    st = L.o_pi_digits(digits, 1 if readme_prec else 0, ctypes.byref(ret), ctypes.byref(ret_len),
                       ctypes.byref(dig), ctypes.byref(cd))
    if st != 0:
        raise RuntimeError(f"oracle pi failed: {st}")
    limbs = np.ctypeslib.as_array(ret, shape=(max(ret_len.value, 1),))[:ret_len.value].copy()
    raw = ctypes.string_at(dig, digits) if digits else b""
    L.o_free(ctypes.cast(ret, ctypes.c_void_p))
    L.o_free(ctypes.cast(dig, ctypes.c_void_p))
    return limbs, raw.decode(), cd.value


def set_threads(threads: int) -> None:
    """Threads for the rows of the oracle's batched products (results do not
    depend on it)."""
    lib().o_set_threads(threads)


def mul(A: np.ndarray, B: np.ndarray) -> np.ndarray:
    """One product A * B of two limb vectors (GMP mpn_mul; releases the GIL
    while it runs, so several may run on threads)."""
    hc = HostArray.empty(1, A.size + B.size)
    assert lib().o_mul(HostArray(A.reshape(1, -1)).oarray(), HostArray(B.reshape(1, -1)).oarray(), hc.oarray()) == 0
    return hc.buf[:A.size + B.size].copy()


def mod_hash(limbs: np.ndarray) -> int:
    limbs = np.ascontiguousarray(limbs, dtype=np.uint32)
    return lib().o_mod_hash(limbs.ctypes.data, len(limbs))
