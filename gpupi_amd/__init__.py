"""gpupi_amd: Python host bindings for libgpupi.so, the MI355X (gfx950) build
of the gpupi batched big-integer and pi library.

The library's interface is the C ABI declared in include/*.h (batch_arith.h,
batch_kernels.h, convert_decimal.h, gpupi_pi.h); this module mirrors it with
ctypes so tests and benchmarks can drive it from Python.  Device memory is
plain hipMalloc memory from the HIP runtime; there is no fallback path: if
libgpupi.so is missing or no GPU is present the calls raise.
"""
from __future__ import annotations

import ctypes
import os
from typing import Optional

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libgpupi.so")
CLI_PATH = os.path.join(_HERE, "pi_bs_gpu")

HIP_SUCCESS = 0
HIP_ERROR_INVALID_VALUE = 1
PREC_REQUESTED = 0
PREC_ROUNDED = 1
FORM_BATCHED = 0
FORM_PLAIN = 1

_H2D, _D2H, _D2D = 1, 2, 3


class BatchMPArray(ctypes.Structure):
    """reference batch_arith.h:10-17 (data is a device pointer)."""

    _fields_ = [
        ("data", ctypes.c_void_p),
        ("length", ctypes.c_uint32),
        ("batch_size", ctypes.c_uint32),
        ("stride", ctypes.c_uint32),
    ]

    def view(self, offset_words: int = 0, length: Optional[int] = None,
             batch_size: Optional[int] = None, stride: Optional[int] = None) -> "BatchMPArray":
        base = (self.data or 0) + 4 * offset_words
        return BatchMPArray(base,
                            self.length if length is None else length,
                            self.batch_size if batch_size is None else batch_size,
                            self.stride if stride is None else stride)


class GpupiPiResult(ctypes.Structure):
    """include/gpupi_pi.h GpupiPiResult."""

    _fields_ = [
        ("requested_digits", ctypes.c_uint32),
        ("convert_digits", ctypes.c_uint32),
        ("leaf_digits", ctypes.c_uint32),
        ("prec_limbs", ctypes.c_uint32),
        ("groups", ctypes.c_uint32),
        ("ret_length", ctypes.c_uint32),
        ("ret", ctypes.c_void_p),
        ("digits", ctypes.c_void_p),
        ("elapsed_ms", ctypes.c_float),
        ("inverse_bits", ctypes.c_uint64),
    ]


class GpupiKernelStat(ctypes.Structure):
    _fields_ = [
        ("name", ctypes.c_char * 48),
        ("launches", ctypes.c_uint64),
        ("total_ms", ctypes.c_double),
        ("total_bytes", ctypes.c_double),
    ]


class GpupiMulItem(ctypes.Structure):
    _fields_ = [
        ("A", ctypes.c_void_p),
        ("B", ctypes.c_void_p),
        ("C", ctypes.c_void_p),
        ("L_a", ctypes.c_uint32),
        ("L_b", ctypes.c_uint32),
    ]


class GpupiServeConfig(ctypes.Structure):
    _fields_ = [("hbm_budget_bytes", ctypes.c_uint64), ("pi_workers", ctypes.c_uint32),
                ("mul_batch_max", ctypes.c_uint32), ("pi_batch_max", ctypes.c_uint32)]


class GpupiJobResult(ctypes.Structure):
    _fields_ = [
        ("id", ctypes.c_int64), ("kind", ctypes.c_int32), ("status", ctypes.c_int32),
        ("submit_ms", ctypes.c_double), ("start_ms", ctypes.c_double), ("finish_ms", ctypes.c_double),
        ("reserved_bytes", ctypes.c_uint64), ("batch_size", ctypes.c_uint32), ("ret_length", ctypes.c_uint32),
        ("ret_hash", ctypes.c_uint32), ("dec_hash", ctypes.c_uint32),
    ]


class GpupiStageEntry(ctypes.Structure):
    _fields_ = [
        ("kind", ctypes.c_char * 16),
        ("label", ctypes.c_char * 24),
        ("n1", ctypes.c_uint32),
        ("batch_size", ctypes.c_uint32),
        ("lhs_length", ctypes.c_uint32),
        ("rhs_length", ctypes.c_uint32),
        ("elapsed_ms", ctypes.c_float),
        ("uses_ntt", ctypes.c_int),
    ]


class NTTPrecomputedTables(ctypes.Structure):
    """include/batch_kernels.h: the NTT's device root tables."""
    _fields_ = [
        ("roots", ctypes.c_void_p),
        ("max_log2", ctypes.c_uint32),
    ]


class ConvertDecimalPowerTable(ctypes.Structure):
    _fields_ = [
        ("powers", ctypes.POINTER(BatchMPArray)),
        ("levels", ctypes.c_uint32),
        ("leaf_digits", ctypes.c_uint32),
    ]


_lib = None
_hip = None


def _declare(lib):
    A = BatchMPArray
    c_ctx = ctypes.c_void_p
    u32, i32, u64 = ctypes.c_uint32, ctypes.c_int32, ctypes.c_uint64
    err = ctypes.c_int
    vp = ctypes.c_void_p
    sig = {
        "batch_mp_init": (c_ctx, []),
        "batch_mp_destroy": (None, [c_ctx]),
        "batch_mp_array_create": (A, [u32, u32, u32]),
        "batch_mp_array_release": (None, [A]),
        "batch_mp_workspace_size": (ctypes.c_size_t, [c_ctx]),
        "batch_mp_ensure_workspace": (err, [c_ctx, ctypes.c_size_t]),
        "batch_mp_mul": (err, [c_ctx, A, A, A]),
        "batch_mp_mul_shared": (err, [c_ctx, A, A, A, A, A]),
        "batch_mp_mul_noalloc": (err, [c_ctx, A, A, A]),
        "batch_mp_mul_small": (err, [c_ctx, A, u32, A]),
        "batch_mp_exactdiv_small": (err, [c_ctx, A, u32]),
        "batch_mp_add": (err, [c_ctx, A, A, A]),
        "batch_mp_add_small": (err, [c_ctx, A, u32, A]),
        "batch_mp_shift_add": (err, [c_ctx, A, A, u32, A]),
        "batch_mp_shift_sub": (err, [c_ctx, A, A, u32, A]),
        "batch_mp_sub": (err, [c_ctx, A, A, A]),
        "batch_mp_sub_small": (err, [c_ctx, A, u32, A]),
        "batch_mp_shift_bits": (err, [c_ctx, A, A, i32]),
        "batch_mp_bitlength_max": (err, [c_ctx, A, ctypes.POINTER(u32)]),
        "batch_mp_limblength_max": (err, [c_ctx, A, ctypes.POINTER(u32)]),
        "batch_mp_array_compact": (err, [c_ctx, ctypes.POINTER(A)]),
        "batch_mp_set_stream": (err, [c_ctx, ctypes.c_void_p]),
        "batch_mp_set_ntt_threshold": (err, [c_ctx, u32]),
        "gpupi_pi_run": (err, [c_ctx, u32, ctypes.c_int, ctypes.POINTER(GpupiPiResult)]),
        "gpupi_pi_run_batch": (err, [c_ctx, ctypes.POINTER(u32), ctypes.c_int, ctypes.c_int,
                                     ctypes.POINTER(GpupiPiResult)]),
        "gpupi_pi_groups_for_digits": (u32, [u32]),
        "gpupi_pi_free": (None, [ctypes.POINTER(GpupiPiResult)]),
        "gpupi_binary_split": (err, [c_ctx, u32, ctypes.POINTER(A), ctypes.POINTER(A)]),
        "gpupi_sqrt_10005": (err, [c_ctx, ctypes.c_int, ctypes.POINTER(A), ctypes.POINTER(A)]),
        "gpupi_recursive_inverse": (err, [c_ctx, A, u64, ctypes.POINTER(A)]),
        "gpupi_pi_fractional": (err, [c_ctx, ctypes.c_int, ctypes.POINTER(A), ctypes.POINTER(u32),
                                      ctypes.POINTER(u64)]),
        "gpupi_construct_leaf_17": (err, [c_ctx, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p,
                                          ctypes.c_void_p, u32]),
        "gpupi_round_up_decimal_digits": (u32, [u32, ctypes.POINTER(u32)]),
        "gpupi_prec_limbs_for_digits": (u32, [u32]),
        "gpupi_calib_mac_peak": (err, [ctypes.POINTER(ctypes.c_double)]),
        "gpupi_calib_copy": (err, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p]),
        "gpupi_device_info": (err, [ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int), ctypes.c_char_p,
                                    ctypes.c_int]),
        "convert_decimal_precompute_powers": (err, [c_ctx, ctypes.POINTER(ConvertDecimalPowerTable), u32, u32]),
        "convert_decimal_release_powers": (None, [ctypes.POINTER(ConvertDecimalPowerTable)]),
        "convert_decimal_equal_split": (err, [c_ctx, A, u32, u32, u32,
                                              ctypes.POINTER(ConvertDecimalPowerTable), ctypes.c_void_p]),
        "batch_decimal_naive": (None, [ctypes.c_void_p, ctypes.c_void_p, u32, ctypes.c_int, u32,
                                       ctypes.c_int, u32]),
        "batch_mul_ntt_workspace_size": (ctypes.c_size_t, [u32, u32, u32]),
        # the per-op L1 layer (include/batch_kernels.h); pointers are device addresses
        "batch_add_simple": (None, [vp, vp, vp, vp] + [u32] * 7),
        "batch_sub_simple": (None, [vp, vp, vp, vp] + [u32] * 7),
        "batch_add_simple_workspace_size": (ctypes.c_size_t, [u32] * 4),
        "batch_sub_simple_workspace_size": (ctypes.c_size_t, [u32] * 4),
        "batch_add_small": (None, [vp, u32, vp, vp] + [u32] * 5),
        "batch_sub_small": (None, [vp, u32, vp, vp] + [u32] * 5),
        "batch_mul_small": (None, [vp, u32, vp, vp] + [u32] * 5),
        "batch_add_small_workspace_size": (ctypes.c_size_t, [u32] * 3),
        "batch_sub_small_workspace_size": (ctypes.c_size_t, [u32] * 3),
        "batch_mul_small_workspace_size": (ctypes.c_size_t, [u32] * 3),
        "batch_shift_add_simple": (None, [vp, vp, vp, vp] + [u32] * 8),
        "batch_shift_sub_simple": (None, [vp, vp, vp, vp] + [u32] * 8),
        "batch_shift_add_simple_workspace_size": (ctypes.c_size_t, [u32] * 5),
        "batch_shift_sub_simple_workspace_size": (ctypes.c_size_t, [u32] * 5),
        "batch_shift_bits": (None, [vp, vp] + [u32] * 5 + [i32]),
        "batch_exactdiv_small": (None, [vp, u32, u32, u32, u32]),
        "batch_bitlength_max": (u32, [vp, vp, u32, u32, u32]),
        "batch_limblength_max": (u32, [vp, vp, u32, u32, u32]),
        "batch_bitlength_workspace_size": (ctypes.c_size_t, [u32, u32]),
        "batch_mul_naive": (None, [vp, vp, vp] + [ctypes.c_int] * 6),
        "init_ntt_precomputed_tables": (None, [ctypes.POINTER(NTTPrecomputedTables)]),
        "release_ntt_precomputed_tables": (None, [ctypes.POINTER(NTTPrecomputedTables)]),
        "batch_mul_ntt": (None, [vp, vp, vp, vp, NTTPrecomputedTables] + [u32] * 6),
        "gpupi_kernel_stats_reset": (None, [ctypes.c_int]),
        "gpupi_kernel_stats_query": (ctypes.c_int, [ctypes.POINTER(GpupiKernelStat), ctypes.c_int]),
        "gpupi_kernel_stats_query_sizes": (ctypes.c_int, [ctypes.POINTER(GpupiKernelStat), ctypes.c_int]),
        "gpupi_ntt_set_split": (err, [ctypes.c_int, ctypes.c_int, ctypes.c_int]),
        "gpupi_ntt_get_split": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(ctypes.c_int)]),
        "gpupi_ntt_set_launch": (err, [ctypes.c_uint32, ctypes.c_int]),
        "gpupi_ntt_get_launch": (None, [ctypes.POINTER(ctypes.c_uint32), ctypes.POINTER(ctypes.c_int)]),
        "gpupi_group_stage_set_shape": (err, [ctypes.c_uint32]),
        "gpupi_group_stage_get_shape": (ctypes.c_uint32, []),
        "gpupi_stage_profile_enable": (err, [c_ctx, ctypes.c_int]),
        "gpupi_set_pipeline_form": (err, [c_ctx, ctypes.c_int]),
        "gpupi_cache_query": (err, [c_ctx, ctypes.POINTER(ctypes.c_size_t), ctypes.POINTER(u64)]),
        "gpupi_cache_trim": (err, [c_ctx]),
        "gpupi_mul_ragged": (err, [c_ctx, ctypes.POINTER(GpupiMulItem), u32]),
        "gpupi_server_create": (ctypes.c_void_p, [ctypes.POINTER(GpupiServeConfig)]),
        "gpupi_server_submit_pi": (ctypes.c_int64, [ctypes.c_void_p, u32, ctypes.c_void_p]),
        "gpupi_server_submit_mul": (ctypes.c_int64, [ctypes.c_void_p, ctypes.c_void_p, u32, ctypes.c_void_p, u32,
                                                     ctypes.c_void_p]),
        "gpupi_server_results": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(GpupiJobResult), ctypes.c_int,
                                                ctypes.c_int]),
        "gpupi_server_drain": (None, [ctypes.c_void_p]),
        "gpupi_server_memory": (err, [ctypes.c_void_p, ctypes.POINTER(u64), ctypes.POINTER(u64)]),
        "gpupi_set_memory_pool": (err, [c_ctx, ctypes.c_void_p]),
        "gpupi_server_destroy": (None, [ctypes.c_void_p]),
        "gpupi_serve_job_bytes": (u64, [ctypes.c_int, u32, u32]),
        "gpupi_serve_admit": (ctypes.c_int, [ctypes.POINTER(u64), ctypes.POINTER(ctypes.c_int32), ctypes.c_int, u64,
                                             u64, ctypes.c_int, ctypes.POINTER(ctypes.c_int)]),
        "gpupi_leaf_stage_create": (err, [c_ctx, u32, ctypes.c_int, ctypes.POINTER(ctypes.c_void_p)]),
        "gpupi_leaf_stage_run": (err, [ctypes.c_void_p, ctypes.POINTER(A), ctypes.POINTER(A), ctypes.POINTER(A)]),
        "gpupi_leaf_stage_destroy": (None, [ctypes.c_void_p]),
        "gpupi_stage_profile_query": (ctypes.c_int, [c_ctx, ctypes.POINTER(GpupiStageEntry), ctypes.c_int]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args


def lib():
    """The loaded libgpupi.so (raises if it was not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with __graft_entry__.build()")
        _lib = ctypes.CDLL(LIB_PATH)
        _declare(_lib)
    return _lib


def hip():
    global _hip
    if _hip is None:
        _hip = ctypes.CDLL("libamdhip64.so", mode=ctypes.RTLD_GLOBAL)
        _hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        _hip.hipFree.argtypes = [ctypes.c_void_p]
        _hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        _hip.hipMemset.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_size_t]
        _hip.hipGetErrorString.restype = ctypes.c_char_p
        _hip.hipGetDeviceCount.argtypes = [ctypes.POINTER(ctypes.c_int)]
        _hip.hipMemGetInfo.argtypes = [ctypes.POINTER(ctypes.c_size_t), ctypes.POINTER(ctypes.c_size_t)]
    return _hip


def check(err: int, what: str = "HIP call") -> None:
    if err != HIP_SUCCESS:
        msg = hip().hipGetErrorString(err).decode()
        raise RuntimeError(f"{what} failed: {err} ({msg})")


def gpu_available() -> bool:
    try:
        n = ctypes.c_int(0)
        return hip().hipGetDeviceCount(ctypes.byref(n)) == 0 and n.value > 0
    except OSError:
        return False


def synchronize() -> None:
    check(hip().hipDeviceSynchronize(), "hipDeviceSynchronize")


class DeviceBuffer:
    """A hipMalloc'd buffer of uint32 words (or bytes)."""

    def __init__(self, nbytes: int):
        self.nbytes = max(int(nbytes), 4)
        p = ctypes.c_void_p()
        check(hip().hipMalloc(ctypes.byref(p), self.nbytes), "hipMalloc")
        self.ptr = p.value

    @classmethod
    def from_array(cls, arr: np.ndarray) -> "DeviceBuffer":
        arr = np.ascontiguousarray(arr)
        buf = cls(arr.nbytes)
        if arr.nbytes:
            check(hip().hipMemcpy(buf.ptr, arr.ctypes.data, arr.nbytes, _H2D), "hipMemcpy H2D")
        return buf

    def fill(self, byte: int) -> None:
        check(hip().hipMemset(self.ptr, byte, self.nbytes), "hipMemset")

    def download(self, dtype=np.uint32, count: Optional[int] = None, offset_bytes: int = 0) -> np.ndarray:
        itemsize = np.dtype(dtype).itemsize
        if count is None:
            count = (self.nbytes - offset_bytes) // itemsize
        out = np.empty(count, dtype=dtype)
        if count:
            check(hip().hipMemcpy(out.ctypes.data, self.ptr + offset_bytes, count * itemsize, _D2H),
                  "hipMemcpy D2H")
        return out

    def free(self) -> None:
        if self.ptr:
            hip().hipFree(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def download_words(ptr: int, count: int) -> np.ndarray:
    out = np.empty(count, dtype=np.uint32)
    if count:
        check(hip().hipMemcpy(out.ctypes.data, ptr, count * 4, _D2H), "hipMemcpy D2H")
    return out


def download_bytes(ptr: int, count: int) -> bytes:
    out = ctypes.create_string_buffer(max(count, 1))
    if count:
        check(hip().hipMemcpy(out, ptr, count, _D2H), "hipMemcpy D2H")
    return out.raw[:count]


class Context:
    """Owns a BatchMPContext (reference batch_arith.h:33-43)."""

    def __init__(self):
        self._lib = lib()
        self.ptr = self._lib.batch_mp_init()
        if not self.ptr:
            raise RuntimeError("batch_mp_init failed (no GPU, or the HIP runtime could not start)")

    def close(self) -> None:
        if self.ptr:
            self._lib.batch_mp_destroy(self.ptr)
            self.ptr = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __getattr__(self, name):
        # ctx.mul(A, B, C) -> batch_mp_mul(ctx, A, B, C), etc.
        fn = getattr(lib(), "batch_mp_" + name, None)
        if fn is None:
            raise AttributeError(name)
        return lambda *args: fn(self.ptr, *args)

    def limblength_max_value(self, A: BatchMPArray) -> int:
        r = ctypes.c_uint32(0)
        check(self._lib.batch_mp_limblength_max(self.ptr, A, ctypes.byref(r)), "limblength_max")
        return r.value

    def pi(self, digits: int, convention: int = PREC_REQUESTED):
        """Run the pipeline; returns (RET limbs, digit string, result struct)."""
        res = GpupiPiResult()
        check(self._lib.gpupi_pi_run(self.ptr, digits, convention, ctypes.byref(res)), "gpupi_pi_run")
        try:
            ret = download_words(res.ret, res.ret_length)
            dec = download_bytes(res.digits, digits).decode() if digits else ""
        finally:
            self._lib.gpupi_pi_free(ctypes.byref(res))
        return ret, dec, res

    def pi_batch(self, digits, convention: int = PREC_REQUESTED):
        """Several pi jobs sharing one binary-splitting tree (gpupi_pi_run_batch);
        returns [(RET limbs, digit string, result struct)] in input order."""
        n = len(digits)
        arr = (ctypes.c_uint32 * n)(*digits)
        res = (GpupiPiResult * n)()
        check(self._lib.gpupi_pi_run_batch(self.ptr, arr, n, convention, res), "gpupi_pi_run_batch")
        out = []
        try:
            for i, d in enumerate(digits):
                ret = download_words(res[i].ret, res[i].ret_length)
                dec = download_bytes(res[i].digits, d).decode() if d else ""
                out.append((ret, dec, res[i]))
        finally:
            for i in range(n):
                self._lib.gpupi_pi_free(ctypes.byref(res[i]))
        return out

    def set_form(self, form: int) -> None:
        """FORM_BATCHED (default) or FORM_PLAIN, the reference's per-op form."""
        check(self._lib.gpupi_set_pipeline_form(self.ptr, form), "gpupi_set_pipeline_form")

    def cache_stats(self) -> tuple:
        """(bytes of device blocks cached now, device allocations made so far)."""
        b, n = ctypes.c_size_t(0), ctypes.c_uint64(0)
        check(self._lib.gpupi_cache_query(self.ptr, ctypes.byref(b), ctypes.byref(n)), "gpupi_cache_query")
        return b.value, n.value

    def cache_trim(self) -> None:
        """Free the device blocks the context keeps for the pipeline."""
        check(self._lib.gpupi_cache_trim(self.ptr), "gpupi_cache_trim")

    def mul_ragged(self, items) -> int:
        """items: sequence of (A_ptr, B_ptr, C_ptr, L_a, L_b); returns the hipError_t."""
        arr = (GpupiMulItem * max(len(items), 1))(*[GpupiMulItem(*it) for it in items])
        return self._lib.gpupi_mul_ragged(self.ptr, arr, len(items))

    def stage_profile(self, enable: bool) -> None:
        """Turn per-stage timing of the pipeline on (clearing it) or off."""
        check(self._lib.gpupi_stage_profile_enable(self.ptr, 1 if enable else 0), "gpupi_stage_profile_enable")

    def stage_entries(self) -> list:
        """[dict(kind, label, n1, batch, lhs, rhs, ms, ntt)] recorded since stage_profile(True)."""
        n = self._lib.gpupi_stage_profile_query(self.ptr, None, 0)
        if n < 0:
            raise RuntimeError("gpupi_stage_profile_query failed")
        arr = (GpupiStageEntry * max(n, 1))()
        self._lib.gpupi_stage_profile_query(self.ptr, arr, n)
        return [dict(kind=e.kind.decode(), label=e.label.decode(), n1=e.n1, batch=e.batch_size, lhs=e.lhs_length,
                     rhs=e.rhs_length, ms=e.elapsed_ms, ntt=bool(e.uses_ntt)) for e in arr[:n]]


LEAF_PLAIN, LEAF_BATCHED, LEAF_GRAPH, LEAF_FUSED = 0, 1, 2, 3
JOB_PI, JOB_MUL = 0, 1


def serve_admit(bytes_, kinds, budget: int, in_use: int, free_pi_workers: int) -> list:
    """The server's admission rule on a queue (host only)."""
    n = len(bytes_)
    b = (ctypes.c_uint64 * max(n, 1))(*bytes_)
    k = (ctypes.c_int32 * max(n, 1))(*kinds)
    out = (ctypes.c_int * max(n, 1))()
    m = lib().gpupi_serve_admit(b, k, n, budget, in_use, free_pi_workers, out)
    return list(out[:m])


class Server:
    """The serving loop (gpupi_serve.h).  Host buffers passed to submit_*
    are kept alive here until their results are collected."""

    def __init__(self, budget_bytes: int = 0, pi_workers: int = 0, mul_batch_max: int = 0, pi_batch_max: int = 0):
        self._lib = lib()
        cfg = GpupiServeConfig(budget_bytes, pi_workers, mul_batch_max, pi_batch_max)
        self.ptr = self._lib.gpupi_server_create(ctypes.byref(cfg))
        if not self.ptr:
            raise RuntimeError("gpupi_server_create failed")
        self.keep = {}

    def submit_pi(self, digits: int, keep_digits: bool = False) -> int:
        out = ctypes.create_string_buffer(max(digits, 1)) if keep_digits else None
        jid = self._lib.gpupi_server_submit_pi(self.ptr, digits, out)
        if jid >= 0:
            self.keep[jid] = out
        return jid

    def submit_mul(self, A: np.ndarray, B: np.ndarray) -> int:
        A = np.ascontiguousarray(A, dtype=np.uint32)
        B = np.ascontiguousarray(B, dtype=np.uint32)
        C = np.zeros(A.size + B.size, dtype=np.uint32)
        jid = self._lib.gpupi_server_submit_mul(self.ptr, A.ctypes.data if A.size else None, A.size,
                                                B.ctypes.data if B.size else None, B.size, C.ctypes.data)
        if jid >= 0:
            self.keep[jid] = (A, B, C)
        return jid

    def results(self, wait: bool = True, max_results: int = 256) -> list:
        """[(GpupiJobResult, output)] of finished jobs; output is the digit
        bytes (pi, if kept) or the product limbs (multiply)."""
        arr = (GpupiJobResult * max_results)()
        n = self._lib.gpupi_server_results(self.ptr, arr, max_results, 1 if wait else 0)
        out = []
        for r in arr[:n]:
            k = self.keep.pop(r.id, None)
            if r.kind == JOB_MUL:
                out.append((r, k[2] if k is not None else None))
            else:
                out.append((r, k.raw if k is not None else None))
        return out

    def memory(self) -> tuple:
        """(peak bytes handed out, peak bytes held) by the server's memory pool."""
        used, reserved = ctypes.c_uint64(0), ctypes.c_uint64(0)
        check(self._lib.gpupi_server_memory(self.ptr, ctypes.byref(used), ctypes.byref(reserved)),
              "gpupi_server_memory")
        return used.value, reserved.value

    def drain(self) -> None:
        self._lib.gpupi_server_drain(self.ptr)

    def close(self) -> None:
        if self.ptr:
            self._lib.gpupi_server_destroy(self.ptr)
            self.ptr = None


class LeafStage:
    """The 17-term group stage for `groups` groups in one of three forms
    (LEAF_PLAIN, LEAF_BATCHED, LEAF_GRAPH, LEAF_FUSED); run() returns (P, Q, R) views."""

    def __init__(self, ctx: "Context", groups: int, mode: int):
        self._lib = lib()
        self.ptr = ctypes.c_void_p()
        check(self._lib.gpupi_leaf_stage_create(ctx.ptr, groups, mode, ctypes.byref(self.ptr)),
              "gpupi_leaf_stage_create")

    def run(self):
        P, Q, R = BatchMPArray(), BatchMPArray(), BatchMPArray()
        check(self._lib.gpupi_leaf_stage_run(self.ptr, ctypes.byref(P), ctypes.byref(Q), ctypes.byref(R)),
              "gpupi_leaf_stage_run")
        return P, Q, R

    def close(self) -> None:
        if self.ptr:
            self._lib.gpupi_leaf_stage_destroy(self.ptr)
            self.ptr = ctypes.c_void_p()


def ret_mod_hash(limbs: np.ndarray) -> int:
    """RET hash of the reference CLI (pi_bs_gpu.cu:852-858): the limbs read as
    one base-2^32 number, modulo 1e9+7.  Evaluated as a pairwise tree (each
    round folds limb pairs a + b*B and squares B), vectorised in uint64:
    every operand stays below 2^30, so no product overflows."""
    mod = 1000000007
    v = np.asarray(limbs, dtype=np.uint64) % mod
    base = (1 << 32) % mod
    if v.size == 0:
        return 0
    while v.size > 1:
        if v.size & 1:
            v = np.append(v, np.uint64(0))
        v = (v[0::2] + (v[1::2] * np.uint64(base)) % mod) % mod
        base = base * base % mod
    return int(v[0])


def rows_of(arr: BatchMPArray) -> np.ndarray:
    """Download the (batch, length) limbs of any device array view."""
    stride = arr.stride if arr.stride else arr.length
    total = (arr.batch_size - 1) * stride + arr.length if arr.batch_size else 0
    full = np.zeros(arr.batch_size * max(stride, 1), dtype=np.uint32)
    if total:
        full[:total] = download_words(arr.data, total)
    return full.reshape(arr.batch_size, max(stride, 1))[:, :arr.length].copy()


def upload_rows(rows: np.ndarray, stride: Optional[int] = None, pad_value: int = 0):
    """rows: (N, L) uint32.  Returns (buffer, BatchMPArray) with the given stride."""
    rows = np.asarray(rows, dtype=np.uint32)
    n, length = rows.shape
    stride = length if stride is None else stride
    host = np.full((n, max(stride, 1)), pad_value, dtype=np.uint32)
    host[:, :length] = rows
    buf = DeviceBuffer.from_array(host)
    return buf, BatchMPArray(buf.ptr, length, n, stride)


def download_rows(arr: BatchMPArray, buf: DeviceBuffer) -> np.ndarray:
    """The (batch, length) limbs of an array view living in buf."""
    stride = arr.stride if arr.stride else arr.length
    total = (arr.batch_size - 1) * stride + arr.length if arr.batch_size else 0
    off = arr.data - buf.ptr
    flat = buf.download(np.uint32, count=total, offset_bytes=off)
    out = np.zeros((arr.batch_size, arr.length), dtype=np.uint32)
    for i in range(arr.batch_size):
        out[i] = flat[i * stride: i * stride + arr.length]
    return out


def device_info() -> dict:
    """Compute units, peak engine clock (kHz) and architecture of the current device."""
    cus, clk = ctypes.c_int(0), ctypes.c_int(0)
    arch = ctypes.create_string_buffer(64)
    check(lib().gpupi_device_info(ctypes.byref(cus), ctypes.byref(clk), arch, 64), "gpupi_device_info")
    return {"compute_units": cus.value, "clock_khz": clk.value, "arch": arch.value.decode()}


def mac_peak() -> float:
    """Measured device peak of v_mad_u64_u32 multiply-adds per second."""
    r = ctypes.c_double(0.0)
    check(lib().gpupi_calib_mac_peak(ctypes.byref(r)), "gpupi_calib_mac_peak")
    return r.value


def kernel_stats_reset(enable: bool) -> None:
    lib().gpupi_kernel_stats_reset(1 if enable else 0)


def kernel_stats(by_size: bool = False) -> list:
    """[(name, launches, total_ms, total_bytes)] for the NTT kernels timed
    since the last kernel_stats_reset(True); with by_size one entry per
    (kernel, transform size), named "<kernel>@2^<log2 K>"."""
    cap = 256
    arr = (GpupiKernelStat * cap)()
    fn = lib().gpupi_kernel_stats_query_sizes if by_size else lib().gpupi_kernel_stats_query
    n = fn(arr, cap)
    if n < 0:
        raise RuntimeError("gpupi_kernel_stats_query failed")
    return [(arr[i].name.decode(), arr[i].launches, arr[i].total_ms, arr[i].total_bytes) for i in range(n)]


def ntt_set_split(log2_k: int, log2_m0: int, log2_m1: int = 0) -> None:
    """Column level sizes for transforms of 2^log2_k points (0 = built-in);
    raises ValueError for a split the kernels do not cover."""
    e = lib().gpupi_ntt_set_split(log2_k, log2_m0, log2_m1)
    if e != 0:
        raise ValueError(f"invalid NTT split 2^{log2_k} = 2^{log2_m0} x 2^{log2_m1} x rest (hip error {e})")


def ntt_set_launch(pair_mask: int, xcd_chunk_log2: int) -> None:
    """NTT launch shape: column level kinds with paired tiles (bit 0 first
    forward, 1 second forward, 2 second inverse, 3 first inverse level) and
    the log2 of the tiles each XCD takes in turn (0..8); raises ValueError
    outside those ranges."""
    e = lib().gpupi_ntt_set_launch(pair_mask, xcd_chunk_log2)
    if e != 0:
        raise ValueError(f"invalid NTT launch shape pair_mask={pair_mask} xcd_chunk_log2={xcd_chunk_log2}")


def ntt_get_launch() -> tuple:
    """(pair_mask, xcd_chunk_log2) in use."""
    m, c = ctypes.c_uint32(), ctypes.c_int()
    lib().gpupi_ntt_get_launch(ctypes.byref(m), ctypes.byref(c))
    return m.value, c.value


def group_stage_set_shape(groups_per_block: int) -> None:
    """Groups per block of the fused group-stage kernel (1, 2, 4, 8, 16)."""
    e = lib().gpupi_group_stage_set_shape(groups_per_block)
    if e != 0:
        raise ValueError(f"invalid group-stage shape {groups_per_block}")


def group_stage_get_shape() -> int:
    return int(lib().gpupi_group_stage_get_shape())


def ntt_get_split(log2_k: int) -> tuple:
    """log2 sizes of the levels used for transforms of 2^log2_k points."""
    m = (ctypes.c_int * 3)()
    n = lib().gpupi_ntt_get_split(log2_k, m)
    return tuple(m[i] for i in range(n))
