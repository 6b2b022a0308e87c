"""ctypes binding of libophip.so (the C-ABI declared in include/ophip.h).

Mirrors how OPLib's C# layer binds opcuda through [DllImport] (e.g.
oplib/opmodel/models/CModel Device.cs:52-105): one declaration per entry
point, 32-bit unsigned device handles, 64-bit host pointers.
``DeviceArray`` plays the role of the reference's ``CArray`` device buffers
(oplib/opmodel/types/CArray.cs:42-1072): alloc, copy in, copy out, free.

Loading fails loudly when the library is missing -- there is no fallback.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libophip.so")

U32, I32, F32, F64 = C.c_uint, C.c_int, C.c_float, C.c_double
ULONG, LONG, LL = C.c_ulong, C.c_long, C.c_longlong
V = None

# name -> (restype, [argtypes]); the order follows include/ophip.h
SIGNATURES = {
    "opcuda_cublas_init": (I32, []),
    "opcuda_shutdown": (I32, []),
    "opcuda_thread_synchronize": (V, []),
    "opcuda_set_device": (V, [I32]),
    "opcuda_get_device": (I32, []),
    "opcuda_device_get_count": (I32, []),
    "opcuda_device_count": (I32, []),
    "opcuda_multi_processor_count": (I32, [I32]),
    "opcuda_max_threads_per_block": (I32, [I32]),
    "opcuda_get_status": (U32, []),
    "opcuda_mem_alloc": (U32, [U32]),
    "opcuda_mem_free_device": (V, [U32]),
    "opcuda_mem_alloc_host": (ULONG, [U32]),
    "opcuda_mem_free_host": (V, [ULONG]),
    "opcuda_memcpy_h2d": (I32, [U32, ULONG, U32]),
    "opcuda_memcpy_d2h": (I32, [U32, ULONG, U32]),
    "opcuda_memcpy_d2d": (I32, [U32, U32, U32]),
    "opcuda_create_timer": (V, []),
    "opcuda_reset_timer": (V, []),
    "opcuda_stop_timer": (V, []),
    "opcuda_get_timer_value": (F64, []),
    "opcuda_scopy": (V, [I32, U32, I32, U32, I32]),
    "opcuda_scopy1": (I32, [U32, U32, U32]),
    "opcuda_ssetall": (V, [U32, I32, F32, I32]),
    "opcuda_ssetone": (V, [U32, I32, F32, I32]),
    "opcuda_sgemm": (V, [I32, I32, I32, F32, U32, I32, U32, I32, F32, U32, I32]),
    "opcuda_sgemv": (V, [I32, I32, F32, U32, I32, U32, I32, F32, U32, I32]),
    "opcuda_sgemvT": (V, [I32, I32, F32, U32, I32, U32, I32, F32, U32, I32]),
    "opcuda_ssqmm": (V, [I32, I32, U32, U32, U32]),
    "opcuda_sgemm4": (V, [I32, U32]),
    "opcuda_sgemv3": (V, [I32, I32, I32, U32, U32, U32, U32]),
    "opcuda_sgemv4": (V, [I32, U32]),
    "opcuda_sek": (V, [I32, U32, F32, U32]),
    "opcuda_sekd": (V, [I32, U32, F32, U32, U32]),
    "opcuda_sekd_lambda": (V, [I32, U32, F32, U32, U32, U32]),
    "opcuda_sck": (V, [I32, I32, U32, U32]),
    "opcuda_sglvg": (V, [U32, I32, I32, U32, U32]),
    "opcuda_sgsvg": (V, [U32, I32, I32, I32, U32, U32]),
    "opcuda_mc_load_mt_gpu": (I32, [C.c_char_p, LONG]),
    "opcuda_mc_nrng": (I32, []),
    "opcuda_mc_status_sz": (I32, []),
    "opcuda_mc_setseed": (V, [ULONG, U32]),
    "opcuda_mc1f": (I32, [U32, I32, U32, I32, I32, I32, U32, U32, ULONG]),
    "opcuda_mt_benchmark": (I32, [U32, U32, I32]),
    "ophip_set_status": (U32, [U32]),
    "ophip_device_address": (ULONG, [U32]),
    "ophip_arena_bytes": (ULONG, []),
    "ophip_stream": (ULONG, []),
    "ophip_dgen_rs": (I32, [I32, I32, U32, U32, U32, F64, U32]),
    "ophip_sek_rs": (I32, [I32, I32, U32, U32, U32, F64, F64, I32, U32]),
    "ophip_dgen_sv": (I32, [I32, I32] + [U32] * 8 + [F64, I32, U32]),
    "ophip_sek_rs_batch": (I32, [I32, C.POINTER(I32), C.POINTER(I32), C.POINTER(U32),
                                 C.POINTER(U32), C.POINTER(U32), C.POINTER(F64), C.POINTER(F64),
                                 I32, C.POINTER(U32)]),
    "ophip_sck_rows": (I32, [I32, I32, U32, U32]),
    "ophip_sgemm_strided_batched": (I32, [I32, I32, I32, F32, U32, I32, LL, U32, I32, LL, F32,
                                          U32, I32, LL, I32]),
    "ophip_sgemm_skinny": (I32, [I32, I32, I32, F32, U32, I32, U32, I32, U32, I32]),
    "ophip_sck_rows_ex": (I32, [I32, I32, U32, U32, I32]),
    "ophip_sgemm_skinny_ex": (I32, [I32, I32, I32, F32, U32, I32, U32, I32, U32, I32, I32]),
    "ophip_saxpby": (I32, [I32, F32, U32, F32, U32, U32]),
    "ophip_wsq_bytes": (ULONG, [I32, I32]),
    "ophip_wsq_w": (U32, [U32, I32, I32, I32]),
    "ophip_wsq_array": (U32, [U32, I32, I32, I32, I32]),
    "ophip_wsq_plan": (ULONG, [I32, C.POINTER(U32), C.POINTER(I32), C.POINTER(I32), I32, I32,
                               I32]),
    "ophip_wsq_exec": (I32, [ULONG, I32]),
    "ophip_wsq_assemble_rs": (I32, [I32, C.POINTER(I32), C.POINTER(I32), C.POINTER(U32),
                                    C.POINTER(U32), C.POINTER(U32), C.POINTER(F64),
                                    C.POINTER(F64), I32, I32, C.POINTER(U32)]),
    "ophip_wsq_exec_ex": (I32, [ULONG, I32, I32]),
    "ophip_wsq_set_skip": (I32, [I32]),
    "ophip_wsq_set_kcut": (I32, [I32]),
    "ophip_wsq_plan_free": (V, [ULONG]),
    "ophip_w8_bytes": (ULONG, [I32]),
    "ophip_w8_w": (U32, [U32, I32, I32]),
    "ophip_w8_array": (U32, [U32, I32, I32]),
    "ophip_w8_exec": (I32, [U32, I32, I32]),
    "ophip_wsq_bwd_layout": (I32, [U32, I32, I32, I32]),
    "ophip_wsq_bwd_layout_perm": (I32, [U32, I32, I32, I32, U32]),
    "ophip_wsq_bwd_layout_regime": (I32, [U32, I32, I32, I32, I32, I32]),
    "ophip_wsq_bwd_layout_regime_batch": (I32, [I32, C.POINTER(U32), I32, I32, I32, I32, I32]),
    "ophip_wsq_bwd_prep": (I32, [I32, I32, I32, U32, U32]),
    "ophip_wsq_bwd_step": (I32, [U32, I32, I32, I32, I32, F32, U32, U32, U32, U32]),
    "ophip_wsq_bwd_step_batch": (I32, [I32, C.POINTER(U32), I32, I32, I32, I32, C.POINTER(F32),
                                       C.POINTER(U32), C.POINTER(U32), C.POINTER(U32),
                                       C.POINTER(U32)]),
    "ophip_wsq_set_bwd_kernel": (I32, [I32]),
    "ophip_wsq_set_bwd_cutoffs": (I32, [I32, I32]),
    "ophip_wsq_get_bwd_cutoffs": (I32, [C.POINTER(I32), C.POINTER(I32)]),
    "ophip_wsq_bwd_kernel_for": (I32, [I32, I32, I32]),
    "ophip_wsq_plan_tiles": (ULONG, [I32, C.POINTER(U32), C.POINTER(I32), C.POINTER(I32), I32,
                                     C.POINTER(I32), I32]),
    "ophip_wsq_autotune": (I32, [I32, C.POINTER(U32), C.POINTER(I32), C.POINTER(I32), I32,
                                 C.POINTER(I32), C.POINTER(I32)]),
    "ophip_fwd_plan": (ULONG, [I32, C.POINTER(I32), C.POINTER(U32), C.POINTER(I32),
                               C.POINTER(U32), C.POINTER(U32), C.POINTER(U32), C.POINTER(F32), F32,
                               I32, C.POINTER(F32), I32]),
    "ophip_fwd_exec": (I32, [ULONG]),
    "ophip_fwd_set_cells": (I32, [ULONG, C.POINTER(U32)]),
    "ophip_fwd_plan_free": (V, [ULONG]),
    "ophip_spot_prices": (I32, [I32, U32, I32, C.POINTER(I32), I32, U32, I32, F32, U32]),
    "ophip_fwd_set_rates": (I32, [ULONG, C.POINTER(F32)]),
    "ophip_mlp_create": (ULONG, [I32, I32]),
    "ophip_mlp_free": (V, [ULONG]),
    "ophip_mlp_param_count": (I32, []),
    "ophip_mlp_params": (I32, [ULONG, I32, I32, C.POINTER(F32), I32]),
    "ophip_mlp_forward": (I32, [ULONG, I32, U32, U32]),
    "ophip_mlp_backward": (I32, [ULONG, I32, U32, U32, U32, U32]),
    "ophip_mlp_adam": (I32, [ULONG, F32, F32, F32, F32]),
    "ophip_mlp_loss": (I32, [ULONG, I32, C.POINTER(F64)]),
    "ophip_surr_eval": (I32, [ULONG, I32, U32, C.POINTER(F32), C.POINTER(F32), F32, F32, I32,
                              C.POINTER(F32), I32, F32, U32, U32, U32]),
    "ophip_surr_base": (I32, [I32, U32, F32, F32, I32, C.POINTER(F32), I32, U32]),
    "ophip_mc_set_nrng": (I32, [I32]),
    "ophip_mc1f_rows": (I32, [U32, I32, U32, I32, I32, I32, U32, U32, ULONG]),
    "ophip_mc_guide": (I32, [I32, I32, U32, U32]),
    "ophip_dcmt_load": (I32, [C.c_void_p, C.c_long]),
    "ophip_dcmt_status_sz": (I32, []),
    "ophip_dcmt_setseed": (I32, [C.POINTER(U32), U32]),
    "ophip_dcmt_uniforms": (I32, [U32, U32, I32]),
    "ophip_mc_payoff": (I32, [U32, I32, I32, I32, C.POINTER(I32), C.POINTER(F64), U32, I32,
                              C.POINTER(F32), I32, U32]),
    "ophip_mc1f_guided": (I32, [U32, I32, U32, I32, I32, I32, U32, U32, U32]),
    "ophip_mc_payoff_noalloc": (I32, [U32, I32, I32, I32, C.POINTER(I32), C.POINTER(F64), U32, I32,
                                      C.POINTER(F32), I32, U32]),
    "ophip_read_probe": (I32, [U32, ULONG, U32]),
    "ophip_reserve_workspace": (I32, [ULONG]),
    "ophip_capture_begin": (I32, []),
    "ophip_capture_end": (ULONG, []),
    "ophip_graph_launch": (I32, [ULONG]),
    "ophip_graph_destroy": (V, [ULONG]),
    "ophip_event_record": (I32, [I32]),
    "ophip_event_elapsed_ms": (F64, [I32, I32]),
}


def _load(path: str = LIB_PATH) -> C.CDLL:
    if not os.path.exists(path):
        raise RuntimeError(
            f"libophip.so not found at {path}: build it first "
            "(python -c 'import __graft_entry__ as g; g.build()')")
    lib = C.CDLL(path)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)  # AttributeError = missing export: fail loudly
        fn.restype = res
        fn.argtypes = args
    return lib


lib = _load()


class OphipError(RuntimeError):
    pass


def check(status=None, what: str = ""):
    """Raise when the thread-local status (or the given one) is non-zero."""
    st = lib.opcuda_get_status() if status is None else status
    if st:
        raise OphipError(f"{what}: ophip status 0x{int(st):x}")
    return st


def init():
    check(lib.opcuda_cublas_init(), "opcuda_cublas_init")


def sync():
    lib.opcuda_thread_synchronize()
    check(None, "opcuda_thread_synchronize")


class DeviceArray:
    """A device buffer addressed by a 32-bit arena handle (cf. CArray)."""

    def __init__(self, nbytes: int, dtype=np.float32, shape=None):
        self.nbytes = int(nbytes)
        self.dtype = np.dtype(dtype)
        self.shape = shape
        self.h = lib.opcuda_mem_alloc(self.nbytes)
        if not self.h:
            raise OphipError(f"opcuda_mem_alloc({nbytes}) failed: 0x{lib.opcuda_get_status():x}")

    @classmethod
    def from_numpy(cls, a: np.ndarray, order="F"):
        a = np.asarray(a)
        flat = np.ascontiguousarray(a.reshape(-1, order=order))
        da = cls(flat.nbytes, flat.dtype, a.shape)
        da.upload(flat)
        return da

    @classmethod
    def empty(cls, shape, dtype=np.float32):
        shape = tuple(np.atleast_1d(shape))
        n = int(np.prod(shape)) * np.dtype(dtype).itemsize
        return cls(n, dtype, shape)

    def upload(self, a: np.ndarray, offset: int = 0):
        a = np.ascontiguousarray(a)
        check(lib.opcuda_memcpy_h2d(self.h + offset, a.ctypes.data, a.nbytes), "h2d")

    def download(self, order="F", dtype=None, count=None) -> np.ndarray:
        dt = np.dtype(dtype) if dtype is not None else self.dtype
        n = count if count is not None else self.nbytes // dt.itemsize
        out = np.empty(n, dtype=dt)
        check(lib.opcuda_memcpy_d2h(self.h, out.ctypes.data, out.nbytes), "d2h")
        if self.shape is not None and count is None and dtype is None:
            return out.reshape(self.shape, order=order)
        return out

    def free(self):
        if self.h:
            lib.opcuda_mem_free_device(self.h)
            self.h = 0

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass
