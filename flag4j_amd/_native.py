"""ctypes binding of libflag4j_amd.so (the C ABI declared in include/flag4j_amd.h).

The library is loaded from this package directory.  There is no fallback: if
the library is missing or fails to load, importing raises, and every call
reports the library's own status code and message.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libflag4j_amd.so")

OK = 0
ERR_INVALID_ARG = -1
ERR_SHAPE = -2
ERR_OUT_OF_MEMORY = -3
ERR_HIP = -4
ERR_NO_DEVICE = -5
ERR_UNSUPPORTED = -6
ERR_ACCURACY = -7
ERR_SINGULAR = -8

OP = {"N": 0, "T": 1, "C": 2}
PREC = {"auto": -1, "fp64": 0, "strict": 1, "split_bf16": 2, "split_fp16": 3, "split_int8": 4}

# Every symbol include/flag4j_amd.h declares, with its ctypes signature.
_i, _i64, _sz, _vp, _cp = ctypes.c_int, ctypes.c_int64, ctypes.c_size_t, ctypes.c_void_p, ctypes.c_char_p
_dp = ctypes.POINTER(ctypes.c_double)


class DeviceInfo(ctypes.Structure):
    _fields_ = [("device", ctypes.c_int), ("compute_units", ctypes.c_int), ("clock_khz", ctypes.c_int),
                ("lds_bytes_per_cu", ctypes.c_int), ("xcds", ctypes.c_int), ("hbm_bytes", ctypes.c_int64),
                ("peak_fp64_tflops", ctypes.c_double), ("peak_bf16_tflops", ctypes.c_double),
                ("peak_hbm_gbs", ctypes.c_double), ("name", ctypes.c_char * 64), ("arch", ctypes.c_char * 32)]


ABI_VERSION = 4          # F4J_ABI_VERSION the structures below match


class SplitReport(ctypes.Structure):
    _fields_ = [("precision", ctypes.c_int), ("slices", ctypes.c_int), ("products", ctypes.c_int),
                ("launches", ctypes.c_int), ("checked_rows", ctypes.c_int), ("max_rel_err", ctypes.c_double),
                ("product_ms", ctypes.c_double)]


SIGNATURES = {
    "f4j_abi_version": ([], _i),
    "f4j_init": ([_i], _i),
    "f4j_shutdown": ([], _i),
    "f4j_last_error": ([], _cp),
    "f4j_device_info_get": ([ctypes.POINTER(DeviceInfo)], _i),
    "f4j_dev_alloc": ([_sz, ctypes.POINTER(_vp)], _i),
    "f4j_dev_free": ([_vp], _i),
    "f4j_memcpy_h2d": ([_vp, _vp, _sz, _vp], _i),
    "f4j_memcpy_d2h": ([_vp, _vp, _sz, _vp], _i),
    "f4j_stream_sync": ([_vp], _i),
    "f4j_dgemm": ([_i, _i, _i64, _i64, _i64, _vp, _i64, _vp, _i64, _vp, _i64, _i, _i], _i),
    "f4j_dgemm_dev": ([_i, _i, _i64, _i64, _i64, _vp, _i64, _vp, _i64, _vp, _i64, _i, _i, _vp], _i),
    "f4j_dgemm_dev_noalloc": ([_i, _i, _i64, _i64, _i64, _vp, _i64, _vp, _i64, _vp, _i64, _i, _i, _vp], _i),
    "f4j_zgemm": ([_i, _i, _i64, _i64, _i64, _vp, _i64, _vp, _i64, _vp, _i64, _i, _i, _i], _i),
    "f4j_zgemm_dev": ([_i, _i, _i64, _i64, _i64, _vp, _i64, _vp, _i64, _vp, _i64, _i, _i, _i, _vp], _i),
    "f4j_dzgemm": ([_i64, _i64, _i64, _vp, _i64, _vp, _i64, _vp, _i64, _i, _i], _i),
    "f4j_dzgemm_dev": ([_i64, _i64, _i64, _vp, _i64, _vp, _i64, _vp, _i64, _i, _i, _vp], _i),
    "f4j_zdgemm": ([_i64, _i64, _i64, _vp, _i64, _vp, _i64, _vp, _i64, _i, _i], _i),
    "f4j_zdgemm_dev": ([_i64, _i64, _i64, _vp, _i64, _vp, _i64, _vp, _i64, _i, _i, _vp], _i),
    "f4j_dgemv": ([_i64, _i64, _vp, _i64, _vp, _i64, _vp, _i64, _i], _i),
    "f4j_dgemv_dev": ([_i64, _i64, _vp, _i64, _vp, _i64, _vp, _i64, _i, _vp], _i),
    "f4j_zgemv": ([_i64, _i64, _vp, _i64, _vp, _i64, _vp, _i64, _i], _i),
    "f4j_zgemv_dev": ([_i64, _i64, _vp, _i64, _vp, _i64, _vp, _i64, _i, _vp], _i),
    "f4j_dlstsq": ([_i64, _i64, _i64, _vp, _i64, _vp, _i64, _vp, _i64, _i], _i),
    "f4j_dlstsq_dev": ([_i64, _i64, _i64, _vp, _i64, _vp, _i64, _vp, _i64, _i, _vp], _i),
    "f4j_dpinv": ([_i64, _i64, _vp, _i64, _vp, _i64], _i),
    "f4j_dpinv_dev": ([_i64, _i64, _vp, _i64, _vp, _i64, _vp], _i),
    "f4j_grouped_plan_create": ([_i, _i, _i, _i64, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _i,
                                 ctypes.POINTER(_vp)], _i),
    "f4j_grouped_plan_create_split": ([_i, _i, _i, _i64, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _i, _i, _i,
                                       ctypes.POINTER(_vp)], _i),
    "f4j_grouped_plan_split_report": ([_vp, _vp, _i64, ctypes.POINTER(ctypes.c_int)], _i),
    "f4j_grouped_plan_run": ([_vp, _vp], _i),
    "f4j_grouped_plan_destroy": ([_vp], _i),
    "f4j_grouped_plan_info": ([_vp, ctypes.POINTER(ctypes.c_double), _vp, _vp, _i], _i),
    "f4j_grouped_plan_tune": ([_vp, _i], _i),
    "f4j_tuner_save": ([_cp], _i),
    "f4j_tuner_load": ([_cp], _i),
    "f4j_tuner_clear": ([], _i),
    "f4j_last_split_report": ([ctypes.POINTER(SplitReport)], _i),
    "f4j_config_set": ([_cp, _cp], _i),
    "f4j_config_get": ([_cp, _cp, _sz], _i),
    "f4j_config_load": ([_cp], _i),
    "f4j_dispatch_name": ([_i, _i, _i, _i64, _i64, _i64, _i, _cp, _sz], _i),
}

_lib = None


def load():
    """Load the native library (once).  Raises OSError if it is missing."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise OSError(f"flag4j_amd native library not built: {LIB_PATH} is missing "
                          "(run __graft_entry__.build() or make -C flag4j_amd/csrc)")
        lib = ctypes.CDLL(LIB_PATH)
        for name, (args, res) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.argtypes = args
            fn.restype = res
        if lib.f4j_abi_version() != ABI_VERSION:
            raise OSError(f"{LIB_PATH} has ABI version {lib.f4j_abi_version()}, these bindings expect "
                          f"{ABI_VERSION}: rebuild the library")
        _lib = lib
    return _lib


class NativeError(RuntimeError):
    """A nonzero f4j_status from the native library."""

    def __init__(self, status, message):
        super().__init__(f"flag4j_amd status {status}: {message}")
        self.status = status


def check(status):
    if status != OK:
        msg = load().f4j_last_error()
        raise NativeError(status, msg.decode() if msg else "")
    return status


def last_split_report():
    rep = SplitReport()
    check(load().f4j_last_split_report(ctypes.byref(rep)))
    return rep


def config_set(key, value):
    check(load().f4j_config_set(key.encode(), str(value).encode()))


def config_get(key):
    buf = ctypes.create_string_buffer(4096)
    check(load().f4j_config_get(key.encode(), buf, 4096))
    return buf.value.decode()


def device_info():
    info = DeviceInfo()
    check(load().f4j_device_info_get(ctypes.byref(info)))
    return info
