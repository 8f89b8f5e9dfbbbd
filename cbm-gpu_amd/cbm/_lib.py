"""Loader for the package's two native libraries (ctypes over their C ABI).

* ``lib/libcbm_host.so`` -- host structure builder (include/cbm_host.h)
* ``lib/libcbm_hip.so``  -- HIP kernels for gfx950 (include/cbm_hip.h)

Both are built in-tree by ``make -C cbm-gpu_amd/csrc``.  There is no
fallback: a missing library raises :class:`NativeLibraryMissing`.
"""
import ctypes
import os
import threading

import torch  # noqa: F401  (torch's HIP runtime is loaded first, libcbm_hip.so binds to it)

LIB_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "lib")
# CBM_HOST_LIB selects another build of the host library, e.g. the
# AddressSanitizer/UBSan one of `make -C cbm-gpu_amd/csrc asan` (lib/asan/).
HOST_LIB = os.environ.get("CBM_HOST_LIB") or os.path.join(LIB_DIR, "libcbm_host.so")
HIP_LIB = os.path.join(LIB_DIR, "libcbm_hip.so")

_lock = threading.Lock()
_host = None
_hip = None

c_i32, c_i64, c_u64, c_f64, c_vp = ctypes.c_int32, ctypes.c_int64, ctypes.c_uint64, ctypes.c_double, ctypes.c_void_p
c_f32 = ctypes.c_float
P_vp = ctypes.POINTER(ctypes.c_void_p)


class NativeLibraryMissing(RuntimeError):
    """A native library of the package has not been built."""


class CbmError(RuntimeError):
    """A native call returned a failure status."""


# Status codes of include/cbm_host.h and include/cbm_hip.h.
CBM_OK, CBM_ERR_ARG, CBM_ERR_VALUE, CBM_ERR_NOMEM, CBM_ERR_INTERNAL, CBM_ERR_DEVICE = 0, -1, -2, -3, -4, -5


def _sig(lib, name, restype, *argtypes):
    fn = getattr(lib, name)
    fn.restype = restype
    fn.argtypes = list(argtypes)


def host():
    """The host builder library, loaded once."""
    global _host
    if _host is not None:
        return _host
    with _lock:
        if _host is None:
            if not os.path.exists(HOST_LIB):
                raise NativeLibraryMissing(
                    f"{HOST_LIB} is not built; run `make -C cbm-gpu_amd/csrc` (or __graft_entry__.build())")
            lib = ctypes.CDLL(HOST_LIB)
            _sig(lib, "cbm_host_last_error", ctypes.c_char_p)
            _sig(lib, "cbm_graph_copying_model", c_i32, c_i64, c_i32, c_f64, c_f64, c_i32, c_u64, P_vp)
            _sig(lib, "cbm_graph_nnz", c_i64, c_vp)
            _sig(lib, "cbm_graph_row", c_vp, c_vp)
            _sig(lib, "cbm_graph_col", c_vp, c_vp)
            _sig(lib, "cbm_graph_free", None, c_vp)
            _sig(lib, "cbm_build_create", c_i32, c_vp, c_vp, c_vp, c_i64, c_i32, c_i32, c_i32, P_vp)
            _sig(lib, "cbm_build_size", c_i64, c_vp, c_i32)
            _sig(lib, "cbm_build_array", c_vp, c_vp, c_i32)
            _sig(lib, "cbm_build_stat", c_i64, c_vp, c_i32)
            _sig(lib, "cbm_build_free", None, c_vp)
            _sig(lib, "cbm_adjacency_create", c_i32, c_vp, c_vp, c_vp, c_i64, c_i32, P_vp)
            _sig(lib, "cbm_adjacency_rows", c_i32, c_vp)
            _sig(lib, "cbm_adjacency_nnz", c_i64, c_vp)
            _sig(lib, "cbm_adjacency_ptr", c_vp, c_vp)
            _sig(lib, "cbm_adjacency_col", c_vp, c_vp)
            _sig(lib, "cbm_adjacency_symmetric", c_i32, c_vp)
            _sig(lib, "cbm_adjacency_t_ptr", c_vp, c_vp)
            _sig(lib, "cbm_adjacency_t_col", c_vp, c_vp)
            _sig(lib, "cbm_adjacency_free", None, c_vp)
            _sig(lib, "cbm_candidates_create", c_i32, c_vp, c_i32, c_i32, P_vp)
            _sig(lib, "cbm_candidates_size", c_i64, c_vp)
            _sig(lib, "cbm_candidates_array", c_vp, c_vp, c_i32)
            _sig(lib, "cbm_candidates_free", None, c_vp)
            _sig(lib, "cbm_build_from_candidates", c_i32, c_vp, c_vp, c_vp, c_vp, c_i64, c_i32, P_vp)
            _sig(lib, "cbm_plan_create", c_i32, c_i32, c_vp, c_vp, c_vp, c_i64, c_vp, c_vp, c_i64, c_i32, P_vp)
            _sig(lib, "cbm_plan_size", c_i64, c_vp, c_i32)
            _sig(lib, "cbm_plan_array", c_vp, c_vp, c_i32)
            _sig(lib, "cbm_plan_depth", c_i32, c_vp)
            _sig(lib, "cbm_plan_stat", c_i64, c_vp, c_i32)
            _sig(lib, "cbm_plan_free", None, c_vp)
            _sig(lib, "cbm_gather_order", c_i32, c_i32, c_vp, c_vp, c_vp, c_i32, c_vp, ctypes.POINTER(c_i32))
            _host = lib
    return _host


def hip():
    """The HIP kernel library, loaded once.  Raises if it is not built."""
    global _hip
    if _hip is not None:
        return _hip
    with _lock:
        if _hip is None:
            if not os.path.exists(HIP_LIB):
                raise NativeLibraryMissing(
                    f"{HIP_LIB} is not built; run `make -C cbm-gpu_amd/csrc` (or __graft_entry__.build())")
            lib = ctypes.CDLL(HIP_LIB)
            from . import _hip_sigs
            _hip_sigs.declare(lib, _sig)
            _hip = lib
    return _hip


def check_host(rc, what=""):
    if rc == CBM_OK:
        return
    msg = host().cbm_host_last_error().decode(errors="replace")
    _raise(rc, msg, what)


def check_hip(rc, what=""):
    if rc == CBM_OK:
        return
    msg = hip().cbm_last_error().decode(errors="replace")
    _raise(rc, msg, what)


def check_candidates(rc, what=""):
    if rc == CBM_OK:
        return
    msg = hip().cbm_candidates_last_error().decode(errors="replace")
    _raise(rc, msg, what)


def _raise(rc, msg, what):
    text = f"{what}: {msg}" if what else msg
    if rc in (CBM_ERR_ARG, CBM_ERR_VALUE):
        raise ValueError(text)
    if rc == CBM_ERR_NOMEM:
        raise MemoryError(text)
    raise CbmError(f"{text} (status {rc})")
