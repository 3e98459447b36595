"""ctypes binding of the C ABI declared in include/isspa_amd.h.

The shared library libisspa_amd.so is built in-tree (``__graft_entry__.build``)
next to this package.  There is no fallback: if the library is missing or
fails to load, importing this module raises, so no caller can silently run a
CPU substitute.
"""
from __future__ import annotations

import ctypes as C
import os

PKG_ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(PKG_ROOT, "libisspa_amd.so")


class IsspaError(RuntimeError):
    pass


if not os.path.exists(LIB_PATH):
    raise ImportError(f"{LIB_PATH} is not built: run __graft_entry__.build() (hipcc, gfx950)")

# One HIP runtime per process: PyTorch's ROCm wheel carries its own
# libamdhip64.so.7 and loads it by path.  Imported first, it is the copy the
# library's libamdhip64.so.7 dependency resolves to; loaded after the library,
# it would bring up a second runtime beside the system one, and the library's
# own calls then fail ("no ROCm-capable device").  Without PyTorch the system
# runtime is the only one.
try:
    import torch  # noqa: F401
except ImportError:
    pass

_lib = C.CDLL(LIB_PATH)

vp = C.c_void_p
i32 = C.c_int
u32 = C.c_uint32
u64 = C.c_uint64
f32 = C.c_float
ip = C.POINTER(C.c_int)
fp = C.POINTER(C.c_float)


class Windows(C.Structure):
    """isspa_windows"""
    _fields_ = [("n_windows", i32), ("window_atoms", vp), ("system_ids", vp), ("lbox", vp)]


class ModelDesc(C.Structure):
    """isspa_model_desc"""
    _fields_ = [("n_atoms", i32), ("n_types", i32), ("n_mc", i32), ("n_g", i32), ("n_e", i32),
                ("n_f", i32), ("g_r0", f32), ("g_dr", f32), ("e_r0", f32), ("e_dr", f32),
                ("f_r0", f32), ("f_dr", f32), ("lbox", f32), ("e0", f32), ("p0", f32),
                ("q0", f32), ("o0", f32), ("ep", f32), ("seed", u64), ("system_id", u32)]


class DiffDesc(C.Structure):
    """isspa_diff_desc"""
    _fields_ = [("n_atoms", i32), ("n_types", i32), ("n_mc", i32), ("lbox", f32), ("e0", f32),
                ("p0", f32), ("q0", f32), ("o0", f32), ("ep", f32), ("sig_s", C.c_double),
                ("eps_s", C.c_double), ("core", C.c_double), ("delta", C.c_double),
                ("seed", u64), ("step", u32)]


_SIGS = {
    "isspa_last_error": ([], C.c_char_p),
    "isspa_abi_version": ([], i32),
    "isspa_dev_malloc": ([C.POINTER(vp), C.c_size_t], i32),
    "isspa_dev_free": ([vp], i32),
    "isspa_memcpy_htod": ([vp, vp, C.c_size_t, vp], i32),
    "isspa_memcpy_dtoh": ([vp, vp, C.c_size_t, vp], i32),
    "isspa_memcpy_dtod": ([vp, vp, C.c_size_t, vp], i32),
    "isspa_zero": ([vp, C.c_size_t, vp], i32),
    "isspa_add4": ([vp, vp, i32, vp], i32),
    "isspa_stream_create": ([C.POINTER(vp)], i32),
    "isspa_stream_destroy": ([vp], i32),
    "isspa_stream_sync": ([vp], i32),
    "isspa_timer_begin": ([C.POINTER(vp), vp], i32),
    "isspa_timer_end": ([vp, vp, fp], i32),
    "isspa_events_create": ([C.POINTER(vp)], i32),
    "isspa_events_mark_begin": ([vp, vp], i32),
    "isspa_events_mark_end": ([vp, vp], i32),
    "isspa_events_elapsed": ([vp, fp], i32),
    "isspa_events_destroy": ([vp], i32),
    "isspa_default_constants": ([fp, fp, fp, fp, fp], i32),
    "isspa_model_create": ([C.POINTER(ModelDesc), vp, vp, vp, vp, vp, C.POINTER(vp)], i32),
    "isspa_model_create_batch": ([C.POINTER(ModelDesc), C.POINTER(Windows), vp, vp, vp, vp, vp,
                                  C.POINTER(vp)], i32),
    "isspa_model_destroy": ([vp], i32),
    "isspa_model_force": ([vp, vp, vp, u32, vp, vp], i32),
    "isspa_model_force_noalloc": ([vp, vp, vp, u32, vp, vp], i32),
    "isspa_model_force_keep": ([vp, vp, vp, vp, u32, vp, vp], i32),
    "isspa_model_force_components": ([vp, vp, vp, vp, vp, u32, vp], i32),
    "isspa_model_get_field": ([vp, vp, vp, vp, vp, vp, vp], i32),
    "isspa_model_profile": ([vp, vp, i32, fp, vp], i32),
    "isspa_model_set_tiling": ([vp, i32, i32], i32),
    "isspa_model_get_tiling": ([vp, ip, ip], i32),
    "isspa_model_set_target_waves": ([vp, i32], i32),
    "nonbond_model_create": ([i32, i32, vp, vp, vp, vp, vp, vp, i32, vp, i32, vp, f32, f32,
                              C.POINTER(vp)], i32),
    "nonbond_model_create_batch": ([i32, i32, vp, vp, vp, vp, vp, vp, i32, vp, i32, vp, f32, f32,
                                    C.POINTER(Windows), C.POINTER(vp)], i32),
    "nonbond_model_destroy": ([vp], i32),
    "nonbond_force": ([vp, vp, vp, vp], i32),
    "nonbond_force_components": ([vp, vp, vp, vp, vp, vp], i32),
    "bonded_model_create": ([i32, i32, vp, i32, vp, i32, vp, i32, vp, i32, vp, vp, i32, vp, vp,
                             i32, vp, vp, vp, vp, f32, C.POINTER(vp)], i32),
    "bonded_model_destroy": ([vp], i32),
    "bonded_force": ([vp, i32, vp, vp, vp, vp], i32),
    "leapfrog_andersen": ([i32, vp, vp, vp, f32, f32, f32, f32, u64, u32, u32, vp, vp], i32),
    "us_model_create": ([i32, i32, vp, vp, vp, f32, f32, C.POINTER(vp)], i32),
    "us_model_create_batch": ([i32, i32, vp, vp, vp, vp, vp, vp, C.POINTER(vp)], i32),
    "us_model_destroy": ([vp], i32),
    "us_force": ([vp, vp, vp, vp, vp], i32),
    "md_engine_create": ([vp, vp, vp, vp, i32, vp, vp, vp, f32, f32, f32, f32, u64, u32,
                          C.POINTER(vp)], i32),
    "md_engine_destroy": ([vp], i32),
    "md_engine_set_step": ([vp, u32, vp], i32),
    "md_engine_step": ([vp, vp], i32),
    "md_engine_capture": ([vp, i32, vp], i32),
    "md_engine_replay": ([vp, i32, vp], i32),
    "md_engine_forces": ([vp, vp], i32),
    "md_engine_get_com": ([vp, vp, vp], i32),
    "isspa_diff_create": ([C.POINTER(DiffDesc), C.POINTER(Windows), vp, vp, vp, vp, vp, i32,
                           C.POINTER(vp)], i32),
    "isspa_diff_destroy": ([vp], i32),
    "isspa_diff_set_tiling": ([vp, i32, i32], i32),
    "isspa_diff_forces": ([vp, vp, vp, vp, vp, vp], i32),
    "isspa_diff_vjp": ([vp, vp, vp, vp, vp, vp, vp, vp, vp], i32),
    "isspa_diff_vjp_after_forces": ([vp, vp, vp, vp, vp, vp, vp, vp, vp], i32),
}

EXPORTED = tuple(_SIGS)

for _name, (_args, _res) in _SIGS.items():
    _fn = getattr(_lib, _name)
    _fn.argtypes = _args
    _fn.restype = _res

lib = _lib

TERM_BOND, TERM_ANGLE, TERM_DIH, TERM_ALL = 1, 2, 4, 7


def check(rc: int, what: str = "") -> None:
    if rc != 0:
        msg = _lib.isspa_last_error().decode(errors="replace")
        raise IsspaError(f"{what or 'isspa call'} failed (status {rc}): {msg}")


def default_constants():
    vals = [C.c_float() for _ in range(5)]
    check(_lib.isspa_default_constants(*[C.byref(v) for v in vals]))
    return dict(zip(("e0", "p0", "q0", "o0", "ep"), (v.value for v in vals)))


def ptr(a) -> int:
    """Address of a host numpy array or a device torch tensor."""
    if a is None:
        return None
    if hasattr(a, "data_ptr"):
        return a.data_ptr()
    return a.ctypes.data
