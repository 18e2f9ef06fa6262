"""ctypes bindings of libmlem.so (include/mlem.h) and libmlem_synth.so.

The libraries are built in-tree by mlem_amd.build; loading fails loudly when
they are missing — there is no Python or CPU fallback for any GPU entry point.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

LIBDIR = Path(__file__).resolve().parent / "lib"
LIBMLEM = LIBDIR / "libmlem.so"
LIBSYNTH = LIBDIR / "libmlem_synth.so"

MLEM_OK = 0
MLEM_EINVAL = -1
MLEM_EIO = -2
MLEM_EFORMAT = -3
MLEM_ENOMEM = -4
MLEM_EDEVICE = -5
MLEM_ESTATE = -6

MODE_FAST = 0
MODE_EXACT = 1
MODE_ATOMIC = 2
MODES = {"fast": MODE_FAST, "exact": MODE_EXACT, "atomic": MODE_ATOMIC}

OPEN_NO_CSC = 0x1
OPEN_KEEP_ORDER = 0x2
OPEN_FP16 = 0x4
OPEN_BF16 = 0x8
OPEN_EXPAND_SYMMETRY = 0x10
INIT_REFERENCE = 0
INIT_GIVEN = 1


class MlemError(RuntimeError):
    """A failing C-ABI call: `code` is the MLEM_E* status, str() its message."""

    def __init__(self, code: int, message: str):
        super().__init__(message)
        self.code = code


class Options(C.Structure):
    _fields_ = [
        ("mode", C.c_int),
        ("init", C.c_int),
        ("checkpoint_interval", C.c_int),
        ("checkpoint_dir", C.c_char_p),
        ("verbose", C.c_int),
        ("use_graph", C.c_int),
        ("loglik", C.POINTER(C.c_double)),
        ("fwd_total", C.POINTER(C.c_double)),
    ]


class BatchOptions(C.Structure):
    _fields_ = [
        ("rel_tol", C.c_double),
        ("block_width", C.c_int),
        ("use_graph", C.c_int),
        ("mode", C.c_int),
    ]


_P = C.c_void_p
_u32, _u64, _i32 = C.c_uint32, C.c_uint64, C.c_int32
_PU32 = C.POINTER(C.c_uint32)

# name -> (restype, argtypes); the list mirrors include/mlem.h.
SIGNATURES = {
    "mlem_last_error": (C.c_char_p, []),
    "mlem_version": (C.c_char_p, []),
    "mlem_probe": (C.c_int, [C.c_char_p, _PU32, _PU32, C.POINTER(C.c_uint64), _PU32, _PU32]),
    "mlem_expand_symmetry_file": (C.c_int, [C.c_char_p, C.c_char_p]),
    "mlem_open": (C.c_int, [C.c_char_p, C.c_uint, C.POINTER(_P)]),
    "mlem_open_host": (C.c_int, [_u32, _u32, _u64, _P, _P, _P, C.c_uint, C.POINTER(_P)]),
    "mlem_open_streamed": (C.c_int, [C.c_char_p, C.c_uint, _u64, _u64, C.POINTER(_P)]),
    "mlem_open_host_streamed": (C.c_int, [_u32, _u32, _u64, _P, _P, _P, C.c_uint, _u64, _u64,
                                          C.POINTER(_P)]),
    "mlem_stream_blocks": (C.c_int, [_P, _PU32, _PU32]),
    "mlem_stream_passes": (C.c_int, [_P, C.POINTER(C.c_int)]),
    "mlem_close": (None, [_P]),
    "mlem_dims": (C.c_int, [_P, _PU32, _PU32, C.POINTER(C.c_uint64)]),
    "mlem_geometry": (C.c_int, [_P, _PU32, _PU32]),
    "mlem_stream": (_P, [_P]),
    "mlem_column_sums": (C.c_int, [_P, _P, _u32]),
    "mlem_init_image": (C.c_int, [_P, _P, _u32, _P, _u32, _P, C.POINTER(C.c_float)]),
    "mlem_fwproj": (C.c_int, [_P, _P, _u32, _P, _u32, C.c_int]),
    "mlem_correl": (C.c_int, [_P, _P, _P, _P, _u32]),
    "mlem_bkproj": (C.c_int, [_P, _P, _u32, _P, _u32, C.c_int]),
    "mlem_update": (C.c_int, [_P, _P, _P, _P, _u32]),
    "mlem_options_default": (None, [C.POINTER(Options)]),
    "mlem_reconstruct": (C.c_int, [_P, _P, _u32, _P, _u32, C.c_int, C.POINTER(Options)]),
    "mlem_reconstruct_noalloc": (C.c_int, [_P, _P, _u32, _P, _u32, C.c_int, C.c_int, _P, _P]),
    "mlem_recon_create": (C.c_int, [_P, _P, _u32, _P, _u32, C.POINTER(Options), C.POINTER(_P)]),
    "mlem_recon_destroy": (None, [_P]),
    "mlem_recon_run": (C.c_int, [_P, C.c_int]),
    "mlem_recon_run_timed": (C.c_int, [_P, C.c_int, _P, _P, C.POINTER(C.c_float)]),
    "mlem_recon_image": (C.c_int, [_P, _P, _u32]),
    "mlem_recon_iterations": (C.c_int, [_P]),
    "mlem_get_launch": (C.c_int, [_P, C.POINTER(C.c_int), C.POINTER(C.c_int),
                                  C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    "mlem_set_launch": (C.c_int, [_P, C.c_int, C.c_int, C.c_int, C.c_int]),
    "mlem_autotune": (C.c_int, [_P]),
    "mlem_batch_options_default": (None, [C.c_void_p]),
    "mlem_batch_create": (C.c_int, [_P, _u32, _P, _u32, _P, _P, C.POINTER(_P)]),
    "mlem_batch_run": (C.c_int, [_P, C.c_int, C.POINTER(C.c_int)]),
    "mlem_batch_image": (C.c_int, [_P, _u32, _P, _u32]),
    "mlem_batch_frame_info": (C.c_int, [_P, _u32, C.POINTER(C.c_int), _P, C.c_int]),
    "mlem_batch_progress": (C.c_int, [_P, C.POINTER(C.c_uint32), C.POINTER(C.c_uint64),
                                      C.POINTER(C.c_int), C.POINTER(C.c_uint32)]),
    "mlem_batch_tuning": (C.c_int, [_P, _P]),
    "mlem_batch_single_ms": (C.c_int, [_P, C.POINTER(C.c_double)]),
    "mlem_batch_destroy": (None, [_P]),
    "mlem_forward_k": (C.c_int, [_P, _u32, _P, _u32, _P, _u32, C.c_int, C.c_int, _P,
                                 C.POINTER(_P)]),
    "mlem_forward_k_mode": (C.c_int, [_P, _u32, _P, _u32, _P, _u32, C.c_int, C.c_int, C.c_int,
                                      _P, C.POINTER(_P)]),
    "mlem_backward_k": (C.c_int, [_P, _P, _P, _P]),
    "mlem_tape_info": (C.c_int, [_P, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int),
                                 C.POINTER(C.c_int), C.POINTER(C.c_uint64)]),
    "mlem_tape_destroy": (None, [_P]),
    "mlem_calibration_copy": (C.c_int, [C.c_size_t, C.c_int, C.POINTER(C.c_float)]),
    "mlem_device_sync": (C.c_int, []),
}

SYNTH_SIGNATURES = {
    "mlem_synth_matrix": (C.c_int, [_u32, _u32, _u32, _u32, _u64, C.c_int,
                                    C.POINTER(_P), C.POINTER(_P), C.POINTER(C.c_uint64)]),
    "mlem_synth_free": (None, [_P]),
    "mlem_synth_phantom": (C.c_int, [_u32, _u32, _u32, _u32, _u64, C.c_float, C.c_float, _P, _P]),
    "mlem_synth_project": (C.c_int, [_u32, _P, _P, _P, _P, C.c_int]),
    "mlem_synth_poisson": (C.c_int, [_u32, _P, C.c_double, _u64, C.c_int, _P,
                                     C.POINTER(C.c_int64)]),
    "mlem_synth_uniform": (C.c_int, [_u64, _u64, _u64, _P]),
    "mlem_synth_siddon": (C.c_int, [_P, _P, _u32, _u32, _u32, _u32, _P, _P]),
    "mlem_synth_write_csr4": (C.c_int, [C.c_char_p, _u32, _u32, _u32, _u32, _u64, _P, _P]),
}

_lib = None
_synth = None


def _bind(path: Path, sigs: dict) -> C.CDLL:
    if not path.exists():
        raise ImportError(f"{path} is not built; run `python -m mlem_amd.build` "
                          "(or __graft_entry__.build())")
    lib = C.CDLL(str(path))
    for name, (res, args) in sigs.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        _lib = _bind(LIBMLEM, SIGNATURES)
    return _lib


def synth_lib() -> C.CDLL:
    global _synth
    if _synth is None:
        _synth = _bind(LIBSYNTH, SYNTH_SIGNATURES)
    return _synth


def check(rc: int) -> None:
    if rc != MLEM_OK:
        raise MlemError(rc, lib().mlem_last_error().decode(errors="replace"))
