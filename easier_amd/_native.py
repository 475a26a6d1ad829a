"""ctypes bridge to libesrhip.so (the C ABI declared in include/esr_hip.h).

Everything device-side in easier_amd goes through this module. The library is
required: there is no fallback implementation, and a missing library raises
at first use.
"""
import ctypes
import os

import torch

from ._build import LIB_PATH

_c_i32 = ctypes.c_int
_c_i64 = ctypes.c_int64
_c_u32 = ctypes.c_uint32
_c_u64 = ctypes.c_uint64
_c_dbl = ctypes.c_double
_c_vp = ctypes.c_void_p
_c_sz = ctypes.c_size_t
_c_cp = ctypes.c_char_p
_P_i64 = ctypes.POINTER(ctypes.c_int64)

ESR_OK, ESR_ERR_INVALID, ESR_ERR_HIP, ESR_ERR_UNSUPPORTED, ESR_ERR_RTC, \
    ESR_ERR_INTERNAL = range(6)

F32, F64, I32, I64, U8, BF16 = range(6)
_DTYPE_CODES = {
    torch.float32: F32, torch.float64: F64, torch.int32: I32,
    torch.int64: I64, torch.bool: U8, torch.uint8: U8, torch.bfloat16: BF16,
}

RED_OPS = {'sum': 0, 'prod': 1, 'mean': 2, 'amax': 3, 'amin': 4}
AGG_SUM, AGG_PROD, AGG_NORM, AGG_MAX, AGG_MIN = range(5)
SEQ_ROW_MAX = 256
AGG_SCRATCH_BYTES = 131072   # ESR_AGG_SCRATCH_BYTES

# elementwise op codes (esr_hip.h ESR_EW_*)
EW = dict(
    copy=0, neg=1, abs=2, exp=3, log=4, sqrt=5, rsqrt=6, sin=7, cos=8, tanh=9,
    sign=10, reciprocal=11, floor=12, ceil=13, logical_not=14, sigmoid=15,
    relu=16, isnan=17,
    add=32, sub=33, mul=34, div=35, pow=36, maximum=37, minimum=38, lt=39,
    le=40, gt=41, ge=42, eq=43, ne=44, logical_and=45, logical_or=46,
    atan2=47, fmod=48, floordiv=49,
    where=64,
)
EW_PREDICATES = {'lt', 'le', 'gt', 'ge', 'eq', 'ne', 'logical_and',
                 'logical_or', 'isnan', 'logical_not'}

# name: (restype, argtypes)
_SIGNATURES = {
    'esr_abi_version': (_c_i32, []),
    'esr_last_error': (_c_cp, []),
    'esr_device_arch': (_c_i32, [_c_cp, _c_i32]),
    'esr_stream_sync': (_c_i32, [_c_vp]),
    'esr_gather': (_c_i32, [_c_i32, _c_vp, _c_i64, _c_i64, _c_vp, _c_i64,
                            _c_vp, _c_vp]),
    'esr_csr_reduce': (_c_i32, [_c_i32, _c_i32, _c_vp, _c_i64, _c_vp, _c_vp,
                                _c_i64, _c_vp, _c_i64, _c_i32, _c_vp, _c_vp]),
    'esr_aggregate_ws_bytes': (_c_i32, [_c_i32, _c_i64, _c_i64, _P_i64]),
    'esr_aggregate': (_c_i32, [_c_i32, _c_i32, _c_vp, _c_i64, _c_i64, _c_dbl,
                               _c_vp, _c_vp, _c_i64, _c_vp]),
    'esr_aggregate_noalloc': (_c_i32, [_c_i32, _c_i32, _c_vp, _c_i64, _c_i64,
                                       _c_dbl, _c_vp, _c_vp]),
    'esr_ewise': (_c_i32, [_c_i32, _c_i32, _c_vp, _c_i32, _c_vp, _c_vp,
                           _c_i32, _c_i32, _c_vp, _c_vp, _c_vp, _c_vp]),
    'esr_fill': (_c_i32, [_c_i32, _c_vp, _c_i64, _c_dbl, _c_vp]),
    'esr_gemv_rows': (_c_i32, [_c_i32, _c_vp, _c_i64, _c_i64, _c_i64, _c_i64,
                               _c_vp, _c_i64, _c_vp, _c_vp]),
    'esr_rtc_compile': (_c_i32, [_c_cp, _c_cp, _c_cp, ctypes.POINTER(_c_vp)]),
    'esr_rtc_check': (_c_i32, [_c_cp, _c_cp, _P_i64]),
    'esr_rtc_function': (_c_i32, [_c_vp, _c_cp, ctypes.POINTER(_c_vp)]),
    'esr_rtc_unload': (_c_i32, [_c_vp]),
    'esr_launch': (_c_i32, [_c_vp, _c_u32, _c_u32, _c_u32, _c_u32, _c_u32,
                            _c_u32, _c_u32, _c_vp, _c_vp, _c_sz]),
    'esr_graph_begin': (_c_i32, [_c_vp]),
    'esr_graph_end': (_c_i32, [_c_vp, ctypes.POINTER(_c_vp)]),
    'esr_graph_launch': (_c_i32, [_c_vp, _c_vp]),
    'esr_graph_destroy': (_c_i32, [_c_vp]),
    'esr_mesh_sizes': (_c_i32, [_c_i32, _c_i64, _P_i64, _P_i64, _P_i64,
                                _P_i64]),
    'esr_mesh_tri4': (_c_i32, [_c_i64] + [_c_vp] * 6),
    'esr_mesh_tri2': (_c_i32, [_c_i64, _c_dbl, _c_u64] + [_c_vp] * 6),
    'esr_csr_from_idx': (_c_i32, [_c_vp, _c_i64, _c_i64, _c_vp, _c_vp]),
    'esr_rcm': (_c_i32, [_c_i64, _c_vp, _c_vp, _c_i64, _c_vp]),
    'esr_bandwidth': (_c_i32, [_c_vp, _c_vp, _c_i64, _P_i64]),
    'esr_block_order': (_c_i32, [_c_i64, _c_vp, _c_vp, _c_i64, _c_i64, _c_vp]),
    'esr_p1_nnz': (_c_i32, [_c_i64, _c_i64, _c_vp, _P_i64]),
    'esr_p1_assemble': (_c_i32, [_c_i64, _c_i64, _c_vp, _c_vp, _c_vp, _c_vp,
                                 _c_i64] + [_c_vp] * 6),
}

_LIB = None
_EXTRA_SIGNATURES = {}


def register_signatures(sigs):
    """Add entry points declared by other modules (solver, SWE, network)."""
    _EXTRA_SIGNATURES.update(sigs)
    if _LIB is not None:
        _bind(_LIB, sigs)


def _bind(lib, sigs):
    for name, (res, args) in sigs.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args


def lib():
    """The loaded library (raises if it has not been built)."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"easier_amd native library not found at {LIB_PATH}; "
                "build it first (python -c 'import __graft_entry__ as g; "
                "g.build()' or python easier_amd/_build.py)")
        handle = ctypes.CDLL(LIB_PATH)
        _bind(handle, _SIGNATURES)
        _bind(handle, _EXTRA_SIGNATURES)
        _LIB = handle
    return _LIB


def exported_symbols():
    return sorted(set(_SIGNATURES) | set(_EXTRA_SIGNATURES))


class EsrNativeError(RuntimeError):
    pass


def check(rc: int, what: str = ''):
    if rc == ESR_OK:
        return
    msg = lib().esr_last_error().decode(errors='replace')
    if what:
        msg = f'{what}: {msg}'
    if rc == ESR_ERR_INVALID:
        raise ValueError(msg)
    raise EsrNativeError(msg)


def call(name, *args):
    check(getattr(lib(), name)(*args), name)


def dtype_code(dtype: torch.dtype) -> int:
    try:
        return _DTYPE_CODES[dtype]
    except KeyError:
        raise TypeError(f'dtype {dtype} is not supported by the hip backend')


def stream_ptr(stream=None) -> int:
    if stream is None:
        stream = torch.cuda.current_stream()
    return stream.cuda_stream


def ptr(t):
    return None if t is None else t.data_ptr()


def require_device(*tensors):
    for t in tensors:
        if t is not None and t.device.type != 'cuda':
            raise RuntimeError(
                f'the hip backend expects device tensors, got {t.device}')
