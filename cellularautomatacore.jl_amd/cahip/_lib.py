"""ctypes binding of ``libcahip.so`` (the C ABI declared in include/cahip.h).

The library is required: importing the backend raises if it is missing, so no
caller can silently fall back to a CPU path.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libcahip.so")

# error codes (include/cahip.h)
CAH_OK = 0
CAH_ERR_ARGUMENT = -1
CAH_ERR_INCOMPATIBLE_SPACE = -2
CAH_ERR_UNSUPPORTED = -3
CAH_ERR_OUT_OF_MEMORY = -4
CAH_ERR_STATE_SHAPE = -5
CAH_ERR_INVALID_STATE = -6
CAH_ERR_DEVICE = -7
CAH_ERR_BUSY = -8

DTYPE_I8, DTYPE_F32 = 1, 2
BOUNDARY_PERIODIC, BOUNDARY_REFLECTING, BOUNDARY_CONSTANT = 0, 1, 2
SCHEME_SYNCHRONOUS, SCHEME_STOCHASTIC, SCHEME_CHECKERBOARD = 0, 1, 2
RULE_LIFE, RULE_GENERATIONS, RULE_FOREST_FIRE, RULE_LENIA, RULE_FIELD_EPIDEMIC = 1, 2, 3, 4, 5
EPI_BERNOULLI, EPI_STATED = 0, 1
FIELD_DIRECT_SEQ, FIELD_DIRECT_F32, FIELD_MFMA_F16, FIELD_MFMA_BF16, FIELD_FFT_F32, \
    FIELD_MFMA_F16X2 = 0, 1, 2, 3, 4, 5


class CahDesc(ctypes.Structure):
    _fields_ = [
        ("ndims", ctypes.c_int32),
        ("dtype", ctypes.c_int32),
        ("dims", ctypes.c_int64 * 2),
        ("boundary", ctypes.c_int32 * 2),
        ("boundary_value", ctypes.c_double * 2),
        ("scheme", ctypes.c_int32),
        ("scheme_p_is_f32", ctypes.c_int32),
        ("scheme_p", ctypes.c_double),
        ("seed", ctypes.c_uint64),
        ("rule", ctypes.c_int32),
        ("gen_states", ctypes.c_int32),
        ("birth", ctypes.c_uint64),
        ("survive", ctypes.c_uint64),
        ("n_offsets", ctypes.c_int32),
        ("lenia_precision", ctypes.c_int32),
        ("offsets", ctypes.POINTER(ctypes.c_int32)),
        ("n_taps", ctypes.c_int32),
        ("field_method", ctypes.c_int32),
        ("tap_offsets", ctypes.POINTER(ctypes.c_int32)),
        ("tap_weights", ctypes.POINTER(ctypes.c_double)),
        ("mu", ctypes.c_double),
        ("sigma", ctypes.c_double),
        ("dt", ctypes.c_double),
        ("mask", ctypes.POINTER(ctypes.c_uint8)),
        ("epi_form", ctypes.c_int32),
        ("epi_recover", ctypes.c_int32),
        ("covariate", ctypes.POINTER(ctypes.c_double)),
        ("covariate_slices", ctypes.c_int64),
    ]


class CahParams(ctypes.Structure):
    _fields_ = [
        ("ff_p", ctypes.c_double),
        ("ff_f", ctypes.c_double),
        ("ff_p_is_f32", ctypes.c_int32),
        ("ff_f_is_f32", ctypes.c_int32),
    ]


_P = ctypes.c_void_p
_I = ctypes.c_int
_I32 = ctypes.c_int32
_I64 = ctypes.c_int64
_U64 = ctypes.c_uint64
_SZ = ctypes.c_size_t
_D = ctypes.c_double
_PD = ctypes.POINTER(CahDesc)
_PP = ctypes.POINTER(CahParams)

# name -> (restype, argtypes); every symbol include/cahip.h declares
SIGNATURES = {
    "cah_open": (_I, [_I, ctypes.POINTER(_P)]),
    "cah_close": (None, [_P]),
    "cah_last_error": (ctypes.c_char_p, []),
    "cah_supports": (_I, [_I]),
    "cah_check_problem": (_I, [_PD]),
    "cah_state_create": (_I, [_P, _PD, ctypes.POINTER(_P)]),
    "cah_state_set": (_I, [_P, _P, _SZ]),
    "cah_state_get": (_I, [_P, _P, _SZ]),
    "cah_state_set_device": (_I, [_P, _P, _SZ]),
    "cah_state_get_device": (_I, [_P, _P, _SZ]),
    "cah_state_counts": (_I, [_P, _P, _P, _P]),
    "cah_state_read_cell": (_I, [_P, _I64, _I64, _P]),
    "cah_state_write_cell": (_I, [_P, _I64, _I64, _P]),
    "cah_advance": (_I, [_P, _I64, _I64, _PP]),
    "cah_state_sync": (_I, [_P]),
    "cah_state_stream": (_I, [_P, ctypes.POINTER(_P)]),
    "cah_state_tune": (_I, [_P, _I32, _I32]),
    "cah_state_sparse": (_I, [_P, ctypes.c_int8, _I32]),
    "cah_state_active_tiles": (_I, [_P, ctypes.POINTER(_I64)]),
    "cah_state_destroy": (_I, [_P]),
    "cah_state_fill_bernoulli": (_I, [_P, _U64, _D]),
    "cah_state_fill_uniform_box": (_I, [_P, _U64, _I64, _I64, _I64, _I64]),
    "cah_batch_create": (_I, [_P, _I32, _PD, ctypes.POINTER(_P)]),
    "cah_batch_set": (_I, [_P, _I32, _P, _SZ]),
    "cah_batch_get": (_I, [_P, _I32, _P, _SZ]),
    "cah_batch_fill_bernoulli": (_I, [_P, _I32, _U64, _D]),
    "cah_batch_advance": (_I, [_P, _I64, _I64, _PP, _I32]),
    "cah_batch_counts": (_I, [_P, _P, _P, _P]),
    "cah_batch_counts_noalloc": (_I, [_P, _P, _P, _P]),
    "cah_batch_autotune": (_I, [_P, _I64, _I32, _PP, ctypes.POINTER(_I32), ctypes.POINTER(_I32)]),
    "cah_batch_tune": (_I, [_P, _I32, _I32]),
    "cah_batch_sync": (_I, [_P]),
    "cah_batch_stream": (_I, [_P, ctypes.POINTER(_P)]),
    "cah_batch_destroy": (_I, [_P]),
    "cah_check_launches": (_I, [_PD, _I64, _I64, _PP, ctypes.c_char_p, _SZ]),
    "cah_check_batch_launches": (_I, [_I32, _PD, _I64, _I64, _PP, _I32, _I32, ctypes.c_char_p, _SZ]),
    "cah_server_create": (_I, [_P, _SZ, _I32, ctypes.POINTER(_P)]),
    "cah_server_submit": (_I, [_P, _PD, _P, _SZ, _I64, _I64, _PP, ctypes.POINTER(_U64)]),
    "cah_server_tune": (_I, [_P, _I32, _D]),
    "cah_server_reserve": (_I, [_P, ctypes.c_size_t, ctypes.c_size_t]),
    "cah_server_tick": (_I, [_P, ctypes.POINTER(_I32), ctypes.POINTER(_I32)]),
    "cah_server_poll": (_I, [_P, ctypes.POINTER(_U64), _I32, ctypes.POINTER(_I32)]),
    "cah_server_status": (_I, [_P, _U64, ctypes.POINTER(_I32)]),
    "cah_server_fetch": (_I, [_P, _U64, _P, _SZ]),
    "cah_server_times": (_I, [_P, _U64, ctypes.POINTER(_D), ctypes.POINTER(_D),
                              ctypes.POINTER(_D)]),
    "cah_server_clock": (_D, []),
    "cah_server_destroy": (_I, [_P]),
}

REQ_PENDING, REQ_ACTIVE, REQ_DONE, REQ_FETCHED = 0, 1, 2, 3
COUNTS_NOALLOC_MAX = 512

_lib = None


def load():
    """The loaded library, with argument types set.  Raises if it is absent."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                f"{LIB_PATH} is missing: build it with `python -m cahip.build` "
                "(the HIP backend has no CPU fallback)")
        lib = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
    return _lib


def last_error():
    msg = load().cah_last_error()
    return msg.decode() if msg else ""
