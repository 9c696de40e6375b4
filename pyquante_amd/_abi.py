"""ctypes binding of libpqhip.so (C ABI in include/pq_abi.h).

The library is required: importing this module raises ImportError when the
HIP library is not built, and every call that needs a GPU fails with
PQError when none is visible.  There is no CPU fallback.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libpqhip.so")

PQ_PREC_FP64 = 0
PQ_PREC_MIXED = 1
PQ_CAP_STAGE_D = 4096  # bra-run cap flag: stage the bra rows of D in LDS

if not os.path.exists(LIB_PATH):
    raise ImportError("libpqhip.so is not built (run __graft_entry__.build() or make -C pyquante_amd/csrc)")

# torch first: the process then has one HIP runtime, the one torch's device
# memory and streams live in, and libpqhip.so binds to it.  Loaded first,
# libpqhip.so would bring in the system runtime, and torch would then find no
# device.
import torch  # noqa: E402,F401

lib = ctypes.CDLL(LIB_PATH)

_c = ctypes
_D, _I, _I64, _U64, _VP = _c.c_double, _c.c_int, _c.c_int64, _c.c_uint64, _c.c_void_p
_PD, _PI, _PU64 = _c.POINTER(_D), _c.POINTER(_I), _c.POINTER(_U64)


class PQError(RuntimeError):
    """A libpqhip call failed; .code holds the PQ_E* status."""

    def __init__(self, code, msg):
        RuntimeError.__init__(self, "%s (code %d)" % (msg, code))
        self.code = code


def _proto(name, argtypes, restype=_I):
    f = getattr(lib, name)
    f.argtypes = argtypes
    f.restype = restype
    return f


_proto("pq_last_error", [], _c.c_char_p)
_proto("pq_version", [])
_proto("pq_device_count", [])
_proto("pq_ijkl2intindex", [_I64, _I64, _I64, _I64, _PU64])
_proto("pq_boys_batch", [_I, _PD, _PD, _I, _PD])
_proto("pq_rys_roots_batch", [_I, _I, _PD, _PD, _PD])
_proto("pq_contr_coulomb_batch", [_I, _PI, _PI, _I, _PD, _PD, _PD, _PD, _PI, _PD])
_proto("pq_tho_contr_coulomb_batch", [_I, _PI, _PI, _I, _PD, _PD, _PD, _PD, _PI, _I, _PD])
_proto("pq_hgp_contr_coulomb_batch", [_I, _PI, _PI, _I, _PD, _PD, _PD, _PD, _PI, _PI, _I, _PD])
_proto("pq_engine_create", [_I, _PI, _PI, _PI, _PI, _PD, _PD, _I, _PD, _PD, _c.POINTER(_VP)])
_proto("pq_engine_create_batch", [_I, _PI, _PI, _PI, _PI, _PD, _PD, _I, _PD, _PD, _PI, _c.POINTER(_VP)])
_proto("pq_engine_layout", [_VP, _PI, _c.POINTER(_I64), _PI, _c.POINTER(_I64), _PI, _PI])
_proto("pq_engine_destroy", [_VP])
_proto("pq_engine_nbf", [_VP, _PI])
_proto("pq_engine_generation", [_VP, _PU64])
_proto("pq_engine_set_geometry", [_VP, _PD])
_proto("pq_build_pairs", [_VP, _D, _D])
_proto("pq_engine_set_streams", [_VP, _I])
_proto("pq_engine_set_lds_limit", [_VP, _I, _PI])
_proto("pq_engine_set_fused", [_VP, _I])
_proto("pq_engine_set_deterministic", [_VP, _I])
_proto("pq_engine_set_prim_screen", [_VP, _D])
_proto("pq_engine_set_group_cap", [_VP, _I])
_proto("pq_engine_set_group_caps", [_VP, _PI])
_proto("pq_engine_autotune", [_VP, _VP, _I, _VP, _PI])
_proto("pq_engine_stats", [_VP, _PU64, _PD])
_proto("pq_get2ints", [_VP, _VP, _U64, _VP])
_proto("pq_get2ints_hgp", [_VP, _VP, _U64, _I, _VP])
_proto("pq_shell_quartet_blocks", [_VP, _I, _PI, _VP, _U64, _VP])
_proto("pq_fock_jk", [_VP, _VP, _VP, _VP, _I, _VP])
_proto("pq_fock_jk_multi", [_VP, _I, _c.POINTER(_VP), _c.POINTER(_VP), _c.POINTER(_VP), _I, _VP])
_proto("pq_engine_class_stats", [_VP, _I, _PI, _PU64, _PD])
_proto("pq_fock_jk_timed", [_VP, _VP, _VP, _VP, _I, _VP, _c.POINTER(_c.c_float)])
_proto("pq_digest_stored", [_I, _VP, _U64, _VP, _VP, _VP, _VP])
_proto("pq_digest_stored_noalloc", [_I, _VP, _U64, _VP, _VP, _VP, _VP])
_proto("pq_unpack_eri", [_I, _VP, _U64, _I, _I, _VP, _VP])
_proto("pq_get1ints", [_VP, _VP, _VP, _VP, _I, _VP])
_proto("pq_forces", [_VP, _VP, _VP, _VP, _VP])
_proto("pq_forces_uhf", [_VP, _VP, _VP, _VP, _VP, _VP, _VP])
_proto("pq_forces_timed", [_VP, _VP, _VP, _VP, _VP, _c.POINTER(_c.c_float)])


def check(rc):
    if rc != 0:
        raise PQError(rc, lib.pq_last_error().decode())
    return rc


def pd(a):
    return a.ctypes.data_as(_PD)


def pi(a):
    return a.ctypes.data_as(_PI)


def require_gpu():
    if lib.pq_device_count() < 1:
        raise PQError(-5, "no HIP device visible: the pyquante_amd engine runs on the GPU only")
