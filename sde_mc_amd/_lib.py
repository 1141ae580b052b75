"""ctypes binding of libsdemc.so (the C ABI declared in include/sdemc.h).

The structures below mirror the header field for field; ``check_layout``
compares their sizes with the library's own ``sizeof`` exports.  Every call
returns an int status that :func:`check` turns into an exception.  There is
no fallback: when the library is missing, importing the solvers works but any
simulation raises :class:`HipUnavailableError`.
"""
import ctypes
import os

MAX_DIM = 4
CHUNK = 64

OK, ERR_ARG, ERR_UNSUPPORTED, ERR_LAUNCH, ERR_WORKSPACE, ERR_OVERFLOW, ERR_DEVICE = range(7)

ABI_VERSION = 5   # SDEMC_ABI_VERSION these structures follow
# stats flag bits (enum sdemc_flag)
FLAG_OVERFLOW, FLAG_JUMP_CSR, FLAG_CLASS, FLAG_TABLE = 1, 2, 4, 8

NOISE_DIAG, NOISE_INDEP_COMMON, NOISE_INDEP_FULL, NOISE_HESTON, NOISE_GENERAL = 0, 1, 2, 3, 4
JUMP_LOGNORMAL, JUMP_LEVY = 0, 1
GRID_UNIFORM, GRID_ADAPTED = 0, 1
(PAYOFF_EURO_CALL, PAYOFF_EURO_PUT, PAYOFF_BINARY_AON, PAYOFF_DIGITAL, PAYOFF_BASKET_ARITH,
 PAYOFF_BASKET_GEO, PAYOFF_RAINBOW, PAYOFF_BEST_OF, PAYOFF_IDENTITY, PAYOFF_ASIAN_CALL,
 PAYOFF_RAINBOW_EVEN) = range(11)

_D4 = ctypes.c_double * MAX_DIM


class Model(ctypes.Structure):
    _fields_ = [
        ("dim", ctypes.c_int32), ("noise", ctypes.c_int32), ("jump_law", ctypes.c_int32),
        ("exact_jumps", ctypes.c_int32), ("rate", ctypes.c_double),
        ("x0", _D4), ("a0", _D4), ("a1", _D4), ("b0", _D4), ("b1", _D4),
        ("c0", _D4), ("c1", _D4), ("j0", _D4), ("j1", _D4),
        ("chol", ctypes.c_double * (MAX_DIM * MAX_DIM)), ("jp", ctypes.c_double * 12),
        ("ax", _D4), ("brown_dim", ctypes.c_int32), ("reserved", ctypes.c_int32),
        ("bm0", ctypes.c_double * (MAX_DIM * MAX_DIM)),
        ("bm1", ctypes.c_double * (MAX_DIM * MAX_DIM)),
    ]


class Payoff(ctypes.Structure):
    _fields_ = [
        ("kind", ctypes.c_int32), ("log", ctypes.c_int32), ("strike", ctypes.c_double),
        ("scale", ctypes.c_double), ("disc", ctypes.c_double), ("param", ctypes.c_double),
    ]


class Job(ctypes.Structure):
    _fields_ = [
        ("model", Model), ("payoff", Payoff), ("T", ctypes.c_double),
        ("steps", ctypes.c_int32), ("grid", ctypes.c_int32), ("coarse_factor", ctypes.c_int32),
        ("max_steps", ctypes.c_int32), ("n_paths", ctypes.c_uint64), ("seed", ctypes.c_uint64),
        ("path_offset", ctypes.c_uint64),
    ]


class PathOut(ctypes.Structure):
    _fields_ = [(n, ctypes.c_void_p) for n in (
        "payoff_dev", "terminal_dev", "nsteps_dev", "paths_dev", "times_dev", "left_dev",
        "jumps_dev", "normals_dev")]


class FixedIn(ctypes.Structure):
    _fields_ = [
        ("dw_dev", ctypes.c_void_p), ("counts_dev", ctypes.c_void_p),
        ("jump_off_dev", ctypes.c_void_p), ("jump_time_dev", ctypes.c_void_p),
        ("jump_size_dev", ctypes.c_void_p), ("n_marks", ctypes.c_int64),
        ("n_rows", ctypes.c_int32), ("max_jumps", ctypes.c_int32),
    ]


SEGMENT_INTS = 6  # sdemc_segment: job, first, count, per_tile, tile_begin, reserved
MAX_GROUPS = 8    # class groups per batched launch
# dedicated classes 0..7: 4 * (dim-2 family) + 2 * grid + coupled; 8: the 2-D
# family on the uniform grid at one level with the arrival clock (4 has per-step
# counts); 9: the generic kernel
CLASS_INDEP2_CLOCK = 8
CLASS_GENERIC = 9


class BatchDev(ctypes.Structure):
    _fields_ = [
        ("jobs_dev", ctypes.c_void_p), ("segs_dev", ctypes.c_void_p),
        ("seg_hdr_dev", ctypes.c_void_p), ("finalize_dev", ctypes.c_void_p),
        ("partials_dev", ctypes.c_void_p), ("chunk_flags_dev", ctypes.c_void_p),
        ("stats_dev", ctypes.c_void_p), ("counter_dev", ctypes.c_void_p),
        ("partial_chunks", ctypes.c_int64), ("max_jobs", ctypes.c_int32),
        ("max_segments", ctypes.c_int32), ("n_groups", ctypes.c_int32),
        ("group_class", ctypes.c_int32 * MAX_GROUPS), ("reserved", ctypes.c_int32),
    ]


class BatchTune(ctypes.Structure):
    _fields_ = [("waves_per_block", ctypes.c_int32), ("blocks_per_cu", ctypes.c_int32),
                ("concurrent_groups", ctypes.c_int32), ("step_unroll", ctypes.c_int32)]


class SdemcError(RuntimeError):
    """A libsdemc entry point returned a non-zero status."""

    def __init__(self, status, where):
        self.status = status
        msg = "%s failed: %s (status %d)" % (where, status_string(status), status)
        if status == ERR_LAUNCH and _LIB is not None:
            msg += ", hip error %d" % _LIB.sdemc_last_hip_error()
        super().__init__(msg)


class HipUnavailableError(RuntimeError):
    """libsdemc.so is not built or no HIP device is present."""


_LIB = None
_LIB_ERR = None

_SIGS = {
    "sdemc_abi_version": (ctypes.c_int, []),
    "sdemc_status_string": (ctypes.c_char_p, [ctypes.c_int]),
    "sdemc_last_hip_error": (ctypes.c_int, []),
    "sdemc_sizeof_job": (ctypes.c_size_t, []),
    "sdemc_sizeof_model": (ctypes.c_size_t, []),
    "sdemc_sizeof_path_out": (ctypes.c_size_t, []),
    "sdemc_sizeof_fixed_in": (ctypes.c_size_t, []),
    "sdemc_sizeof_prepared_job": (ctypes.c_size_t, []),
    "sdemc_workspace_bytes": (ctypes.c_size_t, [ctypes.c_uint64]),
    "sdemc_device_info": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(ctypes.c_int),
                                         ctypes.POINTER(ctypes.c_int),
                                         ctypes.POINTER(ctypes.c_int)]),
    "sdemc_simulate": (ctypes.c_int, [ctypes.POINTER(Job), ctypes.c_void_p,
                                      ctypes.POINTER(PathOut), ctypes.c_void_p, ctypes.c_size_t,
                                      ctypes.c_int, ctypes.c_void_p]),
    "sdemc_simulate_noalloc": (ctypes.c_int, [ctypes.POINTER(Job), ctypes.c_void_p,
                                              ctypes.POINTER(PathOut), ctypes.c_int,
                                              ctypes.c_void_p]),
    "sdemc_simulate_fixed": (ctypes.c_int, [ctypes.POINTER(Job), ctypes.POINTER(FixedIn),
                                            ctypes.c_void_p, ctypes.POINTER(PathOut),
                                            ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p]),
    "sdemc_reduce_finalize": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                             ctypes.c_void_p, ctypes.c_int32, ctypes.c_void_p,
                                             ctypes.c_void_p]),
    "sdemc_job_class": (ctypes.c_int, [ctypes.POINTER(Job)]),
    "sdemc_cv_mlp_words": (ctypes.c_size_t, [ctypes.c_int32]),
    "sdemc_cv_simulate": (ctypes.c_int, [ctypes.POINTER(Job), ctypes.c_void_p, ctypes.c_void_p,
                                         ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t,
                                         ctypes.c_void_p]),
    "sdemc_prepare_jobs": (ctypes.c_int, [ctypes.POINTER(Job), ctypes.c_int32,
                                          ctypes.POINTER(ctypes.c_int64), ctypes.c_void_p,
                                          ctypes.POINTER(ctypes.c_int32)]),
    "sdemc_batch_check": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, ctypes.c_void_p,
                                         ctypes.c_int32, ctypes.c_void_p, ctypes.c_int32,
                                         ctypes.c_void_p, ctypes.c_int64]),
    "sdemc_batch_launch": (ctypes.c_int, [ctypes.POINTER(BatchDev), ctypes.POINTER(BatchTune),
                                          ctypes.c_void_p]),
    "sdemc_batch_plan_create": (ctypes.c_int, [ctypes.POINTER(BatchDev),
                                               ctypes.POINTER(BatchTune),
                                               ctypes.POINTER(ctypes.c_void_p)]),
    "sdemc_batch_plan_launch": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p]),
    "sdemc_batch_plan_destroy": (ctypes.c_int, [ctypes.c_void_p]),
    "sdemc_philox_dev": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint32,
                                        ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p]),
    "sdemc_valu_probe_dev": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32,
                                            ctypes.c_int32, ctypes.c_void_p]),
    "sdemc_hbm_read_probe_dev": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int64,
                                                ctypes.c_void_p, ctypes.c_int32,
                                                ctypes.c_void_p]),
    "sdemc_normals_dev": (ctypes.c_int, [ctypes.c_uint64, ctypes.c_uint64, ctypes.c_int64,
                                         ctypes.c_int32, ctypes.c_void_p, ctypes.c_void_p]),
}


def lib_path():
    return os.path.join(os.path.dirname(os.path.abspath(__file__)), "lib", "libsdemc.so")


def load(path=None):
    """Load libsdemc.so once and bind every entry point; raises if absent."""
    global _LIB, _LIB_ERR
    if _LIB is not None:
        return _LIB
    path = path or lib_path()
    if not os.path.exists(path):
        _LIB_ERR = "libsdemc.so not built (expected at %s); run __graft_entry__.build()" % path
        raise HipUnavailableError(_LIB_ERR)
    lib = ctypes.CDLL(path)
    for name, (res, args) in _SIGS.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _LIB = lib
    check_layout()
    return lib


def lib():
    return _LIB if _LIB is not None else load()


def status_string(status):
    if _LIB is None:
        return "status %d" % status
    return _LIB.sdemc_status_string(int(status)).decode()


def check(status, where):
    if status != OK:
        raise SdemcError(status, where)


def check_layout():
    """The ctypes mirrors must match the compiled structures byte for byte."""
    if _LIB.sdemc_abi_version() != ABI_VERSION:
        raise RuntimeError("ABI mismatch: library version %d, ctypes mirror %d"
                           % (_LIB.sdemc_abi_version(), ABI_VERSION))
    pairs = [("sdemc_sizeof_job", Job), ("sdemc_sizeof_model", Model),
             ("sdemc_sizeof_path_out", PathOut), ("sdemc_sizeof_fixed_in", FixedIn)]
    for fn, cls in pairs:
        got = getattr(_LIB, fn)()
        if got != ctypes.sizeof(cls):
            raise RuntimeError("ABI mismatch: %s = %d, ctypes %s = %d"
                               % (fn, got, cls.__name__, ctypes.sizeof(cls)))


def exported_symbols():
    return list(_SIGS)
