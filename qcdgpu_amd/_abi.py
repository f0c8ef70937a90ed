"""ctypes binding of libqcdhip.so (declarations in include/qcd_hip.h).

The library is the product: this module only loads it and declares argument
types.  It raises if the library is missing -- there is no CPU fallback.
"""
import ctypes
import os
import re

PKG = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(PKG)
LIB_PATH = os.path.join(PKG, "libqcdhip.so")
HEADER = os.path.join(REPO, "include", "qcd_hip.h")

QCD_OK = 0
FP32, FP64 = 0, 1
START_HOT, START_COLD, START_GID, START_RANDOM = 0, 1, 2, 3
PRNG_XOR128, PRNG_RANLUX, PRNG_PM = 1, 2, 3
STREAM_REFERENCE, STREAM_FAST = 0, 1
HEATBATH, OVERRELAX = 0, 1
MEAS_PLAQUETTE, MEAS_POLYAKOV, MEAS_WILSON, MEAS_POLYAKOV_SLICES = 1, 2, 4, 8
MEAS_FMUNU, MEAS_F0MU, MEAS_ACTION_SLICES = 16, 32, 64
TUNE_XCD_MAP, TUNE_BLOCKS_PER_CU, TUNE_UPDATE_SHAPE, TUNE_SHADOW, TUNE_CYCLES_PER_GRAPH = 1, 2, 3, 4, 5
BUFFER_GLOBAL, BUFFER_INPUT, BUFFER_IO, BUFFER_OUTPUT, BUFFER_CONSTANT, BUFFER_LDS, BUFFER_UAV = 1, 2, 3, 4, 5, 6, 7


class Params(ctypes.Structure):
    _fields_ = [
        ("sun", ctypes.c_int),
        ("n", ctypes.c_int * 4),
        ("precision", ctypes.c_int),
        ("storage_reals", ctypes.c_int),
        ("beta", ctypes.c_double),
        ("nhit", ctypes.c_int),
        ("nhitpar", ctypes.c_int),
        ("prng", ctypes.c_int),
        ("ranlux_luxury", ctypes.c_int),
        ("stream_mode", ctypes.c_int),
        ("randseries", ctypes.c_uint),
        ("max_measurements", ctypes.c_int),
        ("prng_instances", ctypes.c_uint),
    ]


class ChainDesc(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int * 4), ("beta", ctypes.c_double), ("randseries", ctypes.c_uint)]


class FlowDesc(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int * 4), ("layers", ctypes.c_int), ("hidden", ctypes.c_int), ("bins", ctypes.c_int),
                ("max_batch", ctypes.c_int)]


class RunInfo(ctypes.Structure):
    """qcd_run_info: the run counters of a .qcg header."""

    _fields_ = [
        (name, ctypes.c_uint)
        for name in (
            "init", "start", "prng_counter", "nav", "nav_counter", "niter", "iter", "iter_counter",
            "wilson_r", "wilson_t", "pl_level", "get_actions", "get_plaquettes", "get_wilson",
            "fmunu_index1", "fmunu_index2", "get_fmunu", "get_f0mu",
        )
    ] + [("phi", ctypes.c_double), ("omega", ctypes.c_double)]

    def as_dict(self):
        return {name: (float if t is ctypes.c_double else int)(getattr(self, name)) for name, t in self._fields_}


class QCDError(RuntimeError):
    pass


_lib = None
_P = ctypes.c_void_p
_D = ctypes.POINTER(ctypes.c_double)
_I = ctypes.POINTER(ctypes.c_int)

_SIGNATURES = {
    "qcd_device_open": [ctypes.c_int, ctypes.c_int],
    "qcd_device_close": [],
    "qcd_device_info": [ctypes.c_char_p, ctypes.c_int, _I, ctypes.POINTER(ctypes.c_size_t)],
    "qcd_abi_version": [],
    "qcd_lattice_create": [ctypes.POINTER(Params), ctypes.POINTER(_P)],
    "qcd_lattice_destroy": [_P],
    "qcd_lattice_sync": [_P],
    "qcd_lattice_table_bytes": [_P, ctypes.POINTER(ctypes.c_size_t)],
    "qcd_lattice_tune": [_P, ctypes.c_int, ctypes.c_int],
    "qcd_kernel_create": [_P, ctypes.c_char_p, ctypes.POINTER(ctypes.c_int)],
    "qcd_kernel_run": [_P, ctypes.c_int],
    "qcd_kernel_enqueue": [_P, ctypes.c_int],
    "qcd_kernel_info": [_P, ctypes.c_int, ctypes.c_char_p, ctypes.c_int, ctypes.POINTER(ctypes.c_uint)],
    "qcd_graph_begin": [_P],
    "qcd_graph_end": [_P],
    "qcd_buffer_create": [ctypes.c_int, ctypes.c_size_t, ctypes.c_size_t, ctypes.c_void_p, ctypes.POINTER(ctypes.c_int)],
    "qcd_buffer_write": [ctypes.c_int],
    "qcd_buffer_map": [ctypes.c_int, ctypes.POINTER(ctypes.c_void_p)],
    "qcd_buffer_unmap": [ctypes.c_int],
    "qcd_buffer_info": [ctypes.c_int, ctypes.POINTER(ctypes.c_size_t), ctypes.POINTER(ctypes.c_void_p)],
    "qcd_buffer_kill": [ctypes.c_int],
    "qcd_buffer_size_align": [ctypes.c_uint, ctypes.c_uint],
    "qcd_lattice_profile": [_P, ctypes.c_int],
    "qcd_profile_count": [_P, ctypes.POINTER(ctypes.c_int)],
    "qcd_profile_entry": [_P, ctypes.c_int, ctypes.c_char_p, ctypes.c_int, ctypes.POINTER(ctypes.c_uint), _D, _D],
    "qcd_lattice_set_slab": [_P, ctypes.c_int],
    "qcd_slab_face_bytes": [_P, ctypes.POINTER(ctypes.c_size_t)],
    "qcd_slab_face": [_P, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_int],
    "qcd_lattice_get_tune": [_P, ctypes.c_int, _I, _I],
    "qcd_lattice_init": [_P, ctypes.c_int],
    "qcd_prng_init": [_P],
    "qcd_lattice_import": [_P, _D],
    "qcd_lattice_export": [_P, _D],
    "qcd_halfstep": [_P, ctypes.c_int, ctypes.c_int, ctypes.c_int],
    "qcd_sweep": [_P, ctypes.c_int],
    "qcd_reunitarize": [_P],
    "qcd_measure": [_P, ctypes.c_int],
    "qcd_measure_noalloc": [_P, ctypes.c_int],
    "qcd_read_measurements": [_P, _D, _D, _I, _I],
    "qcd_reset_measurements": [_P],
    "qcd_lattice_set_wilson": [_P, ctypes.c_int, ctypes.c_int],
    "qcd_read_wilson": [_P, _D, _I],
    "qcd_read_polyakov_slices": [_P, _D, _I],
    "qcd_lattice_set_fmunu": [_P, ctypes.c_int, ctypes.c_int],
    "qcd_lattice_set_twist": [_P, ctypes.c_double, ctypes.c_double],
    "qcd_lattice_get_twist": [_P, _D, _D],
    "qcd_read_fmunu": [_P, _D, _I],
    "qcd_read_action_slices": [_P, _D, _I],
    "qcd_cycle_capture": [_P, ctypes.c_int, ctypes.c_int, ctypes.c_int],
    "qcd_cycle_run": [_P, ctypes.c_int, ctypes.c_int, ctypes.c_int],
    "qcd_cycle_launch": [_P, ctypes.c_int],
    "qcd_cycle_time": [_P, ctypes.c_int, _D],
    "qcd_cycle_profile": [_P, ctypes.c_int, _D, _D, _D, _D],
    "qcd_batch_create": [ctypes.POINTER(Params), ctypes.c_int, ctypes.POINTER(ChainDesc), ctypes.POINTER(_P)],
    "qcd_batch_destroy": [_P],
    "qcd_batch_info": [_P, _I, _I],
    "qcd_batch_init": [_P, ctypes.c_int],
    "qcd_batch_sweep": [_P, ctypes.c_int],
    "qcd_batch_cycle_run": [_P, ctypes.c_int, ctypes.c_int, ctypes.c_int],
    "qcd_batch_cycle_capture": [_P, ctypes.c_int, ctypes.c_int, ctypes.c_int],
    "qcd_batch_cycle_launch": [_P, ctypes.c_int],
    "qcd_batch_cycle_time": [_P, ctypes.c_int, _D],
    "qcd_batch_sync": [_P],
    "qcd_batch_autotune": [_P, ctypes.c_int, _I],
    "qcd_batch_get_tuning": [_P, ctypes.c_int, ctypes.POINTER(ctypes.c_int)],
    "qcd_batch_set_tuning": [_P, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int],
    "qcd_batch_set_merge": [_P, ctypes.c_int],
    "qcd_batch_chain": [_P, ctypes.c_int, ctypes.POINTER(_P)],
    "qcd_lattice_get_params": [_P, ctypes.POINTER(Params)],
    "qcd_prng_counter": [_P, ctypes.POINTER(ctypes.c_uint)],
    "qcd_prng_replay": [_P, ctypes.c_uint, ctypes.c_int],
    "qcd_lattice_state_bytes": [_P, ctypes.POINTER(ctypes.c_size_t)],
    "qcd_lattice_state_save": [_P, ctypes.c_void_p, ctypes.c_size_t],
    "qcd_lattice_state_load": [_P, ctypes.c_void_p, ctypes.c_size_t],
    "qcd_write_measurements": [_P, _D, _D, _D, ctypes.c_int, ctypes.c_int, ctypes.c_int],
    "qcd_qcg_write": [_P, ctypes.c_char_p, ctypes.POINTER(RunInfo)],
    "qcd_qcg_read_header": [ctypes.c_char_p, ctypes.POINTER(Params), ctypes.POINTER(RunInfo), _I],
    "qcd_qcg_read": [_P, ctypes.c_char_p, ctypes.POINTER(RunInfo)],
    "qcd_flow_create": [ctypes.POINTER(FlowDesc), ctypes.POINTER(_P)],
    "qcd_flow_destroy": [_P],
    "qcd_flow_param_count": [_P, ctypes.POINTER(ctypes.c_size_t)],
    "qcd_flow_set_params": [_P, ctypes.POINTER(ctypes.c_float), ctypes.c_size_t],
    "qcd_flow_sample": [_P, ctypes.c_ulonglong, ctypes.c_int, ctypes.c_double, ctypes.POINTER(ctypes.c_float), _D, _D],
    "qcd_flow_evaluate": [_P, ctypes.POINTER(ctypes.c_float), ctypes.c_int, ctypes.c_double, _D, _D],
    "qcd_flow_imh_sweep": [_P, _P, ctypes.c_ulonglong, ctypes.c_double],
    "qcd_flow_imh_sweep_next": [_P, _P, ctypes.c_double],
    "qcd_flow_imh_seed": [_P, ctypes.c_ulonglong],
    "qcd_flow_train_grad": [_P, ctypes.POINTER(ctypes.c_float), ctypes.c_int, _D, ctypes.POINTER(ctypes.c_float)],
    "qcd_flow_train_step": [_P, ctypes.POINTER(ctypes.c_float), ctypes.c_int, ctypes.c_float, _D],
    "qcd_flow_get_params": [_P, ctypes.POINTER(ctypes.c_float), ctypes.c_size_t],
    "qcd_flow_imh_counts": [_P, ctypes.POINTER(ctypes.c_ulonglong), ctypes.POINTER(ctypes.c_ulonglong), ctypes.c_int],
    "qcd_prng_produce_host": [
        ctypes.c_int,
        ctypes.c_int,
        ctypes.c_uint,
        ctypes.c_uint,
        ctypes.c_void_p,
        ctypes.POINTER(ctypes.c_float),
    ],
}


def header_symbols():
    """Function names declared in include/qcd_hip.h."""
    text = open(HEADER).read()
    return sorted(set(re.findall(r"\b(qcd_[a-z0-9_]+)\s*\(", text)))


def lib():
    """Load libqcdhip.so; raise if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise QCDError(
            "libqcdhip.so not found at %s: build it with __graft_entry__.build() "
            "(there is no CPU fallback)" % LIB_PATH
        )
    L = ctypes.CDLL(LIB_PATH)
    for name, args in _SIGNATURES.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = ctypes.c_int
    L.qcd_last_error.restype = ctypes.c_char_p
    L.qcd_last_error.argtypes = []
    L.qcd_buffer_size_align.restype = ctypes.c_uint
    _lib = L
    return L


def check(rc):
    if rc != QCD_OK:
        raise QCDError("qcd error %d: %s" % (rc, lib().qcd_last_error().decode()))
    return rc
