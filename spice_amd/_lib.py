"""ctypes binding of the gfx950 simulator's C ABI (include/spice_hip.h).

The library is built in-tree (spice_amd/libspice_hip.so, see __graft_entry__.build or
`make -C spice_amd/csrc`). There is no fallback: if it is missing, importing the binding
raises, so nothing can silently run anywhere but on the GPU.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libspice_hip.so")

SPICE_OK = 0
SPICE_ERR_INVALID_ARGUMENT = 1
SPICE_ERR_BAD_ALLOC = 2
SPICE_ERR_RUNTIME = 3
SPICE_ERR_BAD_CAST = 4
SPICE_FLAG_FIXED_POINT = 1


class SpiceError(RuntimeError):
    """Raised for SPICE_ERR_RUNTIME (the reference's std::runtime_error)."""


class SpiceCastError(TypeError):
    """Raised for SPICE_ERR_BAD_CAST (the reference's std::bad_cast)."""


class Connection(ctypes.Structure):
    _fields_ = [("src_pop", ctypes.c_uint64), ("dst_pop", ctypes.c_uint64), ("p", ctypes.c_float)]


class Edge(ctypes.Structure):
    _fields_ = [("src_first", ctypes.c_int32), ("src_last", ctypes.c_int32),
                ("dst_first", ctypes.c_int32), ("dst_last", ctypes.c_int32), ("p", ctypes.c_float)]


class NetDesc(ctypes.Structure):
    _fields_ = [("pops", ctypes.POINTER(ctypes.c_uint64)), ("npops", ctypes.c_int32),
                ("conns", ctypes.POINTER(Connection)), ("nconns", ctypes.c_int32), ("delay", ctypes.c_int32),
                ("seed", ctypes.c_uint64), ("params", ctypes.c_float * 4)]


class Stats(ctypes.Structure):
    _fields_ = [(n, ctypes.c_uint64) for n in
                ("steps", "spikes", "events", "plastic_events", "flush_rows", "flush_synapses", "flush_slots",
                 "flush_written")]


class Launch(ctypes.Structure):
    _fields_ = [("tile_w", ctypes.c_int32), ("flush_slots", ctypes.c_int32), ("flush_grid", ctypes.c_int32),
                ("recv_block", ctypes.c_int32), ("recv_parts", ctypes.c_int32), ("reserved", ctypes.c_int32 * 3)]


class JobDesc(ctypes.Structure):
    _fields_ = [("model", ctypes.c_int32), ("net", NetDesc), ("dt", ctypes.c_float), ("flags", ctypes.c_uint32),
                ("steps", ctypes.c_int64)]


class JobResult(ctypes.Structure):
    _fields_ = [("id", ctypes.c_uint64), ("status", ctypes.c_int32), ("reserved0", ctypes.c_int32),
                ("steps", ctypes.c_int64), ("done", ctypes.c_int64),
                ("synapses", ctypes.c_uint64), ("device_bytes", ctypes.c_uint64), ("reserved_bytes", ctypes.c_uint64),
                ("spikes", ctypes.c_uint64), ("events", ctypes.c_uint64), ("plastic_events", ctypes.c_uint64),
                ("digest", ctypes.c_uint64),
                ("t_submit", ctypes.c_double), ("t_admit", ctypes.c_double), ("t_done", ctypes.c_double)]


class ServerInfo(ctypes.Structure):
    _fields_ = [("budget", ctypes.c_uint64), ("reserved", ctypes.c_uint64), ("peak_reserved", ctypes.c_uint64),
                ("queued", ctypes.c_int64), ("resident", ctypes.c_int64), ("finished", ctypes.c_int64),
                ("chunks", ctypes.c_int64), ("rebuilds", ctypes.c_int64), ("steps_launched", ctypes.c_int64),
                ("busy_s", ctypes.c_double), ("now", ctypes.c_double), ("launch_s", ctypes.c_double),
                ("admit_s", ctypes.c_double), ("wait_s", ctypes.c_double), ("collect_s", ctypes.c_double)]


SPIKE_SINK = ctypes.CFUNCTYPE(None, ctypes.c_void_p, ctypes.c_uint64, ctypes.c_int64, ctypes.c_int64,
                              ctypes.POINTER(ctypes.c_uint64), ctypes.POINTER(ctypes.c_int32))
JOB_QUEUED, JOB_RUNNING, JOB_DONE, JOB_FAILED, JOB_CANCELLED = range(5)

_P = ctypes.c_void_p
_I32, _U32, _I64, _U64, _F32 = ctypes.c_int32, ctypes.c_uint32, ctypes.c_int64, ctypes.c_uint64, ctypes.c_float
_PI32, _PU32, _PI64, _PU64, _PF32 = (ctypes.POINTER(t) for t in (_I32, _U32, _I64, _U64, _F32))

# name -> (restype, argtypes)
SIGNATURES = {
    "spice_last_error": (ctypes.c_char_p, []),
    "spice_version": (ctypes.c_char_p, []),
    "spice_layout_flatten": (ctypes.c_int, [_PU64, _I32, ctypes.POINTER(Connection), _I32,
                                            ctypes.POINTER(Edge), _PI32, _PU64, _PU64]),
    "spice_model_info": (ctypes.c_int, [_I32, _PI32, _PI32, _PI32, _PI32]),
    "spice_hip_create": (ctypes.c_int, [_I32, _PU64, _I32, ctypes.POINTER(Connection), _I32, _F32, _I32,
                                        _U64, _U32, ctypes.POINTER(_P)]),
    "spice_hip_create_desc": (ctypes.c_int, [_I32, ctypes.POINTER(NetDesc), _F32, _U32, ctypes.POINTER(_P)]),
    "spice_batch_create": (ctypes.c_int, [_I32, _I32, ctypes.POINTER(NetDesc), _F32, _U32, ctypes.POINTER(_P)]),
    "spice_batch_destroy": (ctypes.c_int, [_P]),
    "spice_batch_run": (ctypes.c_int, [_P, _I64]),
    "spice_batch_synchronize": (ctypes.c_int, [_P]),
    "spice_batch_size": (ctypes.c_int, [_P, _PI32]),
    "spice_batch_net": (ctypes.c_int, [_P, _I32, ctypes.POINTER(_P)]),
    "spice_hip_get_launch": (ctypes.c_int, [_P, ctypes.POINTER(Launch)]),
    "spice_hip_set_launch": (ctypes.c_int, [_P, ctypes.POINTER(Launch)]),
    "spice_hip_autotune": (ctypes.c_int, [_P, _I64, _I64, ctypes.POINTER(Launch), ctypes.POINTER(ctypes.c_double)]),
    "spice_hip_probe_random_lines": (ctypes.c_int, [_U64, _U64, ctypes.c_int32, ctypes.c_int32, ctypes.POINTER(ctypes.c_double)]),
    "spice_server_create": (ctypes.c_int, [_U64, _I64, SPIKE_SINK, _P, ctypes.POINTER(_P)]),
    "spice_server_destroy": (ctypes.c_int, [_P]),
    "spice_server_estimate": (ctypes.c_int, [ctypes.POINTER(JobDesc), _I64, _PU64]),
    "spice_server_submit": (ctypes.c_int, [_P, ctypes.POINTER(JobDesc), _PU64]),
    "spice_server_wait": (ctypes.c_int, [_P, _U64, ctypes.c_double, ctypes.POINTER(JobResult), _PI32]),
    "spice_server_cancel": (ctypes.c_int, [_P, _U64]),
    "spice_server_drain": (ctypes.c_int, [_P]),
    "spice_server_get_info": (ctypes.c_int, [_P, ctypes.POINTER(ServerInfo)]),
    "spice_server_job_error": (ctypes.c_int, [_P, _U64, ctypes.c_char_p, _U64]),
    "spice_hip_create_from_ell": (ctypes.c_int, [_I32, _PI32, _U64, _U64, _F32, _I32, _U64, _U32,
                                                 ctypes.POINTER(_P)]),
    "spice_hip_create_from_csr": (ctypes.c_int, [_I32, _PI64, _PI32, _U64, _U64, _F32, _I32, _U64, _U32,
                                                 ctypes.POINTER(_P)]),
    "spice_hip_destroy": (ctypes.c_int, [_P]),
    "spice_hip_step": (ctypes.c_int, [_P, ctypes.POINTER(_P), ctypes.POINTER(_P)]),
    "spice_hip_copy_spikes": (ctypes.c_int, [_P, _PI32, _U64, _PU64]),
    "spice_hip_run": (ctypes.c_int, [_P, _I64]),
    "spice_hip_run_profiled": (ctypes.c_int, [_P, _I64, ctypes.POINTER(ctypes.c_double)]),
    "spice_hip_synchronize": (ctypes.c_int, [_P]),
    "spice_hip_num_neurons": (ctypes.c_int, [_P, _PU64]),
    "spice_hip_num_synapses": (ctypes.c_int, [_P, _PU64]),
    "spice_hip_width": (ctypes.c_int, [_P, _PU64]),
    "spice_hip_num_edges": (ctypes.c_int, [_P, _PU64]),
    "spice_hip_dt": (ctypes.c_int, [_P, _PF32]),
    "spice_hip_delay": (ctypes.c_int, [_P, _PI32]),
    "spice_hip_model": (ctypes.c_int, [_P, _PI32]),
    "spice_hip_get_clock": (ctypes.c_int, [_P, _PI64, _PF32, _PF32]),
    "spice_hip_set_clock": (ctypes.c_int, [_P, _I64, _F32, _F32]),
    "spice_hip_get_stats": (ctypes.c_int, [_P, ctypes.POINTER(Stats)]),
    "spice_hip_reset_stats": (ctypes.c_int, [_P]),
    "spice_hip_export_adj": (ctypes.c_int, [_P, _PI32, _U64]),
    "spice_hip_export_csr": (ctypes.c_int, [_P, _PI64, _PI32]),
    "spice_hip_generate_graph": (ctypes.c_int, [_PU64, _I32, ctypes.POINTER(Connection), _I32, _U64, _P, _P, _U64, _PU64, _PU64]),
    "spice_hip_generate_graph_noalloc": (ctypes.c_int, [_PU64, _I32, ctypes.POINTER(Connection), _I32, _U64, _P, _P, _U64, _PU64, _PU64]),
    "spice_hip_export_csr_rows": (ctypes.c_int, [_P, ctypes.c_uint64, ctypes.c_uint64, _PI64, _PI32, ctypes.c_uint64]),
    "spice_hip_export_neuron_attr": (ctypes.c_int, [_P, _I32, _P, _U64]),
    "spice_hip_import_neuron_attr": (ctypes.c_int, [_P, _I32, _P, _U64]),
    "spice_hip_export_synapse_attr": (ctypes.c_int, [_P, _I32, _P, _U64, _I32]),
    "spice_hip_import_synapse_attr": (ctypes.c_int, [_P, _I32, _P, _U64, _I32]),
}

_lib = None


def lib():
    """Loads libspice_hip.so (once). Raises OSError if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise OSError(f"{LIB_PATH} not built: run `make -C spice_amd/csrc` or __graft_entry__.build()")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(status):
    """Maps a spice_status to the Python counterpart of the reference's exception types."""
    if status == SPICE_OK:
        return
    msg = lib().spice_last_error().decode(errors="replace")
    if status == SPICE_ERR_INVALID_ARGUMENT:
        raise ValueError(msg)
    if status == SPICE_ERR_BAD_ALLOC:
        raise MemoryError(msg)
    if status == SPICE_ERR_BAD_CAST:
        raise SpiceCastError(msg)
    raise SpiceError(msg)
