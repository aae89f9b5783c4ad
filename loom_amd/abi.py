"""ctypes binding of the C ABI declared in include/loom_amd.h.

This is the Python side of the drop-in boundary: plain pointers and sizes,
no torch types.  Loading fails loudly when the library was not built; there
is no fallback path.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

_LIB_PATH = Path(__file__).resolve().parent / "_build" / "libloom_amd.so"

MAX_IO = 32
MAX_BLOCKS = 64
NUM_COUNTERS = 16

MODE_COMPACT = 0
MODE_HYBRID = 1
# or'ed into loom_kernel_resources' policy: the pass kernel that regroups
RESOURCES_REGROUP = 0x100
RESOURCES_MEMORY = 0x200
RESOURCES_WORKGROUP = 0x400
RESOURCES_WORKGROUP_8 = 0x800
REGROUP_NONE = 0
REGROUP_QUEUES = 1
REGROUP_WORKGROUP = 2
REGFILE_VGPR = 0
REGFILE_MEMORY = 1

RECONV_SUCCESSOR_MAJOR = 0
RECONV_TOPO = 1
RECONV_SWEEP = 2

# loom_counter
C_ACTIVE_OPS, C_SLOT_OPS, C_DISPATCHES, C_THREAD_STEPS = 0, 1, 2, 3
C_SPILL_ROWS, C_SPILL_BYTES, C_FILL_ROWS, C_FILL_BYTES = 4, 5, 6, 7
C_ADMITTED, C_REFILLS, C_BATCHES, C_SPILL_FAILED = 8, 9, 10, 11
C_ACTIVE_FLOPS = 12

E_ARG, E_KERNEL, E_NODEVICE, E_RUNTIME, E_CAPACITY, E_STATE = range(1000, 1006)
COPY_H2D, COPY_D2H = 0, 1

ERR_ASSERT_FP = 1
ERR_STEP_BUDGET = 2

REG_INPUT, REG_OUTPUT, REG_WRITTEN = 1, 2, 4  # loom_kernel_register flags
TRACE_HALT = 0xFFFF  # next-block value of a trace row after the last block


def trace_words(n_registers: int) -> int:
    """64-bit words per step-trace row (LOOM_TRACE_WORDS)."""
    return 2 + n_registers


class KernelInfo(C.Structure):
    _fields_ = [
        ("name", C.c_char * 64),
        ("hash", C.c_uint64),
        ("n_inputs", C.c_int32),
        ("n_outputs", C.c_int32),
        ("input_type", C.c_int32 * MAX_IO),
        ("output_type", C.c_int32 * MAX_IO),
        ("n_blocks", C.c_int32),
        ("n_registers", C.c_int32),
        ("index_bytes", C.c_int32),
        ("state_size", C.c_uint32),
        ("queue_row_bytes", C.c_uint32),
        ("block_ops", C.c_uint16 * MAX_BLOCKS),
        ("block_live_bytes", C.c_uint32 * MAX_BLOCKS),
        ("block_flops", C.c_uint16 * MAX_BLOCKS),
    ]


class Params(C.Structure):
    _fields_ = [
        ("threshold", C.c_uint32),
        ("live_threshold", C.c_uint32),
        ("refill", C.c_uint32),
        ("regroup", C.c_uint32),
        ("reconv", C.c_uint32),
        ("waves_per_cu", C.c_uint32),
        ("collect_stats", C.c_uint32),
        ("queue_capacity", C.c_uint64),
        ("max_steps", C.c_uint64),
        ("register_file", C.c_uint32),
        ("regroup_phase", C.c_uint32),
        ("regroup_waves", C.c_uint32),
    ]


class RunStats(C.Structure):
    _fields_ = [
        ("ms_total", C.c_double),
        ("ms_main", C.c_double),
        ("main_launches", C.c_uint32),
        ("passes", C.c_uint32),
        ("errors", C.c_uint32),
        ("grid_waves", C.c_uint32),
        ("counters", C.c_uint64 * NUM_COUNTERS),
    ]


MAX_JOBS = 64


class Job(C.Structure):
    _fields_ = [
        ("kernel", C.c_int32),
        ("pad", C.c_int32),
        ("begin", C.c_uint64),
        ("end", C.c_uint64),
        ("inputs", C.c_void_p * MAX_IO),
        ("outputs", C.c_void_p * MAX_IO),
    ]


class ServerConfig(C.Structure):
    _fields_ = [
        ("batch_set", C.c_int32),
        ("max_resident", C.c_uint32),
        ("hbm_budget", C.c_uint64),
        ("step_threads", C.c_uint64),
        ("params", Params),
    ]


class Request(C.Structure):
    _fields_ = [
        ("kernel", C.c_int32),
        ("broadcast", C.c_uint32),
        ("n", C.c_uint64),
        ("inputs", C.c_void_p * MAX_IO),
        ("outputs", C.c_void_p * MAX_IO),
        ("scalars", C.c_uint64 * MAX_IO),
        ("arrival_ms", C.c_double),
        ("admit_ms", C.c_double),
        ("done_ms", C.c_double),
        ("steps", C.c_uint32),
        ("status", C.c_int32),
        ("complete", C.c_int32),
        ("reserved", C.c_int32),
    ]


class StreamCounters(C.Structure):
    _fields_ = [
        ("steps", C.c_uint64),
        ("exchanges", C.c_uint64),
        ("rows_in", C.c_uint64),
        ("rows_out", C.c_uint64),
        ("bytes_h2d", C.c_uint64),
        ("bytes_d2h", C.c_uint64),
    ]


class ServerStats(C.Structure):
    _fields_ = [
        ("makespan_ms", C.c_double),
        ("gpu_ms", C.c_double),
        ("steps", C.c_uint64),
        ("threads", C.c_uint64),
        ("peak_resident", C.c_uint32),
        ("pad", C.c_uint32),
        ("peak_bytes", C.c_uint64),
    ]


# every symbol declared in include/loom_amd.h, with its signature
SIGNATURES = {
    "loom_init": (C.c_int, [C.c_int]),
    "loom_last_error": (C.c_char_p, []),
    "loom_device_info": (C.c_int, [C.POINTER(C.c_int), C.POINTER(C.c_int),
                                   C.POINTER(C.c_size_t), C.c_char_p, C.c_size_t]),
    "loom_kernel_count": (C.c_int, []),
    "loom_kernel_info_get": (C.c_int, [C.c_int, C.POINTER(KernelInfo)]),
    "loom_kernel_find": (C.c_int, [C.c_char_p]),
    "loom_kernel_dump": (C.c_int, [C.c_int, C.c_char_p, C.c_size_t, C.POINTER(C.c_size_t)]),
    "loom_kernel_resources": (C.c_int, [C.c_int, C.c_int, C.c_int, C.POINTER(C.c_int),
                                        C.POINTER(C.c_int)]),
    "loom_kernel_register": (C.c_int, [C.c_int, C.c_int, C.c_char_p, C.c_size_t,
                                       C.POINTER(C.c_int), C.POINTER(C.c_int),
                                       C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    "loom_kernel_trace": (C.c_int, [C.c_int, C.POINTER(C.c_void_p), C.c_uint64, C.c_uint32,
                                    C.c_void_p, C.c_size_t, C.c_void_p]),
    "loom_malloc": (C.c_int, [C.POINTER(C.c_void_p), C.c_size_t]),
    "loom_free": (C.c_int, [C.c_void_p]),
    "loom_memcpy_h2d": (C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t]),
    "loom_memcpy_d2h": (C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t]),
    "loom_memcpy_d2d": (C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t]),
    "loom_memset": (C.c_int, [C.c_void_p, C.c_int, C.c_size_t]),
    "loom_synchronize": (C.c_int, []),
    "loom_fill": (C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t, C.c_uint64]),
    "loom_host_alloc": (C.c_int, [C.POINTER(C.c_void_p), C.c_size_t]),
    "loom_copy_async": (C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int,
                                  C.POINTER(C.c_uint64)]),
    "loom_copy_status": (C.c_int, [C.c_uint64, C.POINTER(C.c_int)]),
    "loom_copy_wait": (C.c_int, [C.c_uint64]),
    "loom_host_free": (C.c_int, [C.c_void_p]),
    "loom_run": (C.c_int, [C.c_int, C.c_int, C.POINTER(C.c_void_p), C.POINTER(C.c_void_p),
                           C.c_uint64, C.POINTER(Params), C.POINTER(RunStats)]),
    "loom_run_noalloc": (C.c_int, [C.c_int, C.c_int, C.POINTER(C.c_void_p),
                                   C.POINTER(C.c_void_p), C.c_uint64, C.POINTER(Params),
                                   C.POINTER(RunStats)]),
    "loom_run_host": (C.c_int, [C.c_int, C.c_int, C.POINTER(C.c_void_p),
                                C.POINTER(C.c_void_p), C.c_uint64, C.POINTER(Params),
                                C.POINTER(RunStats)]),
    "loom_batch_set_count": (C.c_int, []),
    "loom_batch_set_find": (C.c_int, [C.c_char_p]),
    "loom_batch_set_members": (C.c_int, [C.c_int, C.POINTER(C.c_int), C.c_int]),
    "loom_batch_run": (C.c_int, [C.c_int, C.POINTER(Job), C.c_int, C.POINTER(Params),
                                 C.POINTER(RunStats)]),
    "loom_batch_graph_create": (C.c_int, [C.c_int, C.POINTER(Job), C.c_int, C.POINTER(Params),
                                          C.POINTER(C.c_void_p)]),
    "loom_graph_launch": (C.c_int, [C.c_void_p, C.POINTER(RunStats)]),
    "loom_graph_destroy": (C.c_int, [C.c_void_p]),
    "loom_server_create": (C.c_int, [C.POINTER(ServerConfig), C.POINTER(C.c_void_p)]),
    "loom_server_run": (C.c_int, [C.c_void_p, C.POINTER(Request), C.c_uint64, C.c_int,
                                  C.POINTER(ServerStats)]),
    "loom_server_start": (C.c_int, [C.c_void_p]),
    "loom_server_submit": (C.c_int, [C.c_void_p, C.POINTER(Request)]),
    "loom_server_poll": (C.c_int, [C.c_void_p, C.POINTER(Request), C.POINTER(C.c_int)]),
    "loom_server_wait": (C.c_int, [C.c_void_p, C.POINTER(Request)]),
    "loom_server_stop": (C.c_int, [C.c_void_p, C.POINTER(ServerStats)]),
    "loom_server_destroy": (C.c_int, [C.c_void_p]),
    "loom_context_create": (C.c_int, [C.POINTER(C.c_void_p)]),
    "loom_context_destroy": (C.c_int, [C.c_void_p]),
    "loom_ctx_run": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.POINTER(C.c_void_p),
                               C.POINTER(C.c_void_p), C.c_uint64, C.POINTER(Params),
                               C.POINTER(RunStats)]),
    "loom_ctx_run_host": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.POINTER(C.c_void_p),
                                    C.POINTER(C.c_void_p), C.c_uint64, C.POINTER(Params),
                                    C.POINTER(RunStats)]),
    "loom_ctx_fill": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t, C.c_uint64]),
    "loom_ctx_batch_run": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(Job), C.c_int,
                                     C.POINTER(Params), C.POINTER(RunStats)]),
    "loom_ctx_copy_async": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t, C.c_int,
                                      C.POINTER(C.c_uint64)]),
    "loom_ctx_copy_status": (C.c_int, [C.c_void_p, C.c_uint64, C.POINTER(C.c_int)]),
    "loom_ctx_copy_wait": (C.c_int, [C.c_void_p, C.c_uint64]),
    "loom_ctx_synchronize": (C.c_int, [C.c_void_p]),
    "loom_stream_create": (C.c_int, [C.c_int, C.c_uint64, C.POINTER(Params),
                                     C.POINTER(C.c_void_p)]),
    "loom_stream_destroy": (C.c_int, [C.c_void_p]),
    "loom_stream_set_inputs": (C.c_int, [C.c_void_p, C.POINTER(C.c_void_p), C.c_uint64,
                                         C.POINTER(C.c_uint64)]),
    "loom_stream_step": (C.c_int, [C.c_void_p, C.POINTER(RunStats)]),
    "loom_stream_sort": (C.c_int, [C.c_void_p, C.POINTER(C.c_int)]),
    "loom_stream_exchange": (C.c_int, [C.c_void_p, C.POINTER(C.c_void_p), C.c_uint64,
                                       C.POINTER(C.c_void_p), C.c_void_p,
                                       C.POINTER(C.c_uint64)]),
    "loom_stream_drain": (C.c_int, [C.c_void_p, C.POINTER(C.c_void_p), C.c_void_p, C.c_uint64,
                                    C.POINTER(C.c_uint64)]),
    "loom_stream_state": (C.c_int, [C.c_void_p, C.POINTER(C.c_uint64),
                                    C.POINTER(C.c_uint64)]),
    "loom_stream_counters_get": (C.c_int, [C.c_void_p, C.POINTER(StreamCounters)]),
}

_lib = None


def library_path() -> Path:
    return _LIB_PATH


def load() -> C.CDLL:
    """Load the kernel library (no GPU call is made here)."""
    global _lib
    if _lib is None:
        if not _LIB_PATH.exists():
            raise RuntimeError(
                f"{_LIB_PATH} is missing: build it with `python -m loom_amd.build`")
        lib = C.CDLL(str(_LIB_PATH))
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
    return _lib


class LoomError(RuntimeError):
    def __init__(self, code: int, where: str):
        msg = load().loom_last_error()
        super().__init__(f"{where} failed ({code}): {msg.decode() if msg else ''}")
        self.code = code


def check(code: int, where: str) -> None:
    if code != 0:
        raise LoomError(code, where)
