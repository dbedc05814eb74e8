"""ctypes binding of libwahgpu (include/wahgpu.h).

The library is built in-tree (indexing_amd/build/libwahgpu.so) by
``__graft_entry__.build()`` (make in indexing_amd/csrc).  Loading fails loudly when it is missing:
there is no fallback implementation.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
BUILD_DIR = os.path.join(HERE, "build")
LIB_PATH = os.path.join(BUILD_DIR, "libwahgpu.so")

OK = 0
EINVAL, ENOMEM, EHIP, ERANGE, ENODEV, EIO, EBUSY = -1, -2, -3, -4, -5, -6, -7
MAX_VALUES = (1 << 30) - 1
NOALLOC_MAX_VALUES = 1 << 22
TIMING_GROUPS = 7

DIST_UNIFORM, DIST_ZIPF, DIST_ZIPF_RUNS = 0, 1, 2
OP_AND, OP_OR, OP_XOR, OP_ANDNOT = 0, 1, 2, 3

c_u32p = ctypes.POINTER(ctypes.c_uint32)
c_u64p = ctypes.POINTER(ctypes.c_uint64)
vp = ctypes.c_void_p


class Counts(ctypes.Structure):
    _fields_ = [("k", ctypes.c_uint64), ("length", ctypes.c_uint64),
                ("keycnt", ctypes.c_uint64), ("bound", ctypes.c_uint32),
                ("passes", ctypes.c_uint32), ("fault", ctypes.c_uint32),
                ("mode", ctypes.c_uint32)]


class PhaseTimes(ctypes.Structure):
    _fields_ = [(name, ctypes.c_double) for name in (
        "transfer_to", "sort", "chunks", "merge", "fills", "fuse", "column_length",
        "transfer_back", "total")]


class GenParams(ctypes.Structure):
    _fields_ = [("dist", ctypes.c_uint32), ("from_", ctypes.c_uint32),
                ("to", ctypes.c_uint32), ("cardinality", ctypes.c_uint32),
                ("zipf_s", ctypes.c_double), ("run_mean", ctypes.c_double),
                ("seed", ctypes.c_uint64)]


class Column(ctypes.Structure):
    _fields_ = [("offset", ctypes.c_uint64), ("n", ctypes.c_uint64),
                ("rid_base", ctypes.c_uint32), ("reserved", ctypes.c_uint32)]


class Result(ctypes.Structure):
    _fields_ = [("keys", vp), ("offsets", vp), ("lengths", vp), ("keys_cap", ctypes.c_uint64),
                ("words", vp), ("words_cap", ctypes.c_uint64), ("counts", Counts)]


class ServerConfig(ctypes.Structure):
    _fields_ = [("device", ctypes.c_int), ("slots", ctypes.c_uint32),
                ("arena_bytes", ctypes.c_uint64), ("max_batch_cols", ctypes.c_uint32),
                ("max_batch_values", ctypes.c_uint64), ("plain", ctypes.c_uint32),
                ("reserved", ctypes.c_uint32)]


class ServerStats(ctypes.Structure):
    _fields_ = [("requests", ctypes.c_uint64), ("batches", ctypes.c_uint64),
                ("values", ctypes.c_uint64), ("max_batch_cols", ctypes.c_uint64),
                ("slot_values", ctypes.c_uint64), ("p50_us", ctypes.c_double),
                ("p99_us", ctypes.c_double), ("mean_us", ctypes.c_double),
                ("max_us", ctypes.c_double), ("busy_s", ctypes.c_double),
                ("memory_waits", ctypes.c_uint64), ("arena_bytes", ctypes.c_uint64),
                ("arena_peak", ctypes.c_uint64)]


TUNE_BUCKETS, TUNE_SHAPES = 5, 6


class Tuning(ctypes.Structure):
    _fields_ = [("bucket_log2", ctypes.c_uint32 * TUNE_BUCKETS),
                ("shape", ctypes.c_uint32 * TUNE_BUCKETS),
                ("digit_bits", ctypes.c_uint32 * TUNE_BUCKETS),
                ("ms", (ctypes.c_double * TUNE_SHAPES) * TUNE_BUCKETS),
                ("shape_threads", ctypes.c_uint32 * TUNE_SHAPES),
                ("shape_items", ctypes.c_uint32 * TUNE_SHAPES),
                ("digit_ms", (ctypes.c_double * 2) * TUNE_BUCKETS)]


# name -> (restype, argtypes); every function the header declares.
SIGNATURES = {
    "wahgpu_device_count": (ctypes.c_int, [ctypes.POINTER(ctypes.c_int)]),
    "wahgpu_device_name": (ctypes.c_int, [ctypes.c_int, ctypes.c_char_p, ctypes.c_size_t]),
    "wahgpu_device_compute_units": (ctypes.c_int, [ctypes.c_int]),
    "wahgpu_find_device": (ctypes.c_int, [ctypes.c_char_p]),
    "wahgpu_create": (ctypes.c_int, [ctypes.POINTER(vp), ctypes.c_int, ctypes.c_uint64]),
    "wahgpu_destroy": (None, [vp]),
    "wahgpu_last_error": (ctypes.c_char_p, [vp]),
    "wahgpu_stream": (vp, [vp]),
    "wahgpu_set_stream": (ctypes.c_int, [vp, vp]),
    "wahgpu_synchronize": (ctypes.c_int, [vp]),
    "wahgpu_arena_bytes": (ctypes.c_uint64, [vp]),
    "wahgpu_rank_mode": (ctypes.c_int, [vp]),
    "wahgpu_build_device": (ctypes.c_int, [vp, vp, ctypes.c_uint64, ctypes.c_uint32, vp, vp,
                                           vp, vp, vp]),
    "wahgpu_build_device_noalloc": (ctypes.c_int, [vp, vp, ctypes.c_uint64, ctypes.c_uint32,
                                                   vp, vp, vp, vp, vp]),
    "wahgpu_timing_begin": (ctypes.c_int, [vp, ctypes.c_uint32]),
    "wahgpu_timing_end": (ctypes.c_int, [vp, ctypes.POINTER(ctypes.c_double),
                                         ctypes.POINTER(ctypes.c_uint32)]),
    "wahgpu_build": (ctypes.c_int, [vp, vp, ctypes.c_uint64, ctypes.c_uint32, vp, vp, vp,
                                    ctypes.c_uint64, vp, ctypes.c_uint64,
                                    ctypes.POINTER(Counts), ctypes.POINTER(PhaseTimes)]),
    "wahgpu_build_staged": (ctypes.c_int, [vp, vp, ctypes.c_uint64, vp, vp, vp,
                                           ctypes.c_uint64, vp, ctypes.c_uint64,
                                           ctypes.POINTER(Counts),
                                           ctypes.POINTER(PhaseTimes)]),
    "wahgpu_build_batch_device": (ctypes.c_int, [vp, vp, ctypes.c_uint64,
                                                 ctypes.POINTER(Column), ctypes.c_uint32,
                                                 vp, vp, vp, vp, vp]),
    "wahgpu_build_batch": (ctypes.c_int, [vp, ctypes.POINTER(vp), c_u64p, c_u32p,
                                          ctypes.c_uint32, ctypes.POINTER(Result)]),
    "wahgpu_batch_reserve": (ctypes.c_int, [vp, ctypes.c_uint64, ctypes.c_uint32]),
    "wahgpu_graph_stats": (ctypes.c_int, [vp, c_u64p, c_u64p]),
    "wahgpu_autotune": (ctypes.c_int, [vp, ctypes.POINTER(Tuning)]),
    "wahgpu_get_tuning": (ctypes.c_int, [vp, ctypes.POINTER(Tuning)]),
    "wahgpu_set_tuning": (ctypes.c_int, [vp, ctypes.POINTER(Tuning)]),
    "wahgpu_server_create": (ctypes.c_int, [ctypes.POINTER(vp), ctypes.POINTER(ServerConfig)]),
    "wahgpu_server_destroy": (None, [vp]),
    "wahgpu_server_submit": (ctypes.c_int, [vp, vp, ctypes.c_uint64, ctypes.c_uint32,
                                            ctypes.POINTER(Result), c_u64p]),
    "wahgpu_server_poll": (ctypes.c_int, [vp, ctypes.c_uint64]),
    "wahgpu_server_wait": (ctypes.c_int, [vp, ctypes.c_uint64]),
    "wahgpu_server_drain": (ctypes.c_int, [vp]),
    "wahgpu_server_last_error": (ctypes.c_char_p, [vp]),
    "wahgpu_server_get_stats": (ctypes.c_int, [vp, ctypes.POINTER(ServerStats)]),
    "wahgpu_server_reset_stats": (ctypes.c_int, [vp]),
    "wahgpu_host_alloc": (ctypes.c_int, [ctypes.c_uint64, ctypes.POINTER(vp)]),
    "wahgpu_host_free": (None, [vp]),
    "wahgpu_sort_rids_by_value": (ctypes.c_int, [vp, vp, ctypes.c_uint64, ctypes.c_uint32,
                                                 vp, vp]),
    "wahgpu_produce_chunk_id_literals": (ctypes.c_int, [vp, vp, ctypes.c_uint64, vp, vp]),
    "wahgpu_merge_lit_by_val_chids": (ctypes.c_int, [vp, vp, vp, vp, ctypes.c_uint64, vp, vp,
                                                     vp, c_u64p]),
    "wahgpu_produce_fills": (ctypes.c_int, [vp, vp, vp, ctypes.c_uint64, vp]),
    "wahgpu_fuse_fill_literals": (ctypes.c_int, [vp, vp, vp, ctypes.c_uint64, vp, c_u64p]),
    "wahgpu_compute_column_length": (ctypes.c_int, [vp, vp, vp, ctypes.c_uint64, vp, vp, vp,
                                                    c_u64p]),
    "wahgpu_stream_compaction": (ctypes.c_int, [vp, vp, vp, ctypes.c_uint64, vp, c_u64p]),
    "wahgpu_exclusive_scan": (ctypes.c_int, [vp, vp, ctypes.c_uint64, vp]),
    "wahgpu_segmented_inclusive_scan": (ctypes.c_int, [vp, vp, vp, ctypes.c_uint64, vp]),
    "wahgpu_merger_create": (ctypes.c_int, [ctypes.POINTER(vp)]),
    "wahgpu_merger_destroy": (None, [vp]),
    "wahgpu_merger_add": (ctypes.c_int, [vp, ctypes.c_uint64, ctypes.c_uint64, vp, vp, vp]),
    "wahgpu_merger_keycnt": (ctypes.c_uint64, [vp]),
    "wahgpu_merger_length": (ctypes.c_uint64, [vp]),
    "wahgpu_merger_export": (ctypes.c_int, [vp, vp, vp, vp, vp]),
    "wahgpu_merger_key_words": (ctypes.c_int, [vp, ctypes.c_uint32, ctypes.POINTER(c_u32p),
                                               c_u64p]),
    "wahgpu_decode_device": (ctypes.c_int, [vp, vp, ctypes.c_uint64, ctypes.c_uint64, vp,
                                            c_u64p]),
    "wahgpu_query_range_device": (ctypes.c_int, [vp, vp, vp, vp, ctypes.c_uint64, vp,
                                                 ctypes.c_uint32, ctypes.c_uint32,
                                                 ctypes.c_uint64, vp, c_u64p]),
    "wahgpu_encode_device": (ctypes.c_int, [vp, vp, ctypes.c_uint64, vp, ctypes.c_uint64,
                                            c_u64p]),
    "wahgpu_combine_device": (ctypes.c_int, [vp, ctypes.c_int, vp, ctypes.c_uint64, vp,
                                             ctypes.c_uint64, ctypes.c_uint64, vp,
                                             ctypes.c_uint64, c_u64p]),
    "wahgpu_read_values_text": (ctypes.c_int, [ctypes.c_char_p, ctypes.POINTER(c_u32p),
                                               c_u64p]),
    "wahgpu_read_values_text_threads": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_int,
                                                       ctypes.POINTER(c_u32p), c_u64p]),
    "wahgpu_write_values_text": (ctypes.c_int, [ctypes.c_char_p, vp, ctypes.c_uint64,
                                                ctypes.c_char_p, ctypes.c_int]),
    "wahgpu_read_values_binary": (ctypes.c_int, [ctypes.c_char_p, ctypes.POINTER(c_u32p),
                                                 c_u64p]),
    "wahgpu_free": (None, [vp]),
    "wahgpu_index_checksum": (ctypes.c_uint64, [vp, vp, vp, ctypes.c_uint64, vp,
                                                ctypes.c_uint64]),
    "wahgpu_decode_bitmap": (ctypes.c_uint64, [vp, ctypes.c_uint64, ctypes.c_char_p,
                                               ctypes.c_uint64]),
    "wahgpu_generate": (ctypes.c_int, [ctypes.POINTER(GenParams), vp, ctypes.c_uint64]),
    "wahgpu_generate_range": (ctypes.c_int, [ctypes.POINTER(GenParams), vp, ctypes.c_uint64,
                                             ctypes.c_uint64]),
}

_lib = None


def lib():
    """The loaded libwahgpu; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"libwahgpu is not built ({LIB_PATH} missing); run "
                "`python -c 'import __graft_entry__ as g; g.build()'`")
        # torch ships its own HIP runtime and asks for it as "libamdhip64.so";
        # loading torch first makes this library bind to that same runtime
        # (one HIP runtime per process) instead of a second copy.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


class WahError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"libwahgpu error {code}: {msg}")
        self.code = code
