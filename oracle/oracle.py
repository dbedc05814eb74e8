"""ctypes binding of the CPU oracle (oracle/wah_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, by the smoke check and by the
benchmark's CPU-baseline leg.  Product code never imports it.  Build with
``make -C oracle``.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "build", "liboracle_wah.so")
CPU_BIN = os.path.join(HERE, "build", "cpu")

SORT_BUBBLE, SORT_PACKED = 0, 1


class Counts(ctypes.Structure):
    _fields_ = [("k", ctypes.c_uint64), ("length", ctypes.c_uint64),
                ("keycnt", ctypes.c_uint64)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"oracle not built: {LIB_PATH} (run make -C oracle)")
        L = ctypes.CDLL(LIB_PATH)
        vp, sz, u32 = ctypes.c_void_p, ctypes.c_size_t, ctypes.c_uint32
        sig = {
            "wah_oracle_sort_rids_by_value": (None, [vp, vp, sz, u32, ctypes.c_int]),
            "wah_oracle_produce_chunk_id_literals": (None, [vp, vp, vp, sz]),
            "wah_oracle_merge_lit_by_val_chids": (sz, [vp, vp, vp, sz]),
            "wah_oracle_produce_fills": (None, [vp, vp, sz]),
            "wah_oracle_fuse_fill_literals": (sz, [vp, vp, vp, sz]),
            "wah_oracle_compute_column_length": (sz, [vp, vp, vp, vp, sz]),
            "wah_oracle_stream_compaction": (sz, [vp, sz, u32]),
            "wah_oracle_compact_valid": (sz, [vp, vp, sz, vp]),
            "wah_oracle_exclusive_scan": (None, [vp, vp, sz]),
            "wah_oracle_segmented_inclusive_scan": (None, [vp, vp, vp, sz]),
            "wah_oracle_build": (ctypes.c_int, [vp, sz, u32, ctypes.c_int, vp, vp, vp, vp,
                                                ctypes.POINTER(Counts)]),
            "wah_oracle_build_timed": (ctypes.c_int, [vp, sz, u32, ctypes.c_int, vp, vp, vp,
                                                      vp, ctypes.POINTER(Counts),
                                                      ctypes.POINTER(ctypes.c_double)]),
            "wah_oracle_build_threads": (ctypes.c_int, [vp, sz, u32, ctypes.c_int, vp, vp, vp,
                                                        vp, ctypes.POINTER(Counts),
                                                        ctypes.POINTER(ctypes.c_double)]),
            "wah_oracle_decoded_bitmap": (sz, [vp, sz, ctypes.c_char_p, sz]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype, fn.argtypes = res, args
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _u32(a):
    return np.ascontiguousarray(a, dtype=np.uint32).copy()


STAGES = ("sort", "chunks", "merge", "fills", "fuse", "column_length")


def build(values, rid_base=0, sort="packed", timed=False, threads=1):
    """Whole pipeline of src/cpu.cpp:313-328 -> dict of numpy arrays + counts
    (with timed=True also "stage_s": seconds per stage, STAGES order).
    threads > 1 sorts on that many threads (packed order; CPU baseline)."""
    v = _u32(values)
    n = v.size
    cap = max(n, 1)
    keys, offs, lens = (np.zeros(cap, np.uint32) for _ in range(3))
    words = np.zeros(2 * cap, np.uint32)
    c = Counts()
    mode = SORT_BUBBLE if sort == "bubble" else SORT_PACKED
    st = (ctypes.c_double * len(STAGES))()
    if threads > 1 and sort != "bubble":
        rc = lib().wah_oracle_build_threads(_p(v), n, rid_base, threads, _p(keys), _p(offs),
                                            _p(lens), _p(words), ctypes.byref(c), st)
    else:
        rc = lib().wah_oracle_build_timed(_p(v), n, rid_base, mode, _p(keys), _p(offs),
                                          _p(lens), _p(words), ctypes.byref(c), st)
    if rc != 0:
        raise MemoryError("oracle build failed")
    K, L = int(c.keycnt), int(c.length)
    out = {"keys": keys[:K], "offsets": offs[:K], "lengths": lens[:K], "words": words[:L],
           "k": int(c.k), "length": L, "keycnt": K}
    if timed:
        out["stage_s"] = dict(zip(STAGES, list(st)))
    return out


def sort_rids_by_value(values, rid_base=0, sort="packed"):
    v = _u32(values)
    r = np.zeros(v.size, np.uint32)
    lib().wah_oracle_sort_rids_by_value(_p(v), _p(r), v.size, rid_base,
                                        SORT_BUBBLE if sort == "bubble" else SORT_PACKED)
    return v, r


def produce_chunk_id_literals(rids):
    r = _u32(rids)
    ch, li = np.zeros(r.size, np.uint32), np.zeros(r.size, np.uint32)
    lib().wah_oracle_produce_chunk_id_literals(_p(r), _p(ch), _p(li), r.size)
    return ch, li


def merge_lit_by_val_chids(values, chids, lits):
    v, c, l = _u32(values), _u32(chids), _u32(lits)
    k = lib().wah_oracle_merge_lit_by_val_chids(_p(v), _p(c), _p(l), v.size)
    return v[:k], c[:k], l[:k]


def produce_fills(values, chids):
    v, c = _u32(values), _u32(chids)
    lib().wah_oracle_produce_fills(_p(v), _p(c), v.size)
    return c


def fuse_fill_literals(fills, lits):
    f, l = _u32(fills), _u32(lits)
    idx = np.zeros(max(2 * f.size, 1), np.uint32)
    L = lib().wah_oracle_fuse_fill_literals(_p(f), _p(l), _p(idx), f.size)
    return idx[:L]


def compute_column_length(values, fills):
    v, f = _u32(values), _u32(fills)
    offs, lens = np.zeros(max(v.size, 1), np.uint32), np.zeros(max(v.size, 1), np.uint32)
    K = lib().wah_oracle_compute_column_length(_p(v), _p(f), _p(offs), _p(lens), v.size)
    return v[:K], offs[:K], lens[:K]


def stream_compaction(data, valid):
    d, m = _u32(data), _u32(valid)
    out = np.zeros(max(d.size, 1), np.uint32)
    cnt = lib().wah_oracle_compact_valid(_p(d), _p(m), d.size, _p(out))
    return out[:cnt]


def exclusive_scan(x):
    a = _u32(x)
    out = np.zeros(a.size, np.uint32)
    lib().wah_oracle_exclusive_scan(_p(a), _p(out), a.size)
    return out


def segmented_inclusive_scan(x, heads):
    a, h = _u32(x), _u32(heads)
    out = np.zeros(a.size, np.uint32)
    lib().wah_oracle_segmented_inclusive_scan(_p(a), _p(h), _p(out), a.size)
    return out


def decoded_bitmap(words):
    w = _u32(words)
    need = lib().wah_oracle_decoded_bitmap(_p(w), w.size, None, 0)
    buf = ctypes.create_string_buffer(max(int(need), 1))
    lib().wah_oracle_decoded_bitmap(_p(w), w.size, buf, need)
    return buf.raw[:need].decode()
