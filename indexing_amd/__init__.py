"""indexing_amd -- MI355X-native WAH bitmap-index construction.

Python face of libwahgpu (include/wahgpu.h), mirroring the stages of the
reference pipeline (src/cpu.cpp:176-272, src/phases.cpp:720-1088):

    sort_rids_by_value -> produce_chunk_id_literals -> merge_lit_by_val_chids
    -> produce_fills -> fuse_fill_literals -> compute_column_length

Stage functions take and return torch CUDA tensors (int32 storage holding
uint32 bit patterns) and run on torch's current stream.  ``Context.build``
runs the fused build on a host array or a device tensor.  Every call goes to
the HIP kernels; without the built library the import of a GPU entry point
raises (there is no CPU fallback).
"""
import ctypes
from dataclasses import dataclass, field

import numpy as np

from . import _native as N
from ._native import WahError, MAX_VALUES

__all__ = [
    "Context", "Server", "Index", "pinned_empty", "WahError", "MAX_VALUES", "Merger", "read_values_text",
    "write_values_text", "index_checksum",
    "read_values_binary", "decode_bitmap", "generate", "device_count",
]


def _torch():
    import torch
    return torch


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None and t.numel() else ctypes.c_void_p(0)


def _as_u32(a):
    return np.ascontiguousarray(a, dtype=np.uint32)


@dataclass
class Index:
    """A built index: words[offsets[i] : offsets[i] + lengths[i]] encode keys[i]."""
    keys: np.ndarray
    offsets: np.ndarray
    lengths: np.ndarray
    words: np.ndarray
    k: int = 0
    length: int = 0
    keycnt: int = 0
    bound: int = 0
    passes: int = 0
    times: dict = field(default_factory=dict)
    mode: int = 0  # 0: one (value, rid) pair per row sorted; 1: one item per run

    def key_words(self, i):
        o = int(self.offsets[i])
        return self.words[o:o + int(self.lengths[i])]


def device_count():
    c = ctypes.c_int(0)
    N.lib().wahgpu_device_count(ctypes.byref(c))
    return c.value


class Context:
    """One HIP stream plus a pooled device arena (wahgpu_create)."""

    def __init__(self, device=0, arena_bytes=0, use_torch_stream=True):
        self._lib = N.lib()
        h = ctypes.c_void_p()
        rc = self._lib.wahgpu_create(ctypes.byref(h), int(device), int(arena_bytes))
        if rc != N.OK:
            raise WahError(rc, f"wahgpu_create(device={device}) failed")
        self._h = h
        self.device = device
        self.use_torch_stream = use_torch_stream

    def close(self):
        if getattr(self, "_h", None):
            self._lib.wahgpu_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # -- plumbing ----------------------------------------------------------
    def _check(self, rc):
        if rc != N.OK:
            msg = self._lib.wahgpu_last_error(self._h)
            raise WahError(rc, msg.decode() if msg else "")

    def _bind_stream(self):
        """Order library work on torch's current stream."""
        if self.use_torch_stream:
            torch = _torch()
            s = torch.cuda.current_stream(self.device).cuda_stream
            self._check(self._lib.wahgpu_set_stream(self._h, ctypes.c_void_p(s)))

    @property
    def stream(self):
        return self._lib.wahgpu_stream(self._h)

    def synchronize(self):
        self._check(self._lib.wahgpu_synchronize(self._h))

    def timing_begin(self, max_calls):
        """Record per-kernel HIP events for the next build_device calls."""
        self._check(self._lib.wahgpu_timing_begin(self._h, int(max_calls)))

    def timing_end(self):
        """-> (per-group summed ms, calls); groups: histogram, pass0..pass3,
        tail, lengths."""
        ms = (ctypes.c_double * N.TIMING_GROUPS)()
        calls = ctypes.c_uint32(0)
        self._check(self._lib.wahgpu_timing_end(self._h, ms, ctypes.byref(calls)))
        return list(ms), calls.value

    def arena_bytes(self):
        return int(self._lib.wahgpu_arena_bytes(self._h))

    @property
    def rank_mode(self):
        """1: radix ranking by LDS fetch-add (device-checked), 0: ballot match."""
        return int(self._lib.wahgpu_rank_mode(self._h))

    def _empty(self, n):
        torch = _torch()
        return torch.empty(max(int(n), 0), dtype=torch.int32, device=f"cuda:{self.device}")

    # -- whole build ---------------------------------------------------------
    def build(self, values, rid_base=0, staged=False, timed=False):
        """Build the index of one column given as a host array (uint32)."""
        v = _as_u32(values)
        n = v.size
        cap = max(n, 1)
        keys = np.empty(cap, np.uint32)
        offs = np.empty(cap, np.uint32)
        lens = np.empty(cap, np.uint32)
        words = np.empty(2 * cap, np.uint32)
        c = N.Counts()
        t = N.PhaseTimes()
        vp = v.ctypes.data_as(ctypes.c_void_p)
        args = (keys.ctypes.data_as(ctypes.c_void_p), offs.ctypes.data_as(ctypes.c_void_p),
                lens.ctypes.data_as(ctypes.c_void_p), cap,
                words.ctypes.data_as(ctypes.c_void_p), 2 * cap, ctypes.byref(c),
                ctypes.byref(t) if timed else None)
        if staged:
            if rid_base:
                raise ValueError("the staged build indexes rows from 0")
            self._check(self._lib.wahgpu_build_staged(self._h, vp, n, *args))
        else:
            self._check(self._lib.wahgpu_build(self._h, vp, n, int(rid_base), *args))
        K, L = int(c.keycnt), int(c.length)
        times = {name: getattr(t, name) for name, _ in N.PhaseTimes._fields_} if timed else {}
        return Index(keys[:K].copy(), offs[:K].copy(), lens[:K].copy(), words[:L].copy(),
                     int(c.k), L, K, int(c.bound), int(c.passes), times, int(c.mode))

    def build_device(self, d_values, rid_base=0, out=None):
        """Fused build on a device tensor; returns (keys, offsets, lengths, words,
        counts) as device tensors sized to the worst case (n, n, n, 2n) plus a
        device Counts buffer.  Asynchronous on torch's current stream."""
        torch = _torch()
        n = int(d_values.numel())
        if out is None:
            cap = max(n, 1)
            out = (self._empty(cap), self._empty(cap), self._empty(cap), self._empty(2 * cap),
                   torch.zeros(ctypes.sizeof(N.Counts) // 4, dtype=torch.int32,
                               device=d_values.device))
        keys, offs, lens, words, counts = out
        self._bind_stream()
        self._check(self._lib.wahgpu_build_device(
            self._h, _ptr(d_values), n, int(rid_base), _ptr(keys), _ptr(offs), _ptr(lens),
            _ptr(words), _ptr(counts)))
        return out

    def build_device_noalloc(self, d_values, rid_base=0, out=None):
        """build_device without any device allocation by the library: the
        scratch is a static 112 MiB device array per GPU
        (wahgpu_build_device_noalloc; columns of at most NOALLOC_MAX_VALUES).
        out, if given, is reused as in build_device.  Calls from other
        streams are ordered by the library; graphs that captured this call
        and are replayed on several streams at once race."""
        torch = _torch()
        n = int(d_values.numel())
        if out is None:
            cap = max(n, 1)
            out = (self._empty(cap), self._empty(cap), self._empty(cap), self._empty(2 * cap),
                   torch.zeros(ctypes.sizeof(N.Counts) // 4, dtype=torch.int32,
                               device=d_values.device))
        keys, offs, lens, words, counts = out
        self._bind_stream()
        self._check(self._lib.wahgpu_build_device_noalloc(
            self._h, _ptr(d_values), n, int(rid_base), _ptr(keys), _ptr(offs), _ptr(lens),
            _ptr(words), _ptr(counts)))
        return out

    @staticmethod
    def read_counts(counts_tensor):
        raw = counts_tensor.cpu().numpy().view(np.uint8).tobytes()
        return N.Counts.from_buffer_copy(raw[:ctypes.sizeof(N.Counts)])

    @staticmethod
    def read_counts_batch(counts_tensor, ncols):
        raw = counts_tensor.cpu().numpy().view(np.uint8).tobytes()
        sz = ctypes.sizeof(N.Counts)
        return [N.Counts.from_buffer_copy(raw[i * sz:(i + 1) * sz]) for i in range(ncols)]

    # -- batched build --------------------------------------------------------
    def build_batch(self, columns, rid_bases=None):
        """Index many host columns (uint32 arrays) in one launch set; returns
        one Index per column (wahgpu_build_batch)."""
        cols = [_as_u32(c) for c in columns]
        m = len(cols)
        if m == 0:
            return []
        vals = (ctypes.c_void_p * m)(*[c.ctypes.data_as(ctypes.c_void_p) for c in cols])
        ns = (ctypes.c_uint64 * m)(*[c.size for c in cols])
        rb = (ctypes.c_uint32 * m)(*[int(r) for r in rid_bases]) if rid_bases is not None else None
        res = (N.Result * m)()
        bufs = []
        for i, c in enumerate(cols):
            cap = max(c.size, 1)
            b = (np.empty(cap, np.uint32), np.empty(cap, np.uint32), np.empty(cap, np.uint32),
                 np.empty(2 * cap, np.uint32))
            bufs.append(b)
            r = res[i]
            r.keys, r.offsets, r.lengths = (x.ctypes.data for x in b[:3])
            r.keys_cap = cap
            r.words = b[3].ctypes.data
            r.words_cap = 2 * cap
        self._check(self._lib.wahgpu_build_batch(
            self._h, ctypes.cast(vals, ctypes.POINTER(ctypes.c_void_p)), ns,
            ctypes.cast(rb, N.c_u32p) if rb is not None else None, m, res))
        out = []
        for i in range(m):
            c = res[i].counts
            K, L = int(c.keycnt), int(c.length)
            k, o, ln, w = bufs[i]
            out.append(Index(k[:K].copy(), o[:K].copy(), ln[:K].copy(), w[:L].copy(), int(c.k),
                             L, K, int(c.bound), int(c.passes), mode=int(c.mode)))
        return out

    def build_batch_device(self, d_values, columns, out=None):
        """Batched build on one device tensor holding every column.  columns is
        a list of (offset, n[, rid_base]) ascending and not overlapping.
        Returns (keys, offsets, lengths, words, counts): column c's keys sit at
        [offset, offset + n), its words at [2 offset, 2 offset + 2 n), its
        counts at counts[c] (use read_counts_batch).  Asynchronous."""
        torch = _torch()
        m = len(columns)
        arr = (N.Column * max(m, 1))()
        for i, col in enumerate(columns):
            arr[i].offset, arr[i].n = int(col[0]), int(col[1])
            arr[i].rid_base = int(col[2]) if len(col) > 2 else 0
            arr[i].reserved = 0
        total = int(d_values.numel())
        if out is None:
            cap = max(total, 1)
            out = (self._empty(cap), self._empty(cap), self._empty(cap), self._empty(2 * cap),
                   torch.zeros(max(m, 1) * ctypes.sizeof(N.Counts) // 4, dtype=torch.int32,
                               device=d_values.device))
        keys, offs, lens, words, counts = out
        self._bind_stream()
        self._check(self._lib.wahgpu_build_batch_device(
            self._h, _ptr(d_values), total, arr, m, _ptr(keys), _ptr(offs), _ptr(lens),
            _ptr(words), _ptr(counts)))
        return out

    def batch_reserve(self, max_values, max_cols):
        """Lay scratch out for batches up to these sizes and replay them from
        one captured hipGraph ((0, 0) turns it off)."""
        self._check(self._lib.wahgpu_batch_reserve(self._h, int(max_values), int(max_cols)))

    def autotune(self):
        """Time every radix tile shape per column-length bucket on this device
        and keep the fastest (wahgpu_autotune); returns tuning()."""
        t = N.Tuning()
        self._bind_stream()
        self._check(self._lib.wahgpu_autotune(self._h, ctypes.byref(t)))
        return self._tuning_dict(t)

    def tuning(self):
        t = N.Tuning()
        self._check(self._lib.wahgpu_get_tuning(self._h, ctypes.byref(t)))
        return self._tuning_dict(t)

    def set_tuning(self, shapes, digit_bits=None):
        """shapes: radix tile shape per bucket, an index into
        tuning()["shapes"] ((threads, keys per thread) of each shape);
        digit_bits: 8 or 11 (wide digits allowed) per bucket, default 8."""
        t = N.Tuning()
        for b, sh in enumerate(shapes):
            t.shape[b] = int(sh)
        for b, d in enumerate(digit_bits or [8] * N.TUNE_BUCKETS):
            t.digit_bits[b] = int(d)
        self._check(self._lib.wahgpu_set_tuning(self._h, ctypes.byref(t)))

    @staticmethod
    def _tuning_dict(t):
        return {"bucket_log2": list(t.bucket_log2), "shape": list(t.shape),
                "digit_bits": list(t.digit_bits),
                "ms": [[round(x, 4) for x in row] for row in t.ms],
                "shapes": [(int(a), int(b)) for a, b in zip(t.shape_threads, t.shape_items)],
                "digit_ms": [[round(x, 4) for x in row] for row in t.digit_ms]}

    def graph_stats(self):
        cap, rep = ctypes.c_uint64(0), ctypes.c_uint64(0)
        self._check(self._lib.wahgpu_graph_stats(self._h, ctypes.byref(cap), ctypes.byref(rep)))
        return int(cap.value), int(rep.value)

    # -- stages (device tensors) -------------------------------------------
    def sort_rids_by_value(self, values, rid_base=0):
        n = values.numel()
        keys, rids = self._empty(n), self._empty(n)
        self._bind_stream()
        self._check(self._lib.wahgpu_sort_rids_by_value(self._h, _ptr(values), n,
                                                        int(rid_base), _ptr(keys), _ptr(rids)))
        return keys, rids

    def produce_chunk_id_literals(self, rids):
        n = rids.numel()
        chids, lits = self._empty(n), self._empty(n)
        self._bind_stream()
        self._check(self._lib.wahgpu_produce_chunk_id_literals(self._h, _ptr(rids), n,
                                                               _ptr(chids), _ptr(lits)))
        return chids, lits

    def merge_lit_by_val_chids(self, values, chids, lits):
        n = values.numel()
        ov, oc, ol = self._empty(n), self._empty(n), self._empty(n)
        k = ctypes.c_uint64(0)
        self._bind_stream()
        self._check(self._lib.wahgpu_merge_lit_by_val_chids(
            self._h, _ptr(values), _ptr(chids), _ptr(lits), n, _ptr(ov), _ptr(oc), _ptr(ol),
            ctypes.byref(k)))
        k = k.value
        return ov[:k], oc[:k], ol[:k]

    def produce_fills(self, values, chids):
        k = values.numel()
        fills = self._empty(k)
        self._bind_stream()
        self._check(self._lib.wahgpu_produce_fills(self._h, _ptr(values), _ptr(chids), k,
                                                   _ptr(fills)))
        return fills

    def fuse_fill_literals(self, fills, lits):
        k = fills.numel()
        index = self._empty(2 * k)
        L = ctypes.c_uint64(0)
        self._bind_stream()
        self._check(self._lib.wahgpu_fuse_fill_literals(self._h, _ptr(fills), _ptr(lits), k,
                                                        _ptr(index), ctypes.byref(L)))
        return index[:L.value]

    def compute_column_length(self, values, fills):
        k = values.numel()
        keys, offs, lens = self._empty(k), self._empty(k), self._empty(k)
        K = ctypes.c_uint64(0)
        self._bind_stream()
        self._check(self._lib.wahgpu_compute_column_length(
            self._h, _ptr(values), _ptr(fills), k, _ptr(keys), _ptr(offs), _ptr(lens),
            ctypes.byref(K)))
        K = K.value
        return keys[:K], offs[:K], lens[:K]

    # -- query side (device tensors) ------------------------------------------
    def decode_device(self, words, nrows):
        """A key's WAH words -> (row bitmap tensor of (nrows + 31) // 32 words,
        one past its last set row)."""
        bm = self._empty((int(nrows) + 31) // 32)
        last = ctypes.c_uint64(0)
        self._bind_stream()
        self._check(self._lib.wahgpu_decode_device(self._h, _ptr(words), int(words.numel()),
                                                   int(nrows), _ptr(bm), ctypes.byref(last)))
        return bm, int(last.value)

    def query_range_device(self, keys, offsets, lengths, words, lo, hi, nrows):
        """OR of the row bitmaps of the keys in [lo, hi] of a device index ->
        (row bitmap tensor, keys matched)."""
        bm = self._empty((int(nrows) + 31) // 32)
        m = ctypes.c_uint64(0)
        self._bind_stream()
        self._check(self._lib.wahgpu_query_range_device(
            self._h, _ptr(keys), _ptr(offsets), _ptr(lengths), int(keys.numel()), _ptr(words),
            int(lo), int(hi), int(nrows), _ptr(bm), ctypes.byref(m)))
        return bm, int(m.value)

    def encode_device(self, bitmap, nrows, cap=None):
        """Row bitmap tensor -> canonical WAH words (a device tensor)."""
        cap = 2 * ((int(nrows) + 30) // 31) if cap is None else int(cap)
        out = self._empty(max(cap, 1))
        L = ctypes.c_uint64(0)
        self._bind_stream()
        self._check(self._lib.wahgpu_encode_device(self._h, _ptr(bitmap), int(nrows), _ptr(out),
                                                   cap, ctypes.byref(L)))
        return out[:L.value]

    def combine_device(self, op, a, b, nrows, cap=None):
        """op in {"and", "or", "xor", "andnot"} of two WAH word tensors over
        rows [0, nrows) -> canonical WAH words (a device tensor)."""
        code = {"and": N.OP_AND, "or": N.OP_OR, "xor": N.OP_XOR, "andnot": N.OP_ANDNOT}[op]
        cap = 2 * (int(a.numel()) + int(b.numel())) if cap is None else int(cap)
        out = self._empty(max(cap, 1))
        L = ctypes.c_uint64(0)
        self._bind_stream()
        self._check(self._lib.wahgpu_combine_device(
            self._h, code, _ptr(a), int(a.numel()), _ptr(b), int(b.numel()), int(nrows),
            _ptr(out), cap, ctypes.byref(L)))
        return out[:L.value]

    # -- primitives ----------------------------------------------------------
    def stream_compaction(self, data, valid):
        n = data.numel()
        out = self._empty(n)
        cnt = ctypes.c_uint64(0)
        self._bind_stream()
        self._check(self._lib.wahgpu_stream_compaction(self._h, _ptr(data), _ptr(valid), n,
                                                       _ptr(out), ctypes.byref(cnt)))
        return out[:cnt.value]

    def exclusive_scan(self, x):
        n = x.numel()
        out = self._empty(n)
        self._bind_stream()
        self._check(self._lib.wahgpu_exclusive_scan(self._h, _ptr(x), n, _ptr(out)))
        return out

    def segmented_inclusive_scan(self, x, heads):
        n = x.numel()
        out = self._empty(n)
        self._bind_stream()
        self._check(self._lib.wahgpu_segmented_inclusive_scan(self._h, _ptr(x), _ptr(heads), n,
                                                              _ptr(out)))
        return out


# -- host-side formats --------------------------------------------------------
def _read(fn, path):
    L = N.lib()
    p = N.c_u32p()
    n = ctypes.c_uint64(0)
    rc = fn(str(path).encode(), ctypes.byref(p), ctypes.byref(n))
    if rc != N.OK:
        raise WahError(rc, f"cannot read {path}")
    try:
        return np.ctypeslib.as_array(p, shape=(n.value,)).copy() if n.value else \
            np.empty(0, np.uint32)
    finally:
        L.wahgpu_free(p)


def read_values_text(path, threads=0):
    """Whitespace-separated decimal uint32s, read like `ifstream >> uint32_t`
    (parsed in parallel; threads=1 is one serial pass, 0 = up to 16)."""
    L = N.lib()
    return _read(lambda p, v, n: L.wahgpu_read_values_text_threads(p, int(threads), v, n), path)


def write_values_text(path, values, sep="\n", threads=0):
    """Writes values like the `generate` program: decimal, separated by sep,
    then a newline."""
    v = _as_u32(values)
    rc = N.lib().wahgpu_write_values_text(str(path).encode(), v.ctypes.data_as(ctypes.c_void_p),
                                          v.size, sep.encode(), int(threads))
    if rc != N.OK:
        raise WahError(rc, f"cannot write {path}")


def read_values_binary(path):
    return _read(N.lib().wahgpu_read_values_binary, path)


def index_checksum(keys, offsets, lengths, words):
    """64-bit order-sensitive checksum of an index (wahgpu_index_checksum):
    equal for equal keys, offsets, lengths and words.  Runs without the GIL."""
    k, o, ln, w = (_as_u32(a) for a in (keys, offsets, lengths, words))
    if not (k.size == o.size == ln.size):
        raise ValueError("keys, offsets and lengths differ in size")
    return int(N.lib().wahgpu_index_checksum(k.ctypes.data_as(ctypes.c_void_p),
                                             o.ctypes.data_as(ctypes.c_void_p),
                                             ln.ctypes.data_as(ctypes.c_void_p), k.size,
                                             w.ctypes.data_as(ctypes.c_void_p), w.size))


def decode_bitmap(words):
    """'0'/'1' text of a key's bitmap, trailing zeros trimmed (src/cpu.cpp:47-74)."""
    w = _as_u32(words)
    L = N.lib()
    need = L.wahgpu_decode_bitmap(w.ctypes.data_as(ctypes.c_void_p), w.size, None, 0)
    buf = ctypes.create_string_buffer(max(int(need), 1))
    L.wahgpu_decode_bitmap(w.ctypes.data_as(ctypes.c_void_p), w.size, buf, need)
    return buf.raw[:need].decode()


def generate(n, dist="uniform", seed=0, lo=0, hi=65535, cardinality=1 << 20, zipf_s=1.2,
             run_mean=4096.0, start=0, out=None):
    """Seeded synthetic column (the `generate` program's distributions).
    With out (a contiguous uint32 array of at least n entries) the values are
    written there."""
    p = N.GenParams()
    p.dist = {"uniform": N.DIST_UNIFORM, "zipf": N.DIST_ZIPF,
              "zipf-runs": N.DIST_ZIPF_RUNS}[dist]
    p.from_, p.to, p.cardinality = lo, hi, cardinality
    p.zipf_s, p.run_mean, p.seed = zipf_s, run_mean, seed
    if out is None:
        out = np.empty(max(int(n), 1), np.uint32)
    elif out.dtype != np.uint32 or not out.flags.c_contiguous or out.size < n:
        raise ValueError("out must be a contiguous uint32 array of at least n entries")
    rc = N.lib().wahgpu_generate_range(ctypes.byref(p), out.ctypes.data_as(ctypes.c_void_p),
                                       int(start), int(n))
    if rc != N.OK:
        raise WahError(rc, "bad generator parameters")
    return out[:n]


class Merger:
    """Exact stitching of partial indexes over consecutive row ranges."""

    def __init__(self):
        self._lib = N.lib()
        h = ctypes.c_void_p()
        rc = self._lib.wahgpu_merger_create(ctypes.byref(h))
        if rc != N.OK:
            raise WahError(rc, "merger")
        self._h = h

    def __del__(self):
        if getattr(self, "_h", None):
            self._lib.wahgpu_merger_destroy(self._h)
            self._h = None

    def add(self, keys, offsets, words):
        k, o, w = _as_u32(keys), _as_u32(offsets), _as_u32(words)
        rc = self._lib.wahgpu_merger_add(self._h, k.size, w.size,
                                         k.ctypes.data_as(ctypes.c_void_p),
                                         o.ctypes.data_as(ctypes.c_void_p),
                                         w.ctypes.data_as(ctypes.c_void_p))
        if rc != N.OK:
            raise WahError(rc, "malformed partial index")

    def result(self):
        K = int(self._lib.wahgpu_merger_keycnt(self._h))
        L = int(self._lib.wahgpu_merger_length(self._h))
        keys, offs, lens = (np.empty(max(K, 1), np.uint32) for _ in range(3))
        words = np.empty(max(L, 1), np.uint32)
        self._lib.wahgpu_merger_export(self._h, *(a.ctypes.data_as(ctypes.c_void_p)
                                                  for a in (keys, offs, lens, words)))
        return Index(keys[:K], offs[:K], lens[:K], words[:L], 0, L, K)


class _PinnedBlock:
    """Owns one wahgpu_host_alloc block; freed with the last array on it."""

    def __init__(self, nbytes):
        self.ptr = ctypes.c_void_p()
        rc = N.lib().wahgpu_host_alloc(int(nbytes), ctypes.byref(self.ptr))
        if rc != N.OK:
            raise WahError(rc, f"pinned allocation of {nbytes} bytes failed")

    def __del__(self):
        if self.ptr:
            N.lib().wahgpu_host_free(self.ptr)
            self.ptr = None


def pinned_empty(n, dtype=np.uint32):
    """Uninitialised numpy array in page-locked host memory."""
    dt = np.dtype(dtype)
    nbytes = max(int(n), 1) * dt.itemsize
    blk = _PinnedBlock(nbytes)
    raw = (ctypes.c_uint8 * nbytes).from_address(blk.ptr.value)
    raw._block = blk  # the ctypes view keeps the block alive for the array
    return np.frombuffer(raw, dtype=dt)[:int(n)]


class _Pending:
    """A submitted request: keeps the values and result buffers alive."""
    __slots__ = ("id", "values", "bufs", "res")


class Server:
    """Continuous-batching build server (wahgpu_server_*): requests queue as
    they are submitted; worker slots build every queued request that fits
    their HBM budget as one batch.  plain=True serves one request at a time
    through the per-column drop-in build instead."""

    def __init__(self, device=0, slots=2, arena_bytes=8 << 30, max_batch_cols=1024,
                 max_batch_values=0, plain=False):
        self._lib = N.lib()
        cfg = N.ServerConfig(int(device), int(slots), int(arena_bytes), int(max_batch_cols),
                             int(max_batch_values), 1 if plain else 0, 0)
        h = ctypes.c_void_p()
        rc = self._lib.wahgpu_server_create(ctypes.byref(h), ctypes.byref(cfg))
        if rc != N.OK:
            raise WahError(rc, "wahgpu_server_create failed")
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            self._lib.wahgpu_server_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def submit(self, values, rid_base=0, out=None):
        """Queue one column.  out = (keys, offsets, lengths, words) result
        arrays (n, n, n, 2n uint32; pinned_empty ones copy fastest), or None
        for fresh pageable arrays."""
        p = _Pending()
        p.values = _as_u32(values)
        cap = max(p.values.size, 1)
        if out is None:
            p.bufs = (np.empty(cap, np.uint32), np.empty(cap, np.uint32),
                      np.empty(cap, np.uint32), np.empty(2 * cap, np.uint32))
        else:
            p.bufs = tuple(out)
            if (any(b.dtype != np.uint32 or not b.flags.c_contiguous for b in p.bufs)
                    or min(b.size for b in p.bufs[:3]) < p.values.size
                    or p.bufs[3].size < 2 * p.values.size):
                raise ValueError("out needs contiguous uint32 arrays of n, n, n and 2n entries")
        r = N.Result()
        r.keys, r.offsets, r.lengths = (x.ctypes.data for x in p.bufs[:3])
        r.keys_cap = min(b.size for b in p.bufs[:3])
        r.words = p.bufs[3].ctypes.data
        r.words_cap = p.bufs[3].size
        p.res = r
        rid = ctypes.c_uint64(0)
        rc = self._lib.wahgpu_server_submit(self._h, p.values.ctypes.data_as(ctypes.c_void_p),
                                            p.values.size, int(rid_base), ctypes.byref(r),
                                            ctypes.byref(rid))
        if rc != N.OK:
            raise WahError(rc, "wahgpu_server_submit failed")
        p.id = rid.value
        return p

    def _index(self, p):
        c = p.res.counts
        K, L = int(c.keycnt), int(c.length)
        k, o, ln, w = p.bufs
        return Index(k[:K], o[:K], ln[:K], w[:L], int(c.k), L, K, int(c.bound), int(c.passes),
                     mode=int(c.mode))

    def poll(self, p):
        """-> Index when done, None while pending (raises on failure)."""
        rc = self._lib.wahgpu_server_poll(self._h, p.id)
        if rc == N.EBUSY:
            return None
        if rc != N.OK:
            self._fail(rc, p)
        return self._index(p)

    def _fail(self, rc, p):
        msg = self._lib.wahgpu_server_last_error(self._h)
        raise WahError(rc, f"request {p.id} failed: {msg.decode() if msg else ''}")

    def wait(self, p):
        rc = self._lib.wahgpu_server_wait(self._h, p.id)
        if rc != N.OK:
            self._fail(rc, p)
        return self._index(p)

    def drain(self):
        rc = self._lib.wahgpu_server_drain(self._h)
        if rc != N.OK:
            raise WahError(rc, "drain failed")

    def stats(self):
        st = N.ServerStats()
        self._lib.wahgpu_server_get_stats(self._h, ctypes.byref(st))
        return {name: getattr(st, name) for name, _ in N.ServerStats._fields_}

    def reset_stats(self):
        self._lib.wahgpu_server_reset_stats(self._h)
