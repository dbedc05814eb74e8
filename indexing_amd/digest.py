"""Content digests of columns and built indexes.

A digest is SHA-256 over the little-endian uint32 arrays, each preceded by
its length as a little-endian uint64, truncated to 32 hex digits.  Two
indexes have the same digest exactly when their keys, offsets, lengths and
words agree entry for entry (up to a hash collision), so a full-size build
can be compared with a recorded one without keeping its gigabytes around.
"""
import hashlib

import numpy as np


def _update(h, a):
    a = np.ascontiguousarray(a, dtype=np.uint32)
    h.update(np.uint64(a.size).tobytes())
    if a.size:
        h.update(memoryview(a).cast("B"))


def array_digest(a):
    """Digest of one uint32 array (e.g. a generated column)."""
    h = hashlib.sha256()
    _update(h, a)
    return h.hexdigest()[:32]


def index_digest(keys, offsets, lengths, words):
    """Digest of an index given as its four uint32 arrays."""
    h = hashlib.sha256()
    for a in (keys, offsets, lengths, words):
        _update(h, a)
    return h.hexdigest()[:32]
