"""Definitions the tests check results against, independent of the oracle's
sort/merge pipeline: the WAH encoding of each key's bitmap written from the
bitmap itself, the decoder, and the structural invariants.
"""
import numpy as np


def encode_from_bitmaps(values, rid_base=0):
    """Canonical index from first principles: for each distinct value (ascending)
    take its set of rows, group them by 31-row chunk, and emit per occupied chunk
    c (ascending) a zero-fill word of the gap since the previous occupied chunk
    (c itself for the first one) when nonzero, then the literal
    0x80000000 | OR of 1 << (row % 31)."""
    v = np.asarray(values, dtype=np.uint64)
    rows = np.arange(v.size, dtype=np.uint64) + rid_base
    keys, offsets, lengths, words = [], [], [], []
    order = np.lexsort((rows, v))
    vs, rs = v[order], rows[order]
    i = 0
    while i < vs.size:
        j = i
        while j < vs.size and vs[j] == vs[i]:
            j += 1
        chunks = {}
        for r in rs[i:j]:
            c = int(r) // 31
            chunks[c] = chunks.get(c, 0) | (1 << (int(r) % 31))
        keys.append(int(vs[i]))
        offsets.append(len(words))
        prev = None
        for c in sorted(chunks):
            gap = c if prev is None else c - prev - 1
            if gap:
                words.append(gap)
            words.append(0x80000000 | chunks[c])
            prev = c
        lengths.append(len(words) - offsets[-1])
        i = j
    u = lambda a: np.array(a, dtype=np.uint32)
    return u(keys), u(offsets), u(lengths), u(words)


def decode_rows(words):
    """Rows set in a key's bitmap (src/cpu.cpp:47-74 semantics, as row numbers)."""
    rows, pos = [], 0
    for w in np.asarray(words, dtype=np.uint64):
        w = int(w)
        if w & 0x80000000:
            for b in range(31):
                if (w >> b) & 1:
                    rows.append(pos + b)
            pos += 31
        else:
            assert not (w & 0x40000000), "the build never emits one-fills"
            pos += 31 * (w & 0x3FFFFFFF)
    return rows


def check_invariants(values, keys, offsets, lengths, words, rid_base=0, full_decode=True):
    """Structural invariants of the index (each key: [fill] literal ... ending in
    a literal; fills nonzero with bits 31/30 clear; keys strictly ascending;
    offsets the exclusive scan of lengths; decoding gives exactly the rows of
    each value)."""
    values = np.asarray(values, dtype=np.uint32)
    keys, offsets, lengths, words = (np.asarray(a, dtype=np.uint32)
                                     for a in (keys, offsets, lengths, words))
    assert keys.size == offsets.size == lengths.size
    assert np.all(np.diff(keys.astype(np.int64)) > 0)
    excl = np.concatenate([[0], np.cumsum(lengths.astype(np.uint64))[:-1]]) if keys.size else []
    assert np.array_equal(offsets.astype(np.uint64), np.asarray(excl, dtype=np.uint64))
    assert int(lengths.astype(np.uint64).sum()) == words.size
    uniq, counts = np.unique(values, return_counts=True)
    assert np.array_equal(uniq, keys)
    lit = (words & 0x80000000) != 0
    assert np.all(words[~lit] & 0x40000000 == 0)
    assert np.all(words[~lit] != 0)
    assert np.all(words[lit] & 0x7FFFFFFF != 0)
    # a fill is always followed by a literal of the same key; keys end in literals
    ends = (offsets.astype(np.int64) + lengths.astype(np.int64) - 1)
    if words.size:
        assert np.all(lit[ends])
        nxt = np.flatnonzero(~lit) + 1
        assert np.all(nxt < words.size) and np.all(lit[nxt])
    # popcount of literal payloads per key = occurrences of the key
    pc = np.zeros(words.size, np.int64)
    if words.size:
        pay = np.where(lit, words & 0x7FFFFFFF, 0).astype(np.uint64)
        pc = np.array([bin(int(x)).count("1") for x in pay]) if words.size < 200000 else \
            np.unpackbits(pay.astype('<u4').view(np.uint8)).reshape(-1, 32).sum(1)
        per_key = np.add.reduceat(pc, offsets.astype(np.int64)) if keys.size else []
        assert np.array_equal(np.asarray(per_key), counts)
    if full_decode:
        for i, key in enumerate(keys):
            w = words[offsets[i]:offsets[i] + lengths[i]]
            got = decode_rows(w)
            want = (np.flatnonzero(values == key) + rid_base).tolist()
            assert got == want, f"key {key}"
