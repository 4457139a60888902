"""Wire formats of tableaux and measurement records.

In memory (and on the GPU) a state is kept as shot-sliced bit-planes: one
uint64 word holds one tableau entry for 64 consecutive shots (bit s of word
w is shot 64w+s).

* planes:  uint64 [W][2n+1][2n+1] = word[w][col][row]; col < n is x_col,
  n <= col < 2n is z_(col-n), col 2n is r; rows 0..n-1 destabilizers,
  n..2n-1 stabilizers, 2n the scratch row.
* records: uint64 [nrec][2][W] = (result bits, determined bits) per record.

The host-side int32 forms are the reference's: x/z [B][2n+1][n], r [B][2n+1]
(CliffordState.hpp:238-259) and per record results/determined [B]
(Measurement::CopyToHost, CliffordState.hpp:39-57).  ``pack_*``/``unpack_*``
convert between the two; shots past B in the last word are zero.

On disk (``save``/``load``) a file holds one of these arrays:

    offset  size  field
    0       8     magic b"PQWIRE01"
    8       4     kind: 1 = planes, 2 = records (uint32, little-endian)
    12      4     reserved (0)
    16      8     n (qubits)
    24      8     B (shots)
    32      8     seed (int64)
    40      24    dims d0, d1, d2 of the word array (uint64 each)
    64      8*d0*d1*d2  words, little-endian uint64, C order
    end     32    SHA-256 of the word bytes

``digest`` is that SHA-256 as hex: parity fixtures pin a GPU result by the
digest of its packed words.
"""
from __future__ import annotations

import hashlib
import struct

import numpy as np

MAGIC = b"PQWIRE01"
KIND_PLANES, KIND_RECORDS = 1, 2
_HEADER = struct.Struct("<8sIIQQqQQQ")


def words_per_plane(B: int) -> int:
    return (int(B) + 63) // 64


def _pack_bits(a: np.ndarray, B: int) -> np.ndarray:
    """int array [B, ...] of 0/1 -> uint64 [W, ...] (bit s of word w = shot 64w+s)."""
    Wd = words_per_plane(B)
    bits = np.zeros((Wd * 64,) + a.shape[1:], np.uint8)
    bits[:B] = np.asarray(a) & 1
    bits = bits.reshape((Wd, 64) + a.shape[1:])
    weights = (np.uint64(1) << np.arange(64, dtype=np.uint64)).reshape((1, 64) + (1,) * (a.ndim - 1))
    return (bits.astype(np.uint64) * weights).sum(axis=1, dtype=np.uint64)


def _unpack_bits(words: np.ndarray, B: int) -> np.ndarray:
    """uint64 [W, ...] -> int32 [B, ...]."""
    Wd = words.shape[0]
    sh = np.arange(64, dtype=np.uint64).reshape((1, 64) + (1,) * (words.ndim - 1))
    bits = (words[:, None, ...] >> sh) & np.uint64(1)
    return bits.reshape((Wd * 64,) + words.shape[1:])[:B].astype(np.int32)


def pack_tableau(x: np.ndarray, z: np.ndarray, r: np.ndarray) -> np.ndarray:
    """int32 x, z [B][2n+1][n] and r [B][2n+1] -> planes uint64 [W][2n+1][2n+1]."""
    x = np.asarray(x)
    z = np.asarray(z)
    r = np.asarray(r)
    B, R, n = x.shape
    cols = np.concatenate([x.transpose(0, 2, 1), z.transpose(0, 2, 1), r[:, None, :]], axis=1)
    return _pack_bits(cols, B)  # [W][col][row]


def unpack_tableau(planes: np.ndarray, B: int):
    """planes uint64 [W][2n+1][2n+1] -> int32 x, z [B][2n+1][n], r [B][2n+1]."""
    cols = _unpack_bits(np.asarray(planes, np.uint64), B)  # [B][col][row]
    n = (cols.shape[1] - 1) // 2
    x = np.ascontiguousarray(cols[:, :n, :].transpose(0, 2, 1))
    z = np.ascontiguousarray(cols[:, n:2 * n, :].transpose(0, 2, 1))
    r = np.ascontiguousarray(cols[:, 2 * n, :])
    return x, z, r


def pack_records(records, B: int) -> np.ndarray:
    """[(results int32[B], determined int32[B]), ...] -> uint64 [nrec][2][W]."""
    Wd = words_per_plane(B)
    out = np.zeros((len(records), 2, Wd), np.uint64)
    for i, (res, det) in enumerate(records):
        both = np.stack([np.asarray(res), np.asarray(det)], axis=1)  # [B][2]
        out[i] = _pack_bits(both, B).T
    return out


def unpack_records(words: np.ndarray, B: int):
    """uint64 [nrec][2][W] -> [(results int32[B], determined int32[B]), ...]."""
    out = []
    for rec in np.asarray(words, np.uint64):
        both = _unpack_bits(rec.T.copy(), B)  # [B][2]
        out.append((both[:, 0].copy(), both[:, 1].copy()))
    return out


def digest(words: np.ndarray) -> str:
    """SHA-256 (hex) of the words as little-endian uint64 bytes, C order."""
    return hashlib.sha256(np.ascontiguousarray(words, dtype="<u8").tobytes()).hexdigest()


def save(path: str, kind: int, n: int, B: int, seed: int, words: np.ndarray) -> str:
    """Write one word array in the on-disk format; returns its digest."""
    w = np.ascontiguousarray(words, dtype="<u8")
    if w.ndim != 3:
        raise ValueError("planes and records are 3-dimensional word arrays")
    payload = w.tobytes()
    h = hashlib.sha256(payload)
    with open(path, "wb") as f:
        f.write(_HEADER.pack(MAGIC, int(kind), 0, int(n), int(B), int(seed), *map(int, w.shape)))
        f.write(payload)
        f.write(h.digest())
    return h.hexdigest()


def load(path: str) -> dict:
    """Read a file written by ``save``; checks magic, size and digest."""
    with open(path, "rb") as f:
        data = f.read()
    if len(data) < _HEADER.size + 32:
        raise ValueError(f"{path}: too short for a wire file")
    magic, kind, _, n, B, seed, d0, d1, d2 = _HEADER.unpack_from(data, 0)
    if magic != MAGIC:
        raise ValueError(f"{path}: not a wire file (magic {magic!r})")
    nbytes = 8 * d0 * d1 * d2
    if len(data) != _HEADER.size + nbytes + 32:
        raise ValueError(f"{path}: size does not match its header")
    payload = data[_HEADER.size:_HEADER.size + nbytes]
    if hashlib.sha256(payload).digest() != data[_HEADER.size + nbytes:]:
        raise ValueError(f"{path}: digest mismatch")
    words = np.frombuffer(payload, dtype="<u8").reshape(d0, d1, d2).astype(np.uint64)
    return {"kind": kind, "n": n, "B": B, "seed": seed, "words": words,
            "digest": hashlib.sha256(payload).hexdigest()}


def save_state(st, path_planes: str | None = None, path_records: str | None = None) -> dict:
    """A CliffordState's planes and/or records to files; returns their digests."""
    out = {}
    n, B, seed = st.get_num_qubits(), st.get_batch_size(), st.get_seed()
    if path_planes:
        out["planes"] = save(path_planes, KIND_PLANES, n, B, seed, st.read_planes())
    if path_records:
        out["records"] = save(path_records, KIND_RECORDS, n, B, seed, st.read_records())
    return out
