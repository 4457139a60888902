"""The wire formats on the CPU: bit packing against the oracle's int32
layout, the on-disk format round trip and its refusals, and the committed
fixtures (tests/golden/digests.json, cfg0_records.pqw) against a fresh
oracle run."""
import json
import os

import numpy as np
import pytest

import oracle as O
from plaquette_simulator import wire, workloads as W

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def test_pack_bit_positions():
    """word(w, col, row) bit s is shot 64w+s; col < n x, then z, then r."""
    rng = np.random.default_rng(3)
    n, B = 4, 150
    x = rng.integers(0, 2, (B, 2 * n + 1, n)).astype(np.int32)
    z = rng.integers(0, 2, (B, 2 * n + 1, n)).astype(np.int32)
    r = rng.integers(0, 2, (B, 2 * n + 1)).astype(np.int32)
    p = wire.pack_tableau(x, z, r)
    assert p.shape == (3, 2 * n + 1, 2 * n + 1) and p.dtype == np.uint64
    for b, row, q in ((0, 0, 0), (63, 8, 3), (64, 5, 1), (149, 2, 2)):
        w, s = divmod(b, 64)
        assert (int(p[w, q, row]) >> s) & 1 == x[b, row, q]
        assert (int(p[w, n + q, row]) >> s) & 1 == z[b, row, q]
        assert (int(p[w, 2 * n, row]) >> s) & 1 == r[b, row]
    # shots past B in the last word are zero
    assert not np.any(p[2] >> np.uint64(150 - 128))
    back = wire.unpack_tableau(p, B)
    assert all(np.array_equal(a, b) for a, b in zip(back, (x, z, r)))


def test_records_round_trip():
    rng = np.random.default_rng(4)
    B = 70
    recs = [(rng.integers(0, 2, B).astype(np.int32), rng.integers(0, 2, B).astype(np.int32))
            for _ in range(5)]
    words = wire.pack_records(recs, B)
    assert words.shape == (5, 2, 2)
    assert (int(words[3, 1, 1]) >> 5) & 1 == recs[3][1][69]
    for (a, b), (c, d) in zip(recs, wire.unpack_records(words, B)):
        assert np.array_equal(a, c) and np.array_equal(b, d)


def test_file_round_trip_and_refusals(tmp_path):
    rng = np.random.default_rng(5)
    words = rng.integers(0, 2 ** 63, (3, 2, 4), dtype=np.uint64)
    path = str(tmp_path / "r.pqw")
    d = wire.save(path, wire.KIND_RECORDS, 7, 250, -3, words)
    f = wire.load(path)
    assert f["digest"] == d == wire.digest(words)
    assert (f["kind"], f["n"], f["B"], f["seed"]) == (wire.KIND_RECORDS, 7, 250, -3)
    assert np.array_equal(f["words"], words)
    raw = open(path, "rb").read()
    assert len(raw) == 64 + words.size * 8 + 32
    bad = bytearray(raw)
    bad[64] ^= 1
    open(path, "wb").write(bytes(bad))
    with pytest.raises(ValueError, match="digest"):
        wire.load(path)
    open(path, "wb").write(b"NOTWIRE!" + raw[8:])
    with pytest.raises(ValueError, match="magic"):
        wire.load(path)
    open(path, "wb").write(raw[:-1])
    with pytest.raises(ValueError, match="size"):
        wire.load(path)


def test_committed_fixtures_match_the_oracle():
    dg = json.load(open(os.path.join(GOLD, "digests.json")))
    w = W.config_random_n64()
    orc = O.OracleState(w.num_qubits, w.batch_size, w.seed)
    orc.run_program(w.ops.astype(np.int32))
    planes = wire.pack_tableau(orc.x, orc.z, orc.r)
    recs = wire.pack_records(orc.records, w.batch_size)
    assert wire.digest(planes) == dg["random_n64"]["planes_sha256"]
    assert wire.digest(recs) == dg["random_n64"]["records_sha256"]
    f = wire.load(os.path.join(GOLD, "cfg0_records.pqw"))
    assert f["kind"] == wire.KIND_RECORDS and (f["n"], f["B"], f["seed"]) == (64, 1024, 123423)
    assert np.array_equal(f["words"], recs)
    for (a, b), (c, d) in zip(orc.records, wire.unpack_records(f["words"], w.batch_size)):
        assert np.array_equal(a, c) and np.array_equal(b, d)
    # the d=9 slice: the first two word groups of the 16384-shot job
    w = W.config_surface_d9()
    orc = O.OracleState(w.num_qubits, 128, w.seed)
    orc.run_program(w.ops.astype(np.int32))
    assert wire.digest(wire.pack_tableau(orc.x, orc.z, orc.r)) == \
        dg["surface_d9_slice"]["planes_sha256"]
    assert wire.digest(wire.pack_records(orc.records, 128)) == \
        dg["surface_d9_slice"]["records_sha256"]
