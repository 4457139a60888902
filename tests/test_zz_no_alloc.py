"""pq_run_program_static, the entry point that allocates no device memory,
called through the C ABI: two calls in a row (the second on a smaller state
and a shorter program, so the shared device array keeps the first call's
contents) each equal pq_run_program in the persistent form, bit-exact on
every tableau word and record bit; plus its refusals."""
import ctypes
import os
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "plaquette-simulator_amd", "libplaquette_amd.so")
u64, i64 = ctypes.c_uint64, ctypes.c_int64
PERSISTENT = 1
ERR_OOM, ERR_UNSUPPORTED = -4, -6


@pytest.fixture(scope="module")
def L():
    lib = ctypes.CDLL(LIB)
    lib.pq_last_error.restype = ctypes.c_char_p
    c = ctypes.c_int()
    lib.pq_device_count(ctypes.byref(c))
    if c.value < 1:
        pytest.skip("no GPU")
    return lib


@pytest.fixture(scope="module")
def W():
    from plaquette_simulator import workloads as W
    return W


def ok(L, rc):
    assert rc == 0, L.pq_last_error().decode()


def say(msg):
    sys.stderr.write(msg + "\n")
    sys.stderr.flush()


def ops_ptr(ops):
    a = np.ascontiguousarray(ops, dtype=np.uint32)
    return a, a.ctypes.data_as(ctypes.c_void_p), u64(a.shape[0])


def create(L, n, B, seed):
    h = ctypes.c_void_p()
    ok(L, L.pq_create(u64(n), u64(B), i64(seed), 1, ctypes.byref(h)))
    return h


def planes(L, h, n, B):
    Wd = (B + 63) // 64
    R = 2 * n + 1
    out = np.zeros(Wd * R * R, np.uint64)
    ok(L, L.pq_read_planes(h, out.ctypes.data_as(ctypes.c_void_p)))
    return out


def records(L, h, B):
    nrec = u64()
    ok(L, L.pq_info(h, None, None, None, ctypes.byref(nrec)))
    res = np.zeros(B, np.int32)
    det = np.zeros(B, np.int32)
    out = []
    for i in range(nrec.value):
        ok(L, L.pq_get_record(h, u64(i), res.ctypes.data_as(ctypes.c_void_p),
                              det.ctypes.data_as(ctypes.c_void_p)))
        out.append((res.copy(), det.copy()))
    return out


def measure_all(W, n, nrec):
    """Ops that create records 0..nrec-1 (to size a record arena)."""
    b = W.OpList()
    for k in range(nrec):
        b.measure(k % n)
    return b.array()


def run_pair(L, W, n, B, seed, ops, nrec):
    """The same set-up on two states, then pq_run_program on one and
    pq_run_program_static on the other."""
    warm = measure_all(W, n, nrec)
    hs = []
    for _ in range(2):
        h = create(L, n, B, seed)
        a, p, k = ops_ptr(warm)
        ok(L, L.pq_run_program(h, p, k, PERSISTENT))
        ok(L, L.pq_reset(h))
        hs.append(h)
    a, p, k = ops_ptr(ops)
    ok(L, L.pq_run_program(hs[0], p, k, PERSISTENT))
    ok(L, L.pq_run_program_static(hs[1], p, k))
    return hs


def test_static_entry_point_equals_pq_run_program(L, W):
    # first call: the d=5 surface code (noise planes for 4096 shots)
    ops1, n1 = W.surface_code_memory_z(5, 5, 0.01)
    B1 = 4096
    nrec1 = int((ops1[:, 0] == W.OP_M).sum())
    say(f"[case] static run 1: d=5 n={n1} B={B1} ops={len(ops1)}")
    h = run_pair(L, W, n1, B1, 7, ops1, nrec1)
    assert np.array_equal(planes(L, h[0], n1, B1), planes(L, h[1], n1, B1))
    r0, r1 = records(L, h[0], B1), records(L, h[1], B1)
    assert len(r0) == len(r1) == nrec1
    for i, ((a, b), (c, d)) in enumerate(zip(r0, r1)):
        assert np.array_equal(a, c) and np.array_equal(b, d), f"record {i}"
    # second call: a smaller state and program (the array keeps the first
    # call's larger contents)
    ops2, n2 = W.surface_code_memory_z(3, 2, 0.02)
    B2 = 200
    nrec2 = int((ops2[:, 0] == W.OP_M).sum())
    say(f"[case] static run 2: d=3 n={n2} B={B2} ops={len(ops2)}")
    g = run_pair(L, W, n2, B2, 11, ops2, nrec2)
    assert np.array_equal(planes(L, g[0], n2, B2), planes(L, g[1], n2, B2))
    s0, s1 = records(L, g[0], B2), records(L, g[1], B2)
    for i, ((a, b), (c, d)) in enumerate(zip(s0, s1)):
        assert np.array_equal(a, c) and np.array_equal(b, d), f"record {i}"
    # and a second run on the first state continues its coin/noise sequence
    # numbers as pq_run_program would
    say("[case] static run 3: d=5 again on the same states")
    a, p, k = ops_ptr(ops1)
    ok(L, L.pq_run_program(h[0], p, k, PERSISTENT))
    ok(L, L.pq_run_program_static(h[1], p, k))
    assert np.array_equal(planes(L, h[0], n1, B1), planes(L, h[1], n1, B1))
    for x in h + g:
        ok(L, L.pq_destroy(x))


def test_static_entry_point_against_oracle(L, W):
    import oracle as O
    ops, n = W.surface_code_memory_z(3, 3, 0.02)
    B = 130
    nrec = int((ops[:, 0] == W.OP_M).sum())
    h = create(L, n, B, 5)
    say(f"[case] static run vs oracle n={n} B={B}")
    a, p, k = ops_ptr(measure_all(W, n, nrec))
    ok(L, L.pq_run_program(h, p, k, PERSISTENT))
    ok(L, L.pq_reset(h))
    a, p, k = ops_ptr(ops)
    ok(L, L.pq_run_program_static(h, p, k))
    orc = O.OracleState(n, B, 5)
    orc.run_program(measure_all(W, n, nrec).astype(np.int32))
    orc.reset_clifford_state()
    orc.run_program(ops.astype(np.int32))
    x = np.zeros(B * (2 * n + 1) * n, np.int32)
    z = np.zeros_like(x)
    r = np.zeros(B * (2 * n + 1), np.int32)
    ok(L, L.pq_device_to_host(h, x.ctypes.data_as(ctypes.c_void_p),
                              z.ctypes.data_as(ctypes.c_void_p), r.ctypes.data_as(ctypes.c_void_p)))
    assert np.array_equal(x, orc.x.ravel()) and np.array_equal(z, orc.z.ravel())
    assert np.array_equal(r, orc.r.ravel())
    got = records(L, h, B)
    for i, (res, det) in enumerate(orc.records):
        assert np.array_equal(got[i][0], res) and np.array_equal(got[i][1], det), f"record {i}"
    ok(L, L.pq_destroy(h))


def test_static_entry_point_refusals(L, W):
    ops, n = W.surface_code_memory_z(3, 2, 0.02)
    h = create(L, n, 64, 1)
    a, p, k = ops_ptr(ops)
    # no record arena yet: refused before anything runs
    assert L.pq_run_program_static(h, p, k) == ERR_OOM
    assert b"record arena" in L.pq_last_error()
    # noise planes beyond the array: 2^20 shots of many noise ops
    big = create(L, 2, 1 << 20, 1)
    b = W.OpList()
    for _ in range(600):
        b.noise(0, 0.01, 0.01, 0.01)
    a2, p2, k2 = ops_ptr(b.array())
    assert L.pq_run_program_static(big, p2, k2) == ERR_UNSUPPORTED
    assert b"scratch" in L.pq_last_error()
    ok(L, L.pq_destroy(h))
    ok(L, L.pq_destroy(big))
