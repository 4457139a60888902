"""The C ABI called directly through ctypes, as a cgo / JNI / ctypes host
would bind it (the binding shape shown for other hosts): the reference's
main calls one by one, a whole program, the circuit parser and the error
convention, checked against the oracle bit-exactly."""
import ctypes
import os

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "plaquette-simulator_amd", "libplaquette_amd.so")
u64, i64, u32, f32 = ctypes.c_uint64, ctypes.c_int64, ctypes.c_uint32, ctypes.c_float


class Op(ctypes.Structure):
    _fields_ = [(k, ctypes.c_uint32) for k in ("code", "q0", "q1", "slot", "t0", "t1", "t2", "reserved")]


@pytest.fixture(scope="module")
def L():
    lib = ctypes.CDLL(LIB)
    lib.pq_last_error.restype = ctypes.c_char_p
    c = ctypes.c_int()
    lib.pq_device_count(ctypes.byref(c))
    if c.value < 1:
        pytest.skip("no GPU")
    return lib


def ok(L, rc):
    assert rc == 0, L.pq_last_error().decode()


def test_calls_one_by_one(L):
    n, B, seed = 5, 100, 7
    h = ctypes.c_void_p()
    ok(L, L.pq_create(u64(n), u64(B), i64(seed), 1, ctypes.byref(h)))
    orc = O.OracleState(n, B, seed)
    ok(L, L.pq_gate(h, u32(1), u64(0), u64(0)))               # H 0
    ok(L, L.pq_gate(h, u32(6), u64(0), u64(1)))               # CX 0 -> 1
    ok(L, L.pq_gate(h, u32(2), u64(2), u64(0)))               # S 2
    orc.apply_hadamard_gate(0)
    orc.apply_control_not_gate(0, 1)
    orc.apply_phase_gate(2)
    rec = u64()
    for q in (1, 0, 2):
        ok(L, L.pq_measure(h, u64(q), f32(0.5), i64(-1), None, ctypes.byref(rec)))
        orc.measure_qubit(q)
    res, det = ctypes.c_int32(), ctypes.c_int32()
    for r in range(3):
        for b in range(B):
            ok(L, L.pq_get_measurement(h, u64(r), u64(b), ctypes.byref(res), ctypes.byref(det)))
            assert (res.value, det.value) == orc.get_measurement(r, b)
    x = np.zeros(B * (2 * n + 1) * n, np.int32)
    z = np.zeros_like(x)
    r = np.zeros(B * (2 * n + 1), np.int32)
    P = ctypes.POINTER(ctypes.c_int32)
    ok(L, L.pq_device_to_host(h, x.ctypes.data_as(P), z.ctypes.data_as(P), r.ctypes.data_as(P)))
    ox, oz, orr = orc.device_to_host()
    assert np.array_equal(x, ox) and np.array_equal(z, oz) and np.array_equal(r, orr)
    # errors come back as status codes with a message, never as exceptions
    assert L.pq_gate(h, u32(1), u64(n), u64(0)) != 0 and L.pq_last_error()
    assert L.pq_get_measurement(h, u64(99), u64(0), ctypes.byref(res), ctypes.byref(det)) != 0
    ok(L, L.pq_destroy(h))


def test_program_and_parser(L):
    text = b"H 0 1 2\nCX 0 3 1 4\nM 3 4\nCZ 2 3\nS 1\nM 0 1 2 3 4\n"
    nops, nlines, used = u64(), u64(), u64()
    ok(L, L.pq_circuit_parse(text, u64(len(text)), u32(0), None, u64(0), ctypes.byref(nops),
                             ctypes.byref(nlines), ctypes.byref(used)))
    ops = (Op * nops.value)()
    ok(L, L.pq_circuit_parse(text, u64(len(text)), u32(0), ops, nops, ctypes.byref(nops),
                             ctypes.byref(nlines), ctypes.byref(used)))
    assert (nlines.value, used.value) == (6, 5)
    n, B, seed = 5, 300, 11
    arr = np.ctypeslib.as_array(ctypes.cast(ops, ctypes.POINTER(ctypes.c_uint32)),
                                shape=(nops.value, 8)).copy()
    for form in (0, 1, 2):  # plain, persistent, graph
        h = ctypes.c_void_p()
        ok(L, L.pq_create(u64(n), u64(B), i64(seed), 1, ctypes.byref(h)))
        ok(L, L.pq_run_program(h, ops, nops, form))
        orc = O.OracleState(n, B, seed)
        orc.run_program(arr.astype(np.int32))
        res = np.zeros(B, np.int32)
        det = np.zeros(B, np.int32)
        P = ctypes.POINTER(ctypes.c_int32)
        for i, (oref, dref) in enumerate(orc.records):
            ok(L, L.pq_get_record(h, u64(i), res.ctypes.data_as(P), det.ctypes.data_as(P)))
            assert np.array_equal(res, oref) and np.array_equal(det, dref), (form, i)
        ok(L, L.pq_destroy(h))


def test_raw_tableau_pointer_sees_resets_and_frame_runs(L):
    """Once the raw tableau pointer is handed out, a reset and a frame-form
    run write the whole tableau at once (no X/Z left for later): the memory
    behind the pointer holds |0...0> after the reset and the run's tableau
    after the run."""
    hip = ctypes.CDLL("libamdhip64.so")
    hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    text = b"H 0 1 2\nCX 0 3 1 4\nM 3 4\nCZ 2 3\nS 1\nM 0 1 2 3 4\n"
    nops, nlines, used = u64(), u64(), u64()
    ok(L, L.pq_circuit_parse(text, u64(len(text)), u32(0), None, u64(0), ctypes.byref(nops),
                             ctypes.byref(nlines), ctypes.byref(used)))
    ops = (Op * nops.value)()
    ok(L, L.pq_circuit_parse(text, u64(len(text)), u32(0), ops, nops, ctypes.byref(nops),
                             ctypes.byref(nlines), ctypes.byref(used)))
    n, B, seed = 5, 200, 3
    h = ctypes.c_void_p()
    ok(L, L.pq_create(u64(n), u64(B), i64(seed), 1, ctypes.byref(h)))
    tab, W, R = ctypes.c_void_p(), u64(), u64()
    ok(L, L.pq_device_handles(h, ctypes.byref(tab), None, ctypes.byref(W), ctypes.byref(R), None))
    words = W.value * R.value * R.value

    def raw():
        ok(L, L.pq_synchronize(h))
        a = np.zeros(words, np.uint64)
        assert hip.hipMemcpy(a.ctypes.data, tab.value, words * 8, 2) == 0
        return a

    ok(L, L.pq_run_program(h, ops, nops, 1))  # persistent: X/Z change
    ok(L, L.pq_reset(h))
    ident = np.zeros((W.value, R.value, R.value), np.uint64)
    for c in range(R.value):
        ident[:, c, c] = np.uint64(0xFFFFFFFFFFFFFFFF)
    assert np.array_equal(raw(), ident.ravel()), "reset left the raw tableau stale"
    ok(L, L.pq_run_program(h, ops, nops, 3))  # frame form
    got = raw()
    planes = np.zeros(words, np.uint64)
    ok(L, L.pq_read_planes(h, planes.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64))))
    assert np.array_equal(got, planes), "frame run left the raw tableau's X/Z stale"
    ok(L, L.pq_destroy(h))
