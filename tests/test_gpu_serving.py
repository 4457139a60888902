"""Serving: requests admitted by memory budget and run concurrently give
results bit-identical to the same requests run alone."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ps():
    import plaquette_simulator as ps
    import plaquette_simulator_bindings as b
    if b.device_count() < 1:
        pytest.skip("no GPU")
    return ps


def say(msg):
    sys.stderr.write(msg + "\n")
    sys.stderr.flush()


def unpack(planes, B):
    """[nrec][2][W] uint64 -> (results, determined) int32 [nrec][B]."""
    bits = np.unpackbits(planes.view(np.uint8), axis=-1, bitorder="little")
    bits = bits.reshape(planes.shape[0], 2, -1)[:, :, :B].astype(np.int32)
    return bits[:, 0], bits[:, 1]


def requests():
    from plaquette_simulator import workloads as W
    rng = np.random.default_rng(9)
    out = []
    for i in range(12):
        d = int(rng.choice([3, 5, 7]))
        B = int(rng.choice([64, 100, 257, 640, 1000]))
        p = float(rng.choice([0.0, 1e-3, 5e-3]))
        ops, n = W.surface_code_memory_z(d, d, p)
        out.append((n, B, int(rng.integers(0, 2 ** 31)), ops))
    return out


def alone(ps, n, B, seed, ops):
    st = ps.CliffordState(n, B, seed, True)
    st.run_program(ops, "persistent")
    return [st.get_record(i) for i in range(st.num_records)]


@pytest.mark.parametrize("policy,form", [("fcfs", "persistent"), ("smallest_first", "persistent"),
                                         ("fcfs", "frame"), ("smallest_first", "frame")])
def test_server_results_equal_standalone(ps, policy, form):
    """Every request's records equal its standalone persistent run, in both
    forms; in the frame form requests of one distance share a plan."""
    reqs = requests()
    if form == "frame":  # and one request the frame form cannot take
        from plaquette_simulator import workloads as W
        ob = W.OpList()
        ob.h(0)
        m = ob.measure(0)
        ob.flagged(W.OP_H_FLAG, 1, m)
        ob.measure(1)
        reqs.append((3, 100, 77, ob.array()))
    biggest = max(((B + 63) // 64) * (2 * n + 1) ** 2 * 8 for n, B, _, _ in reqs)
    sv = ps.Server(budget_bytes=3 * biggest + (64 << 20), slots=4)
    sv.set_policy(policy, 0.05)
    sv.set_form(form)
    say(f"[case] server {len(reqs)} requests, budget {3 * biggest}, {policy}, {form}")
    tickets = [sv.submit(n, B, seed, ops) for n, B, seed, ops in reqs]
    sv.drain()
    stats = sv.stats()
    say(f"[stats] {stats}")
    assert stats["completed"] == len(reqs)
    assert stats["max_running"] >= 2  # requests overlapped on the GPU
    for t, (n, B, seed, ops) in zip(tickets, reqs):
        status, lat, q = sv.status(t)
        assert status == 2 and lat > 0 and q >= 0
        res, det = unpack(sv.result(t), B)
        ref = alone(ps, n, B, seed, ops)
        assert len(ref) == res.shape[0]
        for i, (r, dd) in enumerate(ref):
            assert np.array_equal(res[i], r) and np.array_equal(det[i], dd), (t, i)
        sv.release(t)
    if form == "frame":
        fs = sv.frame_stats()
        say(f"[frame] {fs}")
        assert fs["frame_runs"] == len(reqs) - 1     # all but the flagged-H request
        # one plan per structure: distance, and whether the circuit has noise ops
        shapes = {(n, bool((ops[:, 0] == 10).any())) for n, B, seed, ops in reqs[:-1]}
        assert fs["plan_builds"] == len(shapes)
        assert fs["plan_hits"] == fs["frame_runs"] - fs["plan_builds"]


def test_server_smallest_first_order(ps):
    """With one slot held by a running request, the smallest-first policy
    starts the queued requests in increasing size."""
    import time
    from plaquette_simulator import workloads as W
    big_ops, big_n = W.surface_code_memory_z(11, 11, 0.0)
    ops, n = W.surface_code_memory_z(5, 5, 0.0)
    sv = ps.Server(budget_bytes=1 << 36, slots=1)
    sv.set_policy("smallest_first", 0.0)
    blocker = sv.submit(big_n, 65536, 1, big_ops)
    t0 = time.time()
    while sv.status(blocker)[0] == 0 and time.time() - t0 < 30:
        time.sleep(1e-4)
    sizes = [4096, 2048, 64, 1024, 128]
    tickets = [sv.submit(n, B, 3, ops) for B in sizes]
    queued_behind = sv.status(blocker)[0] == 1
    sv.drain()
    starts = [sv.status(t)[2] for t in tickets]  # submits are microseconds apart
    order = [sizes[i] for i in sorted(range(len(sizes)), key=lambda i: starts[i])]
    say(f"[order] {order} (queued behind the blocker: {queued_behind})")
    if queued_behind:
        assert order == sorted(sizes)
    with pytest.raises(RuntimeError):
        sv.set_policy("lifo")


def test_server_rejects_oversized(ps):
    from plaquette_simulator import workloads as W
    ops, n = W.surface_code_memory_z(5, 5, 0.0)
    sv = ps.Server(budget_bytes=1 << 20, slots=2)
    with pytest.raises(RuntimeError):
        sv.submit(n, 4096, 1, ops)
    with pytest.raises(RuntimeError):
        sv.submit(n, 64, 1, np.array([[6, 0, 0, 0, 0, 0, 0, 0]], np.uint32))  # CX 0 -> 0
