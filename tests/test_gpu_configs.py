"""The measured workloads other than the headline, on the GPU at their real
sizes: the n=1024 random Clifford circuit (4096 shots, 100000 ops), the
256-circuit mixed batch and a serving stream drawn from the serving
workload's distribution.  Each is checked bit-exactly: whole device word
arrays against an independent execution form or a standalone run (compared
on the device), and shot slices against the CPU oracle."""
import sys

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ps():
    import plaquette_simulator as ps
    import plaquette_simulator_bindings as b
    if b.device_count() < 1:
        pytest.skip("no GPU")
    return ps


@pytest.fixture(scope="module")
def W():
    from plaquette_simulator import workloads as W
    return W


def say(msg):
    sys.stderr.write(msg + "\n")
    sys.stderr.flush()


def same_planes(a, b):
    """Every tableau word equal (compared on the device)."""
    return a.compare(b)[0] == 0


def same_records(a, b):
    return a.num_records == b.num_records and np.array_equal(a.read_records(), b.read_records())


def oracle_slice(n, B, seed, ops, shot0=0):
    orc = O.OracleState(n, B, seed, shot0=shot0)
    orc.run_program(ops.astype(np.int32))
    return orc


# ---------------- n=1024, 4096 shots, 100000 ops ----------------

def test_random_n1024_full_persistent_equals_frame(ps, W):
    w = W.config_random_n1024()
    n, B = w.num_qubits, w.batch_size
    say(f"[case] n1024 full: persistent {len(w.ops)} ops, {B} shots")
    a = ps.CliffordState(n, B, w.seed, True)
    prog = ps.Program(a, w.ops)
    prog.set_census(True)
    prog.run(a, "persistent")
    a.sync_records()  # a Program run updates the records behind the state's back
    c = prog.read_census()
    say(f"[census] {c}")
    # the deep-scrambled regime: random measurements rowsum hundreds of rows
    assert c["random"] > 0 and c["random_rows"] / c["random"] > 100
    say("[case] n1024 full: frame form")
    b = ps.CliffordState(n, B, w.seed, True)
    b.run_program(w.ops, "frame")
    assert a.num_records == b.num_records == w.measurements
    assert same_records(a, b)
    assert same_planes(a, b)
    # plain form (one launch per op) on the first 128 shots: shots are
    # independent and coins are keyed by shot index, so a 128-shot state is
    # exactly the first two word groups of the 4096-shot job
    say("[case] n1024 full: plain form on shots [0, 128)")
    c_ = ps.CliffordState(n, 128, w.seed, True)
    c_.run_program(w.ops, "plain")
    full = a.read_planes()
    assert np.array_equal(c_.read_planes(), full[:2])
    assert np.array_equal(c_.read_records(), a.read_records()[:, :, :2])


def test_random_n1024_prefix_against_oracle(ps, W):
    """The longest prefix the oracle runs in a few seconds on 64 shots."""
    w = W.config_random_n1024()
    n = w.num_qubits
    ops = w.ops[:6000]
    say(f"[case] n1024 prefix {len(ops)} ops, 64 shots vs oracle")
    st = ps.CliffordState(n, 64, w.seed, True)
    st.run_program(ops, "persistent")
    orc = oracle_slice(n, 64, w.seed, ops)
    say(f"[oracle stats] {orc.stats.tolist()}")
    x, z, r = st.device_to_host_numpy()
    assert np.array_equal(x, orc.x) and np.array_equal(z, orc.z) and np.array_equal(r, orc.r)
    for i, (res, det) in enumerate(orc.records):
        gres, gdet = st.get_record(i)
        assert np.array_equal(gres, res) and np.array_equal(gdet, det), f"record {i}"


# ---------------- the 256-circuit mixed batch ----------------

def test_mixed_batch_256_members_equal_standalone(ps, W):
    from plaquette_simulator import wire
    ws = W.config_mixed_batch()
    assert len(ws) == 256
    say(f"[case] mixed batch: {len(ws)} circuits, d {min(w.meta['d'] for w in ws)}.."
        f"{max(w.meta['d'] for w in ws)}, shots {min(w.batch_size for w in ws)}.."
        f"{max(w.batch_size for w in ws)}")
    batch = ps.Batch()
    states, progs = [], []
    for w in ws:
        st = ps.CliffordState(w.num_qubits, w.batch_size, w.seed, True)
        pr = ps.Program(st, w.ops)
        batch.add(st, pr)
        states.append(st)
        progs.append(pr)
    batch.run("graph")
    for st in states:
        st.sync_records()
    say("[case] mixed batch: members against standalone runs")
    checked_oracle = 0
    for i, (w, st) in enumerate(zip(ws, states)):
        ref = ps.CliffordState(w.num_qubits, w.batch_size, w.seed, True)
        ref.run_program(w.ops, "persistent")
        assert same_records(st, ref), f"member {i} records"
        assert same_planes(st, ref), f"member {i} planes"
        del ref
        if w.meta["d"] <= 7 and checked_oracle < 12:
            orc = oracle_slice(w.num_qubits, 64, w.seed, w.ops)
            assert np.array_equal(st.read_planes()[:1],
                                  wire.pack_tableau(orc.x, orc.z, orc.r)), f"member {i} vs oracle"
            assert np.array_equal(st.read_records()[:, :, :1],
                                  wire.pack_records(orc.records, 64)), f"member {i} vs oracle"
            checked_oracle += 1
        if i % 32 == 31:
            say(f"[progress] {i + 1} members checked")
    assert checked_oracle >= 8
    # the frame form of the same batch gives the same records
    say("[case] mixed batch: frame form")
    recs = [st.read_records() for st in states]
    del batch, states, progs
    twins = []
    fb = ps.Batch()
    for w in ws:
        t = ps.CliffordState(w.num_qubits, w.batch_size, w.seed, True)
        tp = ps.Program(t, w.ops)
        fb.add(t, tp)
        twins.append((t, tp))
    fb.run("frame")
    for (t, _), r0 in zip(twins, recs):
        t.sync_records()
        assert np.array_equal(t.read_records(), r0)


# ---------------- a serving stream ----------------

def test_serving_stream_results_equal_standalone(ps, W):
    from plaquette_simulator import wire
    stream = W.config_serving_stream(200.0, 30)
    reqs = [(d, shots, p, seed) for _, d, shots, p, seed in stream]
    reqs.append((21, 65536, 1e-3, 12345))  # the largest request the stream allows
    say(f"[case] serving stream: {len(reqs)} requests, d up to {max(r[0] for r in reqs)}, "
        f"shots up to {max(r[1] for r in reqs)}")
    cache = {}

    def ops_of(d, p):
        if (d, p) not in cache:
            cache[(d, p)] = W.surface_code_memory_z(d, d, p)
        return cache[(d, p)]

    for form in ("persistent", "frame"):
        sv = ps.Server(budget_bytes=200 * 10 ** 9, slots=8)
        sv.set_form(form)
        tickets = []
        for d, shots, p, seed in reqs:
            ops, n = ops_of(d, p)
            tickets.append(sv.submit(n, shots, seed, ops))
        sv.drain()
        say(f"[case] serving {form}: {sv.stats()}")
        results = [sv.result(t) for t in tickets]
        for t in tickets:
            sv.release(t)
        del sv
        oracle_checked = 0
        for k, ((d, shots, p, seed), words) in enumerate(zip(reqs, results)):
            ops, n = ops_of(d, p)
            if form == "persistent":
                st = ps.CliffordState(n, shots, seed, True)
                st.run_program(ops, "persistent")
                assert np.array_equal(words, st.read_records()), f"request {k} ({form})"
                del st
                if d <= 7 and oracle_checked < 6:
                    orc = oracle_slice(n, 64, seed, ops)
                    assert np.array_equal(words[:, :, :1], wire.pack_records(orc.records, 64))
                    oracle_checked += 1
            else:
                assert np.array_equal(words, persistent_words[k]), f"request {k} ({form})"
        if form == "persistent":
            persistent_words = results
