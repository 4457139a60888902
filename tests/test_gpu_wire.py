"""GPU output in the wire formats: the digests of the engine's own device
words (planes and records, read back with no unpacking) equal the committed
digests of the oracle's output (tests/golden/digests.json), in every
execution form; DLPack export hands the same words to torch without a copy;
a state's files read back bit-exactly."""
import json
import os
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
DIG = json.load(open(os.path.join(GOLD, "digests.json")))


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


@pytest.mark.parametrize("form", ["plain", "persistent", "graph", "frame"])
def test_random_n64_digests(ps, form):
    from plaquette_simulator import wire, workloads as W
    w = W.config_random_n64()
    say(f"[case] random n64 digests form={form}")
    st = ps.CliffordState(w.num_qubits, w.batch_size, w.seed, True)
    st.run_program(w.ops, form)
    assert wire.digest(st.read_planes()) == DIG["random_n64"]["planes_sha256"]
    assert wire.digest(st.read_records()) == DIG["random_n64"]["records_sha256"]


@pytest.mark.parametrize("form", ["persistent", "frame"])
def test_surface_d9_slice_digests(ps, form):
    from plaquette_simulator import wire, workloads as W
    w = W.config_surface_d9()
    say(f"[case] d9 slice digests form={form}")
    st = ps.CliffordState(w.num_qubits, w.batch_size, w.seed, True)
    st.run_program(w.ops, form)
    planes = st.read_planes()
    recs = st.read_records()
    assert wire.digest(planes[:2]) == DIG["surface_d9_slice"]["planes_sha256"]
    assert wire.digest(recs[:, :, :2]) == DIG["surface_d9_slice"]["records_sha256"]


_DLPACK_CHECK = r"""
import sys
import numpy as np
import torch  # first: its HIP runtime must be the one the process loads
assert torch.cuda.is_available()
sys.path[:0] = sys.argv[1:]
import plaquette_simulator as ps
from plaquette_simulator import workloads as W
ops, n = W.surface_code_memory_z(3, 3, 0.01)
st = ps.CliffordState(n, 300, 5, True)
st.run_program(ops, "persistent")
t = torch.from_dlpack(st.planes_dlpack())
assert t.is_cuda and tuple(t.shape) == (5, 2 * n + 1, 2 * n + 1)
assert np.array_equal(t.cpu().numpy().astype(np.uint64), st.read_planes())
rt = torch.from_dlpack(st.records_dlpack())
assert np.array_equal(rt.cpu().numpy().astype(np.uint64), st.read_records())
# zero copy: the engine's own later writes show through the tensor
st.apply_hadamard_gate(0)
st.synchronize()
assert np.array_equal(t.cpu().numpy().astype(np.uint64), st.read_planes())
del t, rt, st
print("dlpack ok")
"""


def test_dlpack_zero_copy(ps):
    # torch carries its own HIP runtime; in a process where the engine's
    # runtime came first torch may see no device, so the check runs in a
    # child process that imports torch before the engine
    pytest.importorskip("torch")
    import subprocess
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    say("[case] dlpack export (child process)")
    r = subprocess.run([sys.executable, "-c", _DLPACK_CHECK,
                        os.path.join(root, "plaquette-simulator_amd")],
                       capture_output=True, text=True, timeout=240)
    assert r.returncode == 0 and "dlpack ok" in r.stdout, r.stdout[-2000:] + r.stderr[-2000:]


def test_state_files(ps, tmp_path):
    from plaquette_simulator import wire, workloads as W
    ops, n = W.surface_code_memory_z(3, 2, 0.02)
    st = ps.CliffordState(n, 200, 9, True)
    st.run_program(ops, "persistent")
    d = wire.save_state(st, str(tmp_path / "t.pqw"), str(tmp_path / "r.pqw"))
    t = wire.load(str(tmp_path / "t.pqw"))
    r = wire.load(str(tmp_path / "r.pqw"))
    assert t["digest"] == d["planes"] and r["digest"] == d["records"]
    x, z, rr = wire.unpack_tableau(t["words"], 200)
    gx, gz, gr = st.device_to_host_numpy()
    assert np.array_equal(x, gx) and np.array_equal(z, gz) and np.array_equal(rr, gr)
    for i, (res, det) in enumerate(wire.unpack_records(r["words"], 200)):
        a, b = st.get_record(i)
        assert np.array_equal(res, a) and np.array_equal(det, b)


def test_compare_states(ps):
    from plaquette_simulator import workloads as W
    ops, n = W.surface_code_memory_z(3, 3, 0.01)
    a = ps.CliffordState(n, 200, 4, True)
    b = ps.CliffordState(n, 200, 4, True)
    a.run_program(ops, "persistent")
    b.run_program(ops, "graph")
    assert a.compare(b) == (0, 0)
    b.apply_hadamard_gate(0)
    dp, dr = a.compare(b)
    assert dp > 0 and dr == 0
    c = ps.CliffordState(n, 200, 5, True)  # another seed: other coins and noise
    c.run_program(ops, "persistent")
    assert a.compare(c)[1] > 0
