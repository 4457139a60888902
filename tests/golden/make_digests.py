"""Write digests.json and cfg0_records.pqw: parity fixtures in the wire format.

The CPU oracle (oracle/, the reference's algorithm in its own int32 layout)
runs two workloads; its tableau and records are packed into the engine's
bit-plane words (plaquette_simulator.wire) and pinned by SHA-256:

* random_n64: n=64, 1024 shots, 2000 uniform H/S/CX ops with a measurement
  of a uniform qubit after every 100th op, seed 123423, the whole job;
* surface_d9_slice: the d=9 surface-code memory-Z experiment (161 qubits, 9
  rounds, p=1e-3, seed 7) on shots [0, 128), i.e. the first two 64-shot word
  groups of its 16384-shot run (coins and noise are keyed by shot index).

cfg0's records are also written as a wire file.  A GPU test compares the
digests of its own words with these; a CPU test re-packs the oracle's output
and reads the wire file back.  Run from the repository root:
    python tests/golden/make_digests.py
"""
import json
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
for p in (os.path.join(ROOT, "plaquette-simulator_amd"), os.path.join(ROOT, "oracle")):
    sys.path.insert(0, p)

import oracle as O  # noqa: E402
from plaquette_simulator import wire, workloads as W  # noqa: E402

HERE = os.path.dirname(os.path.abspath(__file__))


def oracle_words(n, B, seed, ops, shot0=0):
    orc = O.OracleState(n, B, seed, shot0=shot0)
    orc.run_program(ops.astype(np.int32))
    return wire.pack_tableau(orc.x, orc.z, orc.r), wire.pack_records(orc.records, B)


def main():
    out = {}
    w = W.config_random_n64()
    planes, recs = oracle_words(w.num_qubits, w.batch_size, w.seed, w.ops)
    out["random_n64"] = {"n": w.num_qubits, "B": w.batch_size, "seed": w.seed,
                         "shots": [0, w.batch_size], "words": list(planes.shape),
                         "records": list(recs.shape),
                         "planes_sha256": wire.digest(planes),
                         "records_sha256": wire.digest(recs)}
    wire.save(os.path.join(HERE, "cfg0_records.pqw"), wire.KIND_RECORDS, w.num_qubits,
              w.batch_size, w.seed, recs)
    w = W.config_surface_d9()
    planes, recs = oracle_words(w.num_qubits, 128, w.seed, w.ops)
    out["surface_d9_slice"] = {"n": w.num_qubits, "B": w.batch_size, "seed": w.seed,
                               "shots": [0, 128], "words": list(planes.shape),
                               "records": list(recs.shape),
                               "planes_sha256": wire.digest(planes),
                               "records_sha256": wire.digest(recs)}
    with open(os.path.join(HERE, "digests.json"), "w") as f:
        json.dump(out, f, indent=1)
        f.write("\n")
    print(json.dumps(out, indent=1))


if __name__ == "__main__":
    main()
