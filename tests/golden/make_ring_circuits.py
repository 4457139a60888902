"""Writes ring_circuits.json: the reference's benchmark circuits
(benchmarks/circuits.tar.gz, read by benchmarks/run_circuit.cpp), as
{file name: circuit text}, for parity runs of the circuit front end.

The archive holds 60 members; one of them, width_depth_4_10.circuit, is
itself a gzip'd tar of the other 59 circuits rather than a circuit, so it is
recorded under "_not_a_circuit" (the reference's parser would reject it)."""
import io
import json
import os
import tarfile

SRC = "/root/reference/benchmarks/circuits.tar.gz"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "ring_circuits.json")


def main():
    circuits, skipped = {}, []
    with tarfile.open(SRC, "r:gz") as tf:
        for m in tf.getmembers():
            if not m.isfile():
                continue
            data = tf.extractfile(m).read()
            if data[:2] == b"\x1f\x8b":
                skipped.append(m.name)
                continue
            circuits[os.path.basename(m.name)] = data.decode("ascii")
    out = {"source": "benchmarks/circuits.tar.gz", "_not_a_circuit": skipped,
           "circuits": dict(sorted(circuits.items()))}
    with open(OUT, "w") as f:
        json.dump(out, f, indent=0)
    print(len(circuits), "circuits;", "skipped", skipped)


if __name__ == "__main__":
    main()
