"""Profiling helpers for the engine (run under rocprofv3).

  python profiling.py probe
      One persistent run of the d=9 surface-code workload (16384 shots),
      preceded by a warm-up run and by a device-to-device copy of the whole
      tableau (W*(2n+1)^2*8 bytes) that serves as the counter calibration
      dispatch.  Run it under `rocprofv3 --pmc FETCH_SIZE` and again under
      `--pmc WRITE_SIZE`.

  python profiling.py kernels <rocprof csv dir> <out.md>
      Summarise a `--kernel-trace --stats --output-format csv` run.

  python profiling.py traffic <fetch csv dir> <write csv dir> <out.json> <out.md>
      Per-dispatch HBM traffic of k_program, each counter scaled by what the
      same counter reads for the calibration copy of known size.
"""
import csv
import glob
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))


def _state_and_program(shots=16384):
    sys.path.insert(0, os.path.dirname(HERE))
    import plaquette_simulator as ps
    from plaquette_simulator import workloads as W
    w = W.config_surface_d9(shots=shots)
    st = ps.CliffordState(w.num_qubits, w.batch_size, w.seed, True)
    prog = ps.Program(st, w.ops)
    return ps, w, st, prog


def probe(shots=16384):
    ps, w, st, prog = _state_and_program(shots)
    prog.run(st, "persistent")              # warm-up (also creates records)
    st.synchronize()
    other = ps.CliffordState(w.num_qubits, w.batch_size, w.seed, False)
    other.synchronize()
    other.device_to_device(st)              # calibration copy
    other.synchronize()
    st.reset_clifford_state()
    prog.run(st, "persistent")              # measured dispatch
    st.synchronize()
    n = w.num_qubits
    words = (w.batch_size + 63) // 64
    copy_bytes = words * (2 * n + 1) ** 2 * 8
    print(json.dumps({"workload": w.name, "copy_bytes": copy_bytes}))


def _find(d, pattern):
    hits = glob.glob(os.path.join(d, "**", pattern), recursive=True)
    if not hits:
        raise FileNotFoundError(f"{pattern} under {d}")
    return hits[0]


def kernels(d, out):
    path = _find(d, "*kernel_stats.csv")
    rows = list(csv.DictReader(open(path)))
    lines = ["| kernel | calls | total ms | avg us | min us | max us | % |", "|---|---|---|---|---|---|---|"]
    for r in rows:
        name = r["Name"].replace("(anonymous namespace)::", "").replace("pq::", "")
        name = name.split("(")[0]
        lines.append("| %s | %s | %.3f | %.2f | %.2f | %.2f | %.2f |" % (
            name[:60], r["Calls"], float(r["TotalDurationNs"]) / 1e6,
            float(r["AverageNs"]) / 1e3, float(r["MinNs"]) / 1e3,
            float(r["MaxNs"]) / 1e3, float(r["Percentage"])))
    # per-launch durations of the persistent kernel, in launch order
    try:
        trace = list(csv.DictReader(open(_find(d, "*kernel_trace.csv"))))
    except (FileNotFoundError, StopIteration, IndexError):
        trace = []
    for kname in ("k_program", "k_frame"):
        durs = [(int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e3
                for r in sorted(trace, key=lambda r: int(r["Start_Timestamp"]))
                if kname + "_t" in r["Kernel_Name"] or kname + "(" in r["Kernel_Name"]]
        if durs:
            srt = sorted(durs)
            lines += ["", "%s launches in order (us): " % kname + ", ".join("%.1f" % v for v in durs),
                      "%s median: %.1f us over %d launches" % (kname, srt[len(srt) // 2], len(srt))]
    with open(out, "w") as f:
        f.write("\n".join(lines) + "\n")
    print("\n".join(lines))


def _counter_rows(d):
    path = _find(d, "*counter_collection.csv")
    return list(csv.DictReader(open(path)))


def _per_dispatch(rows, counter):
    out = {}
    for r in rows:
        if r.get("Counter_Name") != counter:
            continue
        key = (int(r["Dispatch_Id"]), r["Kernel_Name"])
        out[key] = out.get(key, 0.0) + float(r["Counter_Value"])
    return out


def traffic(fetch_dir, write_dir, out_json, out_md, copy_bytes=None):
    if copy_bytes is None:
        n, words = 161, 256
        copy_bytes = words * (2 * n + 1) ** 2 * 8
    res = {"copy_bytes": copy_bytes}
    for label, d, counter in (("fetch", fetch_dir, "FETCH_SIZE"), ("write", write_dir, "WRITE_SIZE")):
        disp = _per_dispatch(_counter_rows(d), counter)
        copies = [v for (i, k), v in sorted(disp.items()) if "copyBuffer" in k]
        progs = [v for (i, k), v in sorted(disp.items()) if "k_program" in k and "noise" not in k]
        if not copies or not progs:
            raise RuntimeError(f"{counter}: calibration copy or k_program dispatch missing")
        kib_copy = copies[-1]
        kib_prog = progs[-1]
        scale = copy_bytes / (kib_copy * 1024.0)
        res[label] = {"counter_kib_k_program": kib_prog, "counter_kib_copy": kib_copy,
                      "scale": scale, "bytes": kib_prog * 1024.0 * scale}
    res["bytes_per_launch"] = res["fetch"]["bytes"] + res["write"]["bytes"]
    with open(out_json, "w") as f:
        json.dump(res, f, indent=1)
    md = [
        "| counter | k_program (KiB) | calibration copy (KiB) | scale (copy bytes / counter bytes) | k_program bytes |",
        "|---|---|---|---|---|",
    ]
    for label, counter in (("fetch", "FETCH_SIZE"), ("write", "WRITE_SIZE")):
        r = res[label]
        md.append("| %s | %.0f | %.0f | %.4f | %.4g |" % (counter, r["counter_kib_k_program"],
                                                        r["counter_kib_copy"], r["scale"], r["bytes"]))
    md.append("")
    md.append("traffic per k_program launch: %.4g bytes (calibration copy: %d bytes each way)"
              % (res["bytes_per_launch"], copy_bytes))
    with open(out_md, "w") as f:
        f.write("\n".join(md) + "\n")
    print("\n".join(md))


if __name__ == "__main__":
    cmd = sys.argv[1] if len(sys.argv) > 1 else "probe"
    if cmd == "probe":
        probe()
    elif cmd == "kernels":
        kernels(sys.argv[2], sys.argv[3])
    elif cmd == "traffic":
        traffic(*sys.argv[2:6])
    else:
        raise SystemExit(__doc__)
