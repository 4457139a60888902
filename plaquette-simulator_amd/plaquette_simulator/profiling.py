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

  python profiling.py sizes <read-size csv dir> <write-size csv dir> <out.json> <out.md>
      HBM traffic of k_program from the request counters by size, no
      calibration: reads 32 x RDREQ_32B + 64 x RDREQ_64B + 128 x RDREQ_128B,
      writes 64 x WRREQ_64B + 32 x (WRREQ - WRREQ_64B) (TCC_EA0_*, summed
      over channels), beside the same for the calibration copy.

  python profiling.py probe-n1024 [slices ...]
      The n=1024 random circuit (4096 shots, 100000 ops): a device-to-device
      copy of the whole tableau (calibration), then one persistent run per
      workgroups-per-word-group count (default 1 2).

  python profiling.py dispatches <out.md> <csv dir> [<csv dir> ...]
      Every counter of every `--pmc` pass for each k_program dispatch and the
      calibration copy, side by side.

  python profiling.py pmc <out.md> <csv dir> [<csv dir> ...]
      Every counter of every `--pmc` pass, for the last k_program dispatch and
      the calibration copy, side by side.
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


def probe_n1024(slices=(1, 2)):
    """The n=1024 random circuit (4096 shots, 100000 ops), one persistent run
    per workgroups-per-word-group count in `slices`, after a device-to-device
    copy of the whole 2.15 GB tableau (the calibration dispatch)."""
    sys.path.insert(0, os.path.dirname(HERE))
    import plaquette_simulator as ps
    from plaquette_simulator import workloads as W
    w = W.config_random_n1024()
    n, B = w.num_qubits, w.batch_size
    st = ps.CliffordState(n, B, w.seed, True)
    other = ps.CliffordState(n, B, w.seed, False)
    other.synchronize()
    other.device_to_device(st)              # calibration copy
    other.synchronize()
    del other
    prog = ps.Program(st, w.ops)
    for s in slices:
        prog.set_slices(s)
        st.reset_clifford_state()
        prog.run(st, "persistent")
        st.synchronize()
        used, failed = prog.slice_status(st)
        print(json.dumps({"workload": w.name, "slices": used, "barrier_failures": failed}),
              flush=True)
    words = (B + 63) // 64
    print(json.dumps({"probe": "done", "copy_bytes": words * (2 * n + 1) ** 2 * 8}), flush=True)


def dispatches(out, *dirs):
    """Every counter of every `--pmc` pass for each k_program dispatch (kernel
    build and grid size name it) and the calibration copy."""
    table = {}
    order = []
    for d in dirs:
        rs = _counter_rows(d)
        for r in rs:
            k = r["Kernel_Name"]
            if not ("k_program" in k and "noise" not in k) and "copyBuffer" not in k:
                continue
            who = ("calibration copy" if "copyBuffer" in k else
                   "k_program %s grid %s" % ("sliced" if "true>" in k else "one group per workgroup",
                                            r["Grid_Size"]))
            if who not in table:
                table[who] = {}
                order.append(who)
            table[who][r["Counter_Name"]] = table[who].get(r["Counter_Name"], 0.0) + float(r["Counter_Value"])
    names = sorted({c for v in table.values() for c in v})
    md = ["| counter | " + " | ".join(order) + " |", "|---" * (len(order) + 1) + "|"]
    for c in names:
        md.append("| %s | " % c + " | ".join("%.6g" % table[w].get(c, float("nan")) for w in order) + " |")
    with open(out, "w") as f:
        f.write("\n".join(md) + "\n")
    print("\n".join(md))


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
    return res


def sizes(rd_dir, wr_dir, out_json, out_md):
    def last(d, name):
        disp = _per_dispatch(_counter_rows(d), name)
        prog = [v for (i, k), v in sorted(disp.items()) if "k_program" in k and "noise" not in k]
        copy = [v for (i, k), v in sorted(disp.items()) if "copyBuffer" in k]
        return (prog[-1] if prog else 0.0), (copy[-1] if copy else 0.0)
    res = {}
    for who, idx in (("k_program", 0), ("calibration_copy", 1)):
        rd = {b: last(rd_dir, "TCC_EA0_RDREQ_%sB_sum" % b)[idx] for b in (32, 64, 128)}
        wr = last(wr_dir, "TCC_EA0_WRREQ_sum")[idx]
        wr64 = last(wr_dir, "TCC_EA0_WRREQ_64B_sum")[idx]
        read = 32 * rd[32] + 64 * rd[64] + 128 * rd[128]
        write = 64 * wr64 + 32 * (wr - wr64)
        res[who] = {"rdreq_32b": rd[32], "rdreq_64b": rd[64], "rdreq_128b": rd[128],
                    "wrreq": wr, "wrreq_64b": wr64, "read_bytes": read, "write_bytes": write,
                    "bytes": read + write}
    with open(out_json, "w") as f:
        json.dump(res, f, indent=1)
    md = ["| dispatch | RDREQ 32B | RDREQ 64B | RDREQ 128B | WRREQ | WRREQ 64B | read bytes | write bytes | total |",
          "|---|---|---|---|---|---|---|---|---|"]
    for who, r in res.items():
        md.append("| %s | %.6g | %.6g | %.6g | %.6g | %.6g | %.4g | %.4g | %.4g |" % (
            who, r["rdreq_32b"], r["rdreq_64b"], r["rdreq_128b"], r["wrreq"], r["wrreq_64b"],
            r["read_bytes"], r["write_bytes"], r["bytes"]))
    with open(out_md, "w") as f:
        f.write("\n".join(md) + "\n")
    print("\n".join(md))


def pmc(out, *dirs):
    rows = []
    for d in dirs:
        rs = _counter_rows(d)
        names = sorted({r["Counter_Name"] for r in rs})
        for name in names:
            disp = _per_dispatch(rs, name)
            prog = [v for (i, k), v in sorted(disp.items()) if "k_program" in k and "noise" not in k]
            copy = [v for (i, k), v in sorted(disp.items()) if "copyBuffer" in k]
            rows.append((name, prog[-1] if prog else float("nan"), copy[-1] if copy else float("nan")))
    md = ["| counter | k_program (last dispatch) | calibration copy |", "|---|---|---|"]
    md += ["| %s | %.6g | %.6g |" % r for r in rows]
    with open(out, "w") as f:
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
    elif cmd == "sizes":
        sizes(*sys.argv[2:6])
    elif cmd == "pmc":
        pmc(sys.argv[2], *sys.argv[3:])
    elif cmd == "probe-n1024":
        probe_n1024(tuple(int(a) for a in sys.argv[2:]) or (1, 2))
    elif cmd == "dispatches":
        dispatches(sys.argv[2], *sys.argv[3:])
    else:
        raise SystemExit(__doc__)
