#!/usr/bin/env python3
"""Benchmark of the batched stabilizer-tableau engine on one MI355X.

Workload: rotated surface-code memory-Z experiment, distance 9 (161 qubits),
9 rounds, 16384 shots, Pauli X/Y/Z noise p=1e-3 each after every gate layer,
seed 7 (3312 gates, 801 measurements, 5904 noise ops).  One step = reset the
tableau to |0...0> and run the whole circuit (synthetic, generated in
process).  The timed form is the batched persistent kernel; the plain form
(one kernel launch per op, as the reference issues them) and the hipGraph
replay of the persistent form are timed beside it.

Output: one JSON line (rank 0) with gate-shots/s as `value`, a `roofline`
object for the persistent kernel (algorithmic bytes per launch from the
engine's own work census / HIP-event kernel time) and a `cpu_baseline`
object (the CPU oracle, OpenMP over shots/rows, on a bounded shot sample).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "plaquette-simulator_amd"))

HBM_PEAK_GBS = 8000.0


def live_traffic(timeout_s=150):
    """HBM bytes of one k_program launch of the d=9 workload, measured now:
    two rocprofv3 --pmc runs of `profiling.py probe` (FETCH_SIZE, then
    WRITE_SIZE), each counter scaled by what it reads for the probe's
    device-to-device copy of the whole tableau (a known byte count).  Child
    processes, each under a time limit; returns (result dict, None) or
    (None, reason)."""
    import shutil
    import subprocess
    import tempfile
    prof = shutil.which("rocprofv3")
    if not prof:
        return None, "rocprofv3 not found"
    probe = os.path.join(ROOT, "plaquette-simulator_amd", "plaquette_simulator", "profiling.py")
    work = tempfile.mkdtemp(prefix="pq_traffic_", dir="/tmp")
    env = dict(os.environ, TMPDIR="/tmp")
    try:
        dirs = []
        for ctr in ("FETCH_SIZE", "WRITE_SIZE"):
            out = os.path.join(work, ctr.lower())
            progress(f"traffic: rocprofv3 --pmc {ctr}")
            try:
                r = subprocess.run([prof, "--pmc", ctr, "--output-format", "csv", "-d", out,
                                    "-o", "run", "--", sys.executable, probe, "probe"],
                                   cwd="/tmp", env=env, stdout=subprocess.DEVNULL,
                                   stderr=subprocess.DEVNULL, timeout=timeout_s)
            except subprocess.TimeoutExpired:
                return None, f"rocprofv3 {ctr} pass over {timeout_s} s"
            if r.returncode != 0:
                return None, f"rocprofv3 {ctr} pass exit {r.returncode}"
            dirs.append(out)
        import contextlib
        sys.path.insert(0, os.path.dirname(probe))
        import profiling
        with contextlib.redirect_stdout(sys.stderr):  # stdout carries the JSON line only
            res = profiling.traffic(dirs[0], dirs[1], os.path.join(work, "t.json"),
                                    os.path.join(work, "t.md"))
        return res, None
    except Exception as e:  # a missing counter file or the like: report, do not fail the line
        return None, f"{type(e).__name__}: {e}"
    finally:
        shutil.rmtree(work, ignore_errors=True)


def progress(msg):
    """One progress line on stderr (long modes print one per phase)."""
    print(f"[bench] {msg}", file=sys.stderr, flush=True)
METRIC = "gate-shots/s on d=9 surface code (%HBM peak); batched vs plain speedup"

# bytes per op per (row x word group), unfused, at bit-plane density
GATE_PLANES = {1: 6, 2: 5, 3: 3, 4: 4, 5: 3, 6: 8, 7: 8}


def algorithmic_bytes(ops, n, W, census):
    """Algorithmic HBM bytes of one run of `ops` on n qubits, W word groups.

    Gates: planes touched x R rows x W words x 8 B (H 6, S 5, CX/CZ 8,
    X/Z 3, Y 4).  Flagged H/S: as H/S plus the flag plane (8W).  Conditional
    Paulis and noise: read x_q, z_q, r and write r (4 planes) plus 8W per
    mask (1 for conditional, 2 for noise).  Measurements, per word group:
    random 8[(4+2K)R + 2], determined 8[2n + (K'+1)R + 2], K = rows rowsummed,
    K' = rowsum steps, summed over word groups by the engine's census."""
    import numpy as np
    R = 2 * n + 1
    codes = ops[:, 0]
    out = {"gates": 0, "flagged": 0, "noise": 0, "measurements": 0}
    for c, planes in GATE_PLANES.items():
        out["gates"] += int((codes == c).sum()) * planes * 8 * R * W
    out["flagged"] += int((codes == 8).sum()) * (48 * R * W + 8 * W)
    out["flagged"] += int((codes == 9).sum()) * (40 * R * W + 8 * W)
    out["flagged"] += int(np.isin(codes, (11, 12, 13)).sum()) * (32 * R * W + 8 * W)
    out["noise"] += int((codes == 10).sum()) * (32 * R * W + 16 * W)
    nr, nd = census["random"], census["determined"]
    kr, kd = census["random_rows"], census["determined_steps"]
    out["measurements"] = 8 * ((4 * R + 2) * nr + 2 * R * kr) + 8 * ((2 * n + R + 2) * nd + R * kd)
    out["total"] = sum(out.values())
    return out


# columns a gate reads and writes (x of qubit a is column a, z is n + a)
def _gate_columns(code, a, b, n):
    x, z = (lambda q: q), (lambda q: n + q)
    base = {8: 1, 9: 2, 11: 3, 12: 4, 13: 5, 15: 6, 16: 7}.get(code, code)
    if base == 1:
        return {x(a), z(a)}, {x(a), z(a)}
    if base == 2:
        return {x(a), z(a)}, {z(a)}
    if base == 3:
        return {z(a)}, set()
    if base in (4, 10):
        return {x(a), z(a)}, set()
    if base == 5:
        return {x(a)}, set()
    if base == 6:
        return {x(a), z(a), x(b), z(b)}, {x(b), z(a)}
    if base == 7:
        return {x(a), z(a), x(b), z(b)}, {z(a), z(b)}
    return set(), set()


def fused_minimum_bytes(ops, n, W, census):
    """Least HBM bytes one run of `ops` can move in the fused persistent form.

    k_program fuses each gate segment -- a maximal run of non-measurement
    ops -- into one pass over the tableau, so the model charges each segment
    one read of every column its ops read and one write of every column they
    change (R = 2n+1 words per column per 64-shot word group), one read and
    one write of the r column, and the per-group mask words it needs (16 B
    per noise op, 8 B per flag).  Measurements are charged their algorithmic
    bytes from the engine's census, as in `algorithmic_bytes`.  Noise planes
    are written by a separate prologue kernel and are not part of k_program."""
    R = 2 * n + 1
    col = R * W * 8
    out = {"segment_columns": 0, "masks": 0, "measurements": 0, "segments": 0}
    i, N = 0, len(ops)
    while i < N:
        if ops[i, 0] == 14:
            i += 1
            continue
        rd, wr, masks = set(), set(), 0
        while i < N and ops[i, 0] != 14:
            c, a, b = int(ops[i, 0]), int(ops[i, 1]), int(ops[i, 2])
            r_, w_ = _gate_columns(c, a, b, n)
            rd |= r_
            wr |= w_
            if c == 10:
                masks += 16 * W
            elif c in (8, 9, 11, 12, 13, 15, 16):
                masks += 8 * W
            i += 1
        out["segments"] += 1
        out["segment_columns"] += (len(rd) + len(wr) + 2) * col
        out["masks"] += masks
    nr, nd = census["random"], census["determined"]
    kr, kd = census["random_rows"], census["determined_steps"]
    out["measurements"] = 8 * ((4 * R + 2) * nr + 2 * R * kr) + 8 * ((2 * n + R + 2) * nd + R * kd)
    out["total"] = out["segment_columns"] + out["masks"] + out["measurements"]
    return out


def cpu_model():
    """The host CPU's model name (lscpu, else /proc/cpuinfo)."""
    import subprocess
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in out.splitlines():
            if line.startswith("Model name:"):
                return line.split(":", 1)[1].strip()
    except (OSError, subprocess.SubprocessError):
        pass
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def _oracle():
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    cores = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or (os.cpu_count() or 1)
    O.set_threads(cores)
    return O


def cpu_baseline(w, target_s, runs=3):
    """The CPU oracle (reference int32 layout [B][2n+1][n], OpenMP over
    (shot, row) for gates and over shots for measurements, as the Kokkos
    index spaces run) on the first S shots of the same circuit.  S is scaled
    from a 64-shot calibration run so one timed run takes about `target_s`
    seconds; shots are independent (coins and noise keyed by shot index), so
    S shots are an exact subset of the job and the rate scales by B/S.  As
    benchmarks/run_circuit.cpp:25-29 times it: circuit execution only (state
    set-up outside), median of `runs` runs, threads bound close
    (OMP_PROC_BIND, set before the OpenMP runtime starts)."""
    import numpy as np
    O = _oracle()
    ops32 = w.ops.astype(np.int32)
    cal = O.OracleState(w.num_qubits, 64, w.seed)
    t = time.perf_counter()
    cal.run_program(ops32)
    t64 = time.perf_counter() - t
    S = int(min(w.batch_size, max(64, 64 * round(target_s / max(t64, 1e-3)))))
    times = []
    for _ in range(runs):
        st = O.OracleState(w.num_qubits, S, w.seed)
        t = time.perf_counter()
        st.run_program(ops32)
        times.append(time.perf_counter() - t)
        del st
    dt = float(np.median(times))
    return {
        "value": w.gates * S / dt,
        "unit": "gate-shots/s",
        "cores": O.get_threads(),
        "kind": "port",
        "sample": f"{w.name}: full circuit on shots [0,{S}) of {w.batch_size} "
                  f"(rate scaled by {w.batch_size / S:.3g}), median of {runs} runs "
                  f"{', '.join(f'{x:.2f}' for x in times)} s, int32 [B][2n+1][n] layout",
        "meas_shots_per_s": w.measurements * S / dt,
        "cpu_model": cpu_model(),
        "omp_proc_bind": os.environ.get("OMP_PROC_BIND", ""),
        "runs_s": times,
        "shot_sample": [S, w.batch_size],
    }


def bench_mixed(args):
    """256 surface-code circuits, d uniform over odd {3..25}, rounds d,
    p=1e-3, shots log-uniform in [64, 4096], seed 11: all in one persistent,
    graph-captured batch launch over a packed (circuit, word group) work list,
    launch shape from the autotuner.  One step = reset every member + run.
    Beside it: the same circuits run one persistent launch at a time."""
    import numpy as np
    import plaquette_simulator as ps
    from plaquette_simulator import workloads as W

    t0 = time.perf_counter()
    ws = W.config_mixed_batch(count=args.mixed_count)
    states, progs = [], []
    batch = ps.Batch()
    for w in ws:
        st = ps.CliffordState(w.num_qubits, w.batch_size, w.seed, True)
        pr = ps.Program(st, w.ops)
        batch.add(st, pr)
        states.append(st)
        progs.append(pr)
    setup_s = time.perf_counter() - t0
    progress(f"mixed: {len(ws)} circuits built in {setup_s:.1f} s")
    gate_shots = sum(w.gates * w.batch_size for w in ws)
    meas_shots = sum(w.measurements * w.batch_size for w in ws)

    # sequential baseline with census (algorithmic bytes) then without
    abytes = 0
    fbytes = 0
    for w, st, pr in zip(ws, states, progs):
        pr.set_census(True)
        st.reset_clifford_state()
        pr.run(st, "persistent")
        c = pr.read_census()
        pr.set_census(False)
        words = (w.batch_size + 63) // 64
        abytes += algorithmic_bytes(w.ops, w.num_qubits, words, c)["total"]
        fbytes += fused_minimum_bytes(w.ops, w.num_qubits, words, c)["total"]
    for st in states:
        st.synchronize()
    progress("mixed: census done")
    t = time.perf_counter()
    for st, pr in zip(states, progs):
        st.reset_clifford_state()
        pr.run(st, "persistent")
        st.synchronize()
    seq_ms = 1e3 * (time.perf_counter() - t)
    progress(f"mixed: one launch per circuit {seq_ms:.0f} ms")

    choice = batch.autotune(1)
    progress(f"mixed: autotune {choice}")
    batch.run("graph")
    for _ in range(args.warmup):
        for st in states:
            st.reset_clifford_state()
        batch.run("graph")
    batch.set_timing(True)
    kms = []
    t = time.perf_counter()
    for i in range(args.steps):
        for st in states:
            st.reset_clifford_state()
        batch.run("graph")
        kms.append(batch.last_ms())
        if i % 5 == 4:
            progress(f"mixed: step {i + 1}/{args.steps}")
    wall = time.perf_counter() - t
    ms_step = 1e3 * wall / args.steps
    kernel_ms = float(np.mean(kms))
    info = batch.info()
    batch.set_timing(False)

    # frame form of the same batch (all members allow it): one noise
    # prologue, one frame kernel over the packed list, plans shared by the
    # circuits of one distance
    frame = {}
    try:
        t = time.perf_counter()
        for st in states:
            st.reset_clifford_state()
        batch.run("frame")
        batch.wait()
        frame["first_run_s"] = time.perf_counter() - t
        for _ in range(args.warmup):
            for st in states:
                st.reset_clifford_state()
            batch.run("frame")
        batch.wait()
        t = time.perf_counter()
        for _ in range(args.steps):
            for st in states:
                st.reset_clifford_state()
            batch.run("frame")
        batch.wait()
        fms = 1e3 * (time.perf_counter() - t) / args.steps
        frame.update({"ms_per_step": fms, "gate_shots_per_s": gate_shots / (fms / 1e3),
                      "speedup_vs_batch": ms_step / fms})
        progress(f"mixed: frame form {fms:.1f} ms/step")
    except RuntimeError as e:
        frame = {"error": str(e)}
    out = {
        "metric": "gate-shots/s, 256 mixed surface-code circuits in one batched launch",
        "value": gate_shots / (ms_step / 1e3),
        "unit": "gate-shots/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "dtype": "u64",
        "data": "synthetic",
        "config": {"workload": "mixed_surface_code_batch", "circuits": len(ws), "d": "3..25 odd",
                   "shots": "log-uniform [64, 4096]", "p": 1e-3, "seed": 11,
                   "word_groups": info["items"], "gate_shots": gate_shots,
                   "meas_shots": meas_shots,
                   "tableau_bytes": sum(((w.batch_size + 63) // 64) * (2 * w.num_qubits + 1) ** 2 * 8
                                        for w in ws)},
        "batch_kernel_ms": kernel_ms,
        "roofline": {"bound": "hbm", "achieved": fbytes / (kernel_ms / 1e3) / 1e9,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": fbytes / (kernel_ms / 1e3) / 1e9 / HBM_PEAK_GBS, "traffic": None,
                     "kernel": "k_batch", "kernel_ms": kernel_ms,
                     "model": "fused minimum: each gate segment reads every column it "
                              "uses and writes every column it changes once; "
                              "measurements by census",
                     "algorithmic_bytes": fbytes,
                     "per_op_model": {"bytes": abytes,
                                      "achieved": abytes / (kernel_ms / 1e3) / 1e9,
                                      "frac": abytes / (kernel_ms / 1e3) / 1e9 / HBM_PEAK_GBS}},
        "meas_shots_per_s": meas_shots / (ms_step / 1e3),
        "sequential_ms": seq_ms,
        "speedup_vs_sequential": seq_ms / ms_step,
        "autotune": choice,
        "launch": info,
        "frame_form": frame,
        "setup_s": setup_s,
    }
    print(json.dumps(out))


def bench_serving(args):
    """Poisson arrivals of surface-code sampling requests (d uniform over odd
    {3..21}, rounds d, shots log-uniform in [256, 65536], p uniform in
    [5e-4, 5e-3], seed 2024) at rates lambda in {50, 200, 800}/s, served by
    the engine's request loop with a 200 GB tableau budget, under both
    admission policies (arrival order; smallest estimated work first with an
    age limit).  Latency = arrival -> results on the host; throughput = all
    gate-shots / (last result - first arrival)."""
    import math
    import numpy as np
    import plaquette_simulator as ps
    from plaquette_simulator import workloads as W

    rng = np.random.default_rng(2024)
    budget = int(args.serving_budget_gb * 1e9)
    cache = {}
    streams = {}
    for lam in (50.0, 200.0, 800.0):
        reqs = []
        t_arr = 0.0
        for _ in range(args.serving_requests):
            t_arr += rng.exponential(1.0 / lam)
            d = int(rng.choice(list(range(3, 22, 2))))
            shots = int(round(math.exp(rng.uniform(math.log(256), math.log(65536)))))
            p = float(rng.uniform(5e-4, 5e-3))
            seed = int(rng.integers(0, 2 ** 31 - 1))
            key = (d, round(p, 9))
            if key not in cache:
                cache[key] = W.surface_code_memory_z(d, d, p)
            ops, n = cache[key]
            gates = int(np.isin(ops[:, 0], W.GATE_CODES).sum())
            reqs.append((t_arr, n, shots, seed, ops, gates))
        streams[lam] = reqs

    def serve(reqs, policy, form):
        sv = ps.Server(budget_bytes=budget, slots=args.serving_slots)
        sv.set_policy(policy, args.serving_age_limit if policy != "fcfs" else 0.0)
        sv.set_form(form)
        tickets = []
        t0 = time.perf_counter()
        for t_arr, n, shots, seed, ops, gates in reqs:
            now = time.perf_counter() - t0
            if t_arr > now:
                time.sleep(t_arr - now)
            tickets.append(sv.submit(n, shots, seed, ops))
        sv.drain()
        t_end = time.perf_counter() - t0
        lat = np.array([sv.status(t)[1] for t in tickets])
        queue = np.array([sv.status(t)[2] for t in tickets])
        stats = sv.stats()
        fstats = sv.frame_stats()
        for t in tickets:
            sv.release(t)
        del sv
        gate_shots = sum(g * s for _, _, s, _, _, g in reqs)
        span = t_end - reqs[0][0]  # first arrival -> last result
        return {"policy": policy, "form": form, "requests": len(reqs),
                "p50_latency_ms": float(np.percentile(lat, 50) * 1e3),
                "p99_latency_ms": float(np.percentile(lat, 99) * 1e3),
                "mean_latency_ms": float(lat.mean() * 1e3),
                "p50_queue_ms": float(np.percentile(queue, 50) * 1e3),
                "gate_shots_per_s": gate_shots / span,
                "gate_shots": gate_shots, "span_s": span,
                "offered_s": reqs[-1][0],
                "max_concurrent": stats["max_running"],
                "peak_bytes": stats["peak_bytes"],
                **({"frame": fstats} if form == "frame" else {})}

    lines = []
    for form in args.serving_forms.split(","):
        for lam, reqs in streams.items():
            for policy in ("fcfs", "smallest_first"):
                line = serve(reqs, policy, form)
                line["lambda_per_s"] = lam
                lines.append(line)
                print(f"[serving] {form} lambda {lam:g}/s {policy}: p50 {line['p50_latency_ms']:.0f} ms "
                      f"p99 {line['p99_latency_ms']:.0f} ms {line['gate_shots_per_s']:.3e} gate-shots/s",
                      file=sys.stderr, flush=True)

    # The first rate's requests once more, one at a time on one stream
    # (state allocation + persistent run + sync per request), and their
    # fused-minimum and per-op bytes from the engine's work census.
    first = streams[50.0]
    abytes = fbytes = 0
    for t_arr, n, shots, seed, ops, gates in first:
        st = ps.CliffordState(n, shots, seed, True)
        pr = ps.Program(st, ops)
        pr.set_census(True)
        pr.run(st, "persistent")
        census = pr.read_census()
        abytes += algorithmic_bytes(ops, n, (shots + 63) // 64, census)["total"]
        fbytes += fused_minimum_bytes(ops, n, (shots + 63) // 64, census)["total"]
        del pr, st
    t = time.perf_counter()
    for t_arr, n, shots, seed, ops, gates in first:
        st = ps.CliffordState(n, shots, seed, True)
        st.run_program(ops, "persistent")
        st.synchronize()
        del st
    alone_s = time.perf_counter() - t
    gs0 = lines[0]["gate_shots"]
    one_at_a_time = {"gate_shots_per_s": gs0 / alone_s, "seconds": alone_s,
                     "fused_minimum_bytes": fbytes, "per_op_bytes": abytes}
    for line in lines:
        if line["lambda_per_s"] == 50.0 and line["form"] == "persistent":
            # fused minimum (each gate segment's columns once) over the span,
            # and the per-op count beside it
            line["hbm_frac_served"] = fbytes / line["span_s"] / (HBM_PEAK_GBS * 1e9)
            line["per_op_frac_served"] = abytes / line["span_s"] / (HBM_PEAK_GBS * 1e9)
    out = {
        "metric": "serving: p50/p99 latency and aggregate gate-shots/s under Poisson load",
        "value": max((l["gate_shots_per_s"] for l in lines
                      if l["policy"] == "fcfs" and l["form"] == "persistent"),
                     default=max(l["gate_shots_per_s"] for l in lines)),
        "unit": "gate-shots/s",
        "n_gpus": 1,
        "higher_is_better": True,
        "dtype": "u64",
        "data": "synthetic",
        "config": {"workload": "surface_code_requests_poisson", "d": "3..21 odd",
                   "shots": "log-uniform [256, 65536]", "p": "U[5e-4, 5e-3]", "seed": 2024,
                   "budget_bytes": budget, "slots": args.serving_slots,
                   "requests_per_rate": args.serving_requests,
                   "age_limit_s": args.serving_age_limit},
        "rates": lines,
        "one_at_a_time_lambda50": one_at_a_time,
    }
    print(json.dumps(out))


def bench_sampling(args):
    """Many shots in one sampling state: a rotated surface-code memory-Z
    experiment (distance d, d rounds, p=1e-3, seed 7; d=25 and 2^20 shots by
    default) in a state that keeps only the phase column per 64-shot group,
    run in the frame form.  One step = reset + whole circuit.  The full
    bit-plane tableau of this job (W (2n+1)^2 words) would not fit one GPU."""
    import numpy as np
    import plaquette_simulator as ps
    from plaquette_simulator import workloads as W

    d, B = args.sampling_d, args.sampling_shots
    ops, n = W.surface_code_memory_z(d, d, 1e-3)
    words, R = (B + 63) // 64, 2 * n + 1
    gates = int(np.isin(ops[:, 0], W.GATE_CODES).sum())
    meas = int((ops[:, 0] == W.OP_M).sum())
    nnoise = int((ops[:, 0] == W.OP_NOISE).sum())
    st = ps.SamplingState(n, B, 7)
    prog = ps.Program(st, ops)
    t0 = time.perf_counter()
    st.reset_clifford_state()
    prog.run(st, "frame")  # builds the plan on the host, then runs
    st.synchronize()
    first = time.perf_counter() - t0
    progress(f"sampling: d={d} {B} shots, first run {first:.2f} s")
    steps = max(1, min(args.steps, 3))
    prog.kernel_ms()
    prog.set_timing(True)
    t0 = time.perf_counter()
    for _ in range(steps):
        st.reset_clifford_state()
        prog.run(st, "frame")
    st.synchronize()
    ms_step = 1e3 * (time.perf_counter() - t0) / steps
    kms, launches = prog.kernel_ms()
    prog.set_timing(False)
    progress(f"sampling: {ms_step:.1f} ms/step")
    out = {
        "metric": f"gate-shots/s, d={d} surface-code memory, {B} shots in one sampling state",
        "value": gates * B / (ms_step / 1e3),
        "unit": "gate-shots/s",
        "n_gpus": 1, "steps": steps, "warmup": 1, "ms_per_step": ms_step,
        "higher_is_better": True, "dtype": "u64", "data": "synthetic",
        "config": {"workload": f"surface_code_memory_z_d{d}_sampling", "qubits": n,
                   "rounds": d, "shots": B, "p": 1e-3, "seed": 7, "gates": gates,
                   "measurements": meas, "noise_ops": nnoise, "form": "frame",
                   "state": "phase column only"},
        "shots_per_s": B / (ms_step / 1e3),
        "meas_shots_per_s": meas * B / (ms_step / 1e3),
        # HIP events around the run's kernels on the state's stream: with
        # more word groups than the GPU holds at once they span the noise
        # kernel and k_frame pipelined slice by slice
        "kernel": "k_frame_noise + k_frame (pipelined slices)", "kernel_ms": kms / max(1, launches),
        "first_run_s": first,
        "device_bytes": {"phase_columns": words * R * 8,
                         "records": meas * 2 * words * 8,
                         "noise_planes": nnoise * 2 * words * 8},
        "full_tableau_bytes": words * R * R * 8,
    }
    print(json.dumps(out))


def cpu_baseline_prefix(w, target_s, shots=64):
    """The CPU oracle on `shots` shots and a prefix of the op list, the prefix
    scaled from a 1000-op calibration run to about `target_s` seconds (the
    ops are uniform random, so a prefix has the circuit's mix)."""
    import numpy as np
    O = _oracle()
    ops32 = w.ops.astype(np.int32)
    cal = O.OracleState(w.num_qubits, shots, w.seed)
    k0 = min(1000, len(ops32))
    t = time.perf_counter()
    cal.run_program(ops32[:k0])
    t0 = time.perf_counter() - t
    k = int(min(len(ops32), max(k0, k0 * target_s / max(t0, 1e-3))))
    st = O.OracleState(w.num_qubits, shots, w.seed)
    t = time.perf_counter()
    st.run_program(ops32[:k])
    dt = time.perf_counter() - t
    codes = ops32[:k, 0]
    gates = int(np.isin(codes, (1, 2, 3, 4, 5, 6, 7)).sum())
    meas = int((codes == 14).sum())
    return {"value": gates * shots / dt, "unit": "gate-shots/s", "cores": O.get_threads(),
            "kind": "port", "cpu_model": cpu_model(),
            "omp_proc_bind": os.environ.get("OMP_PROC_BIND", ""),
            "sample": f"{w.name}: first {k} ops ({gates} gates, {meas} measurements) on "
                      f"shots [0,{shots}), {dt:.2f} s, int32 [B][2n+1][n] layout",
            "meas_shots_per_s": meas * shots / dt}


def bench_random_n1024(args):
    """Random Clifford circuit on 1024 qubits, 4096 shots, 100000 ops (10%
    measurements of uniform qubits, else H/S/CX on uniform qubits), seed 42:
    the measurement-bound shape.  One step = reset + whole circuit in the
    persistent form; census run first for the algorithmic bytes."""
    import numpy as np
    import plaquette_simulator as ps
    from plaquette_simulator import workloads as W

    w = W.config_random_n1024()
    n, B = w.num_qubits, w.batch_size
    words = (B + 63) // 64
    st = ps.CliffordState(n, B, w.seed, True)
    prog = ps.Program(st, w.ops)
    # workgroups per word group (0: chosen for the device; 64 word groups
    # cannot fill 256 compute units one workgroup each)
    prog.set_slices(args.slices)
    prog.set_census(True)
    st.reset_clifford_state()
    prog.run(st, "persistent")
    census = prog.read_census()
    prog.set_census(False)
    slices, failed = prog.slice_status(st)
    if failed:
        raise RuntimeError("a barrier between the slices of a word group gave up")
    progress(f"n1024: census run done ({slices} workgroups per word group)")
    abytes = algorithmic_bytes(w.ops, n, words, census)
    fbytes = fused_minimum_bytes(w.ops, n, words, census)
    steps = max(1, min(args.steps, 3))
    prog.kernel_ms()
    prog.set_timing(True)
    t0 = time.perf_counter()
    for _ in range(steps):
        st.reset_clifford_state()
        prog.run(st, "persistent")
    st.synchronize()
    wall = time.perf_counter() - t0
    kms, launches = prog.kernel_ms()
    kernel_ms = kms / max(1, launches)
    ms_step = 1e3 * wall / steps
    achieved = fbytes["total"] / (kernel_ms / 1e3) / 1e9
    per_op_gbs = abytes["total"] / (kernel_ms / 1e3) / 1e9
    progress(f"n1024: persistent {ms_step:.1f} ms/step")
    # frame form (no flagged ops here, so X/Z are the same in every shot)
    frame = {}
    try:
        t0 = time.perf_counter()
        st.reset_clifford_state()
        prog.run(st, "frame")  # compiles the plan on the host, then runs
        st.synchronize()
        frame["first_run_s"] = time.perf_counter() - t0
        prog.kernel_ms()
        prog.set_timing(True)
        t0 = time.perf_counter()
        for _ in range(steps):
            st.reset_clifford_state()
            prog.run(st, "frame")
        st.synchronize()
        fms = 1e3 * (time.perf_counter() - t0) / steps
        fk, fl = prog.kernel_ms()
        prog.set_timing(False)
        frame.update({"ms_per_step": fms, "gate_shots_per_s": w.gates * B / (fms / 1e3),
                      "meas_shots_per_s": w.measurements * B / (fms / 1e3),
                      "kernel": "k_frame", "kernel_ms": fk / max(1, fl),
                      "speedup_vs_persistent": ms_step / fms})
        progress(f"n1024: frame {fms:.2f} ms/step")
    except RuntimeError as e:
        frame = {"error": str(e)}
    out = {
        "metric": "gate-shots/s and measurement-shots/s, random Clifford circuit n=1024",
        "value": w.gates * B / (ms_step / 1e3),
        "unit": "gate-shots/s",
        "n_gpus": 1, "steps": steps, "warmup": 0, "ms_per_step": ms_step,
        "higher_is_better": True, "dtype": "u64", "data": "synthetic",
        "config": {"workload": "random_clifford_n1024", "qubits": n, "shots": B,
                   "ops": int(len(w.ops)), "gates": w.gates, "measurements": w.measurements,
                   "seed": w.seed, "word_groups": words, "slices": slices},
        "meas_shots_per_s": w.measurements * B / (ms_step / 1e3),
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None, "kernel": "k_program",
                     "kernel_ms": kernel_ms,
                     "model": "fused minimum: each gate segment reads every column it "
                              "uses and writes every column it changes once; "
                              "measurements by census",
                     "algorithmic_bytes": fbytes["total"],
                     "per_op_model": {"bytes": abytes["total"], "achieved": per_op_gbs,
                                      "frac": per_op_gbs / HBM_PEAK_GBS}},
        "algorithmic_bytes": abytes,
        "fused_minimum_bytes": fbytes,
        "census": census,
        "frame_form": frame,
    }
    # measured HBM bytes of one k_program launch of this workload (rocprofv3
    # request counters by size, one workgroup per word group)
    traffic_file = os.path.join(ROOT, "profiles", "k_program_random_n1024_traffic.json")
    if os.path.exists(traffic_file):
        tr = json.load(open(traffic_file))
        rf = out["roofline"]
        rf["traffic"] = tr.get("bytes_per_launch")
        rf["traffic_source"] = ("profiles/k_program_random_n1024_traffic.json: " +
                                tr.get("method", "rocprofv3 counters"))
        if tr.get("bytes_per_launch"):
            gbs = tr["bytes_per_launch"] / (kernel_ms / 1e3) / 1e9
            rf["traffic_gbs"] = gbs
            rf["traffic_frac"] = gbs / HBM_PEAK_GBS
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline_prefix(w, args.cpu_seconds)
    print(json.dumps(out))


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--plain-steps", type=int, default=2)
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-traffic", action="store_true",
                    help="skip the rocprofv3 passes that measure the headline kernel's HBM bytes")
    ap.add_argument("--shots", type=int, default=16384)
    ap.add_argument("--threads", type=int, default=0, help="persistent kernel block size")
    ap.add_argument("--slices", type=int, default=0,
                    help="workgroups per 64-shot word group (0: chosen for the device)")
    ap.add_argument("--mixed", action="store_true",
                    help="measure the 256-circuit mixed batch instead of the d=9 workload")
    ap.add_argument("--mixed-count", type=int, default=256)
    ap.add_argument("--random-n1024", action="store_true",
                    help="measure the n=1024 random Clifford circuit instead")
    ap.add_argument("--serving", action="store_true",
                    help="measure the serving loop under Poisson arrivals instead")
    ap.add_argument("--serving-requests", type=int, default=100)
    ap.add_argument("--serving-slots", type=int, default=8)
    ap.add_argument("--serving-budget-gb", type=float, default=200.0)
    ap.add_argument("--serving-age-limit", type=float, default=2.0,
                    help="age limit (s) of the smallest-first admission policy")
    ap.add_argument("--serving-forms", default="persistent,frame",
                    help="execution forms to serve with (comma-separated)")
    ap.add_argument("--sampling", action="store_true",
                    help="measure many shots of one surface code in a sampling state instead")
    ap.add_argument("--sampling-d", type=int, default=25)
    ap.add_argument("--sampling-shots", type=int, default=1 << 20)
    args = ap.parse_args()
    # the CPU baseline's OpenMP threads stay on neighbouring cores; the
    # runtime reads this when it starts, so it is set before anything loads it
    os.environ.setdefault("OMP_PROC_BIND", "close")
    if args.gpus != 1:
        print(f"bench.py measures one GPU: this engine runs on a single MI355X "
              f"(--gpus {args.gpus} is not supported; the path would shard on shots, each "
              f"GPU a range of 64-shot word groups, with no exchange until the records "
              f"are gathered)", file=sys.stderr)
        sys.exit(2)

    if args.mixed:
        bench_mixed(args)
        return
    if args.serving:
        bench_serving(args)
        return
    if args.random_n1024:
        bench_random_n1024(args)
        return
    if args.sampling:
        bench_sampling(args)
        return

    import plaquette_simulator as ps
    from plaquette_simulator import workloads as W

    w = W.config_surface_d9(shots=args.shots)
    n, B = w.num_qubits, w.batch_size
    words = (B + 63) // 64
    st = ps.CliffordState(n, B, w.seed, True)
    prog = ps.Program(st, w.ops)
    if args.threads:
        prog.set_launch(args.threads, 0)

    # work census (untimed run)
    prog.set_census(True)
    st.reset_clifford_state()
    prog.run(st, "persistent")
    census = prog.read_census()
    prog.set_census(False)
    abytes = algorithmic_bytes(w.ops, n, words, census)

    # persistent form: warmup, then K timed steps with HIP events per kernel
    for _ in range(args.warmup):
        st.reset_clifford_state()
        prog.run(st, "persistent")
    st.synchronize()
    prog.kernel_ms()
    prog.set_timing(True)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        st.reset_clifford_state()
        prog.run(st, "persistent")
    st.synchronize()
    wall = time.perf_counter() - t0
    kms, launches = prog.kernel_ms()
    prog.set_timing(False)
    ms_step = 1e3 * wall / args.steps
    kernel_ms = kms / max(1, launches)

    # graph form (same kernels, captured once, replayed)
    prog.run(st, "graph")
    st.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        st.reset_clifford_state()
        prog.run(st, "graph")
    st.synchronize()
    graph_ms = 1e3 * (time.perf_counter() - t0) / args.steps

    # frame form: X/Z are the same in every shot here (no flagged H/S/CX/CZ),
    # so they are simulated once on the host and only the phase column runs
    # on the GPU; bit-identical results, reported beside the roofline metric
    frame = {}
    try:
        t0 = time.perf_counter()
        st.reset_clifford_state()
        prog.run(st, "frame")  # builds and uploads the plan
        st.synchronize()
        frame["plan_build_s"] = time.perf_counter() - t0
        for _ in range(args.warmup):
            st.reset_clifford_state()
            prog.run(st, "frame")
        st.synchronize()
        # wall clock without timing events (at 0.2 ms a step, their markers
        # show), then the kernels' span from a second pass with them
        t0 = time.perf_counter()
        for _ in range(args.steps):
            st.reset_clifford_state()
            prog.run(st, "frame")
        st.synchronize()
        fwall = time.perf_counter() - t0
        prog.kernel_ms()
        prog.set_timing(True)
        for _ in range(args.steps):
            st.reset_clifford_state()
            prog.run(st, "frame")
        st.synchronize()
        fk, fl = prog.kernel_ms()
        prog.set_timing(False)
        frame_ms = 1e3 * fwall / args.steps
        # the same steps with the shared X/Z written into every word group, so
        # the tableau ends as the persistent form leaves it: a checksum of the
        # first 64 shots reads the tableau, which writes X/Z (k_frame_xz) first,
        # and waits for it
        t0 = time.perf_counter()
        for _ in range(args.steps):
            st.reset_clifford_state()
            prog.run(st, "frame")
            st.check_sum(0, 64)
        fxz_ms = 1e3 * (time.perf_counter() - t0) / args.steps
        frame.update({"ms_per_step": frame_ms,
                      "step": "records only: reset + noise + k_frame; the shared X/Z are "
                              "written when the tableau is next read",
                      "gate_shots_per_s": w.gates * B / (frame_ms / 1e3),
                      "kernel": "k_frame", "kernel_ms": fk / max(1, fl),
                      "speedup_vs_persistent": ms_step / frame_ms,
                      "ms_per_step_with_tableau": fxz_ms,
                      "speedup_vs_persistent_with_tableau": ms_step / fxz_ms})
        progress(f"frame form {frame_ms:.3f} ms/step")
    except RuntimeError as e:
        frame = {"error": str(e)}

    # plain form: one kernel launch per op
    st.reset_clifford_state()
    prog.run(st, "plain")
    st.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.plain_steps):
        st.reset_clifford_state()
        prog.run(st, "plain")
    st.synchronize()
    plain_ms = 1e3 * (time.perf_counter() - t0) / args.plain_steps

    gate_shots = w.gates * B
    value = gate_shots / (ms_step / 1e3)
    fbytes = fused_minimum_bytes(w.ops, n, words, census)
    achieved = fbytes["total"] / (kernel_ms / 1e3) / 1e9
    per_op_gbs = abytes["total"] / (kernel_ms / 1e3) / 1e9
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "gate-shots/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u64",
        "data": "synthetic",
        "config": {"workload": "surface_code_memory_z_d9", "qubits": n, "rounds": 9,
                   "shots": B, "p": 1e-3, "seed": w.seed, "gates": w.gates,
                   "measurements": w.measurements, "noise_ops": w.noise_ops,
                   "form": "persistent"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "kernel": "k_program", "kernel_ms": kernel_ms,
                     "model": "fused minimum: each gate segment reads every column it "
                              "uses and writes every column it changes once; "
                              "measurements by census",
                     "algorithmic_bytes": fbytes["total"],
                     # the per-op, unfused count (every op's columns from HBM),
                     # labelled apart: a fused kernel moves far less than this
                     "per_op_model": {"bytes": abytes["total"], "achieved": per_op_gbs,
                                      "frac": per_op_gbs / HBM_PEAK_GBS}},
        "meas_shots_per_s": w.measurements * B / (ms_step / 1e3),
        "hbm_frac_step": fbytes["total"] / (ms_step / 1e3) / (HBM_PEAK_GBS * 1e9),
        "fused_minimum_bytes": fbytes,
        "graph_ms_per_step": graph_ms,
        "graph_gate_shots_per_s": gate_shots / (graph_ms / 1e3),
        "plain_ms_per_step": plain_ms,
        "plain_gate_shots_per_s": gate_shots / (plain_ms / 1e3),
        "speedup_vs_plain": plain_ms / ms_step,
        "frame_form": frame,
        "algorithmic_bytes": abytes,
        "census": census,
    }
    # measured HBM bytes of one k_program launch of this workload: counted
    # now by rocprofv3 in child processes (counters cannot be read inside the
    # timed run), else the committed request-size count of the same kernel
    live, why = (None, "--no-traffic") if args.no_traffic else live_traffic()
    rf = out["roofline"]
    if live:
        rf["traffic"] = live["bytes_per_launch"]
        rf["traffic_source"] = ("measured in this run: rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE "
                                "of profiling.py probe, each scaled by the same counter for a "
                                "device-to-device copy of the tableau")
        rf["traffic_detail"] = live
    else:
        traffic_file = os.path.join(ROOT, "profiles", "k_program_surface_d9_traffic.json")
        if os.path.exists(traffic_file):
            tr = json.load(open(traffic_file))
            rf["traffic"] = tr.get("bytes_per_launch")
            rf["traffic_source"] = ("profiles/k_program_surface_d9_traffic.json: " +
                                    tr.get("method", "rocprofv3 counters") +
                                    f" (not measured in this run: {why})")
    if rf.get("traffic"):
        # measured HBM bytes (same kernel and workload) per launch time
        gbs = rf["traffic"] / (kernel_ms / 1e3) / 1e9
        rf["traffic_gbs"] = gbs
        rf["traffic_frac"] = gbs / HBM_PEAK_GBS
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(w, args.cpu_seconds)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
