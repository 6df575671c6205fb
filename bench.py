"""Benchmark: 30-qubit complex128 random circuit on one MI355X.

Workload (the configuration the headline metric is quoted on): a 30-qubit
complex128 ket (16 GiB), depth-40 random circuit -- per layer a Haar 1-qubit
gate on every qubit, then CNOTs on a random perfect pairing -- 1200 dense
1-qubit gates + 600 CNOTs.  One step applies the whole circuit to the
resident state (inputs already in HBM).

Metric: achieved HBM GB/s of the gate-application passes (sum of algorithmic
bytes / wall time; a dense 1-qubit pass reads and writes every amplitude
once, 2*16*2^30 B, a CNOT pass touches the control-satisfying half,
2*16*2^29 B) and its fraction of the 8 TB/s HBM3E peak.  Gate applications
per second are reported beside it.

Modes:
  --mode fused  (default) gates grouped into tile passes by the engine's
                scheduler; one pass streams the state once
  --mode gates  gate-by-gate through qs_apply / qs_apply_ctrl

The roofline object measures the dominant kernel live with HIP events on
the library's stream; cpu_baseline times the oracle (qpp 7.0.3's qubit
kernels restated in C++/OpenMP) on the host for a bounded sample of the
same circuit.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6  # MI355X dense FP64 (vector and matrix)
METRIC = "30q gate-apply GB/s (% HBM peak); served circuits/s + p99 at FP64-oracle tol"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def profile_region(qa, fn):
    """Run fn() between two HIP events on the library stream; returns ms."""
    import ctypes
    L = qa.lib()
    L.qs_profile_begin.restype = ctypes.c_int
    L.qs_profile_end.argtypes = [ctypes.POINTER(ctypes.c_double)]
    qa._check(L.qs_profile_begin())
    fn()
    ms = ctypes.c_double()
    qa._check(L.qs_profile_end(ctypes.byref(ms)))
    return ms.value


def cpu_model():
    """The host CPU's model name (the first 'model name' of /proc/cpuinfo)."""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(gates, n, picks):
    """Oracle (qpp qubit kernels restated, OpenMP) on the gates `picks`
    (indices into the circuit) of the same circuit, applied in that order
    to an n-qubit state; returns per-gate times and the circuit time
    extrapolated from the mean time per gate kind (flagged as such)."""
    import ctypes
    from ctypes import POINTER, c_double, c_uint32
    import oracle
    L = oracle.lib()
    L.orc_time_gates.argtypes = [c_uint32, c_uint32, POINTER(c_uint32), POINTER(c_uint32),
                                 POINTER(c_uint32), POINTER(c_double), POINTER(c_double)]
    L.orc_time_gates.restype = ctypes.c_int
    g = [gates[i] for i in picks]
    kind = np.array([0 if x[0] == "u" else 1 for x in g], dtype=np.uint32)
    tg = np.array([x[1][0] for x in g], dtype=np.uint32)
    ct = np.array([x[2][0] if x[2] else 0 for x in g], dtype=np.uint32)
    m = np.ascontiguousarray(np.concatenate(
        [np.asfortranarray(x[3]).reshape(-1, order="F") for x in g]).astype(np.complex128))
    sec = np.zeros(len(g))
    threads = L.orc_time_gates(n, len(g), kind.ctypes.data_as(POINTER(c_uint32)),
                               tg.ctypes.data_as(POINTER(c_uint32)),
                               ct.ctypes.data_as(POINTER(c_uint32)),
                               m.ctypes.data_as(POINTER(c_double)),
                               sec.ctypes.data_as(POINTER(c_double)))
    from qpp_amd.workloads import gate_bytes
    nbytes = sum(gate_bytes(x, n) for x in g)
    total = float(sec.sum())
    n1, n2 = int((kind == 0).sum()), int((kind == 1).sum())
    t1 = float(sec[kind == 0].mean()) if n1 else 0.0
    t2 = float(sec[kind == 1].mean()) if n2 else 0.0
    c1 = sum(1 for x in gates if x[0] == "u")
    c2 = len(gates) - c1
    full = len(picks) == len(gates)
    est = total if full else c1 * t1 + c2 * t2
    return {
        "value": round(nbytes / total / 1e9, 3),
        "unit": "GB/s",
        "cores": int(threads),
        "cpu_model": cpu_model(),
        "kind": "port",
        "sample": (f"the whole {n}-qubit circuit ({n1} dense 1q + {n2} CNOT gates): {total:.2f} s"
                   if full else
                   f"{len(g)} gates of the same {n}-qubit circuit ({n1} dense 1q at qubits "
                   f"{[x[1][0] for x in g if x[0] == 'u']}, {n2} CNOT), {total:.2f} s: "
                   f"{t1:.3f} s per dense gate, {t2:.3f} s per CNOT")
                  + "; qpp 7.0.3 qubit kernels restated (apply_psi_1q / apply_ctrl_psi_1q, "
                    "out-of-place, OpenMP)",
        "circuit_s": round(est, 2),
        "circuit_s_extrapolated": not full,
        "circuit_s_definition": ("measured" if full else
                                 f"EXTRAPOLATED: {c1} x mean dense-gate time + {c2} x mean "
                                 f"CNOT time from the sample above"),
    }


def oracle_threads():
    """OpenMP threads the oracle runs with (orc_time_gates with no gates
    only reports them)."""
    import ctypes
    from ctypes import POINTER, c_double, c_uint32
    import oracle
    L = oracle.lib()
    L.orc_time_gates.argtypes = [c_uint32, c_uint32, POINTER(c_uint32), POINTER(c_uint32),
                                 POINTER(c_uint32), POINTER(c_double), POINTER(c_double)]
    L.orc_time_gates.restype = ctypes.c_int
    return int(L.orc_time_gates(1, 0, None, None, None, None, None))


def run_gates(qa, st, gates):
    from qpp_amd.workloads import apply_gates
    apply_gates(st, gates)


def bench_batch(args, qa):
    """Batched noisy trajectories: 4096 x 16 qubits + 512 x {12, 20, 22}
    qubits, depth 30 (Haar 1q on every qubit + CNOT ladder), depolarizing
    p = 1e-3 before every step on every qubit (QNoisyEngine), measurement of
    every qubit; per run the 20- and 22-qubit members' passes (one launch
    per pass and circuit, a workgroup per tile) and one persistent launch
    for the rest and every measurement."""
    from qpp_amd import workloads
    sizes = [(16, 4096), (12, 512), (20, 512), (22, 512)]
    scale = args.traj_scale
    depth = args.traj_depth
    p = 1e-3
    dt = qa.C64 if args.dtype == "c64" else qa.C128
    circs = [workloads.trajectory_circuit(qa, n, depth, 300 + n) for n, _ in sizes]
    cof = np.concatenate([np.full(max(1, int(c * scale)), i, dtype=np.uint32)
                          for i, (_, c) in enumerate(sizes)])
    seeds = np.arange(len(cof), dtype=np.uint64) + 1000
    ntraj = len(cof)
    es = 8 if dt == qa.C64 else 16
    b = qa.Batch(circs, dtype=dt, low_bits=args.traj_low_bits, max_ops=args.traj_max_ops)
    passes = [b.info(i)["passes"] for i in range(len(circs))]
    algo_bytes = float(sum(passes[c] * 2 * es * 2 ** sizes[c][0] for c in cof))
    for w in range(args.warmup):
        b.run(cof, seeds, p)
        log(f"warmup {w + 1}/{args.warmup}: {b.info(0)}")
    t0 = time.perf_counter()
    dev_ms = []
    for _ in range(args.steps):
        idx, prob = b.run(cof, seeds, p)
        dev_ms.append(b.info(0)["device_ms"])
    wall = (time.perf_counter() - t0) / args.steps
    inf = b.info(0)
    log(f"batch: {ntraj} trajectories, wall {wall * 1e3:.1f} ms/run, kernel {np.mean(dev_ms):.1f} ms, "
        f"host plan {inf['host_ms']:.1f} ms, exact replays {inf['exact_replays']}")

    # the plain GPU path on a sample: one QNoisyEngine per trajectory (c128)
    plain = {}
    for i, (n, c) in enumerate(sizes):
        k = 2 if n >= 20 else 4
        tt = time.perf_counter()
        for j in range(k):
            qa.QEngine(circs[i], noise_p=p, seed=int(seeds[np.nonzero(cof == i)[0][j]])).execute(1)
        qa.synchronize()
        plain[n] = (time.perf_counter() - tt) / k
        log(f"plain engine {n}q: {plain[n] * 1e3:.1f} ms/trajectory")
    plain_sec = sum(plain[sizes[c][0]] for c in cof)
    # precision: the noiseless circuits in c64 vs c128 (FP64 comparison)
    prec = {}
    for i, (n, _) in enumerate(sizes):
        gates = workloads.ladder_layers(n, depth, 300 + n)
        s64 = qa.State([2] * n)
        qa.Program(n, gates).run(s64)
        s32 = qa.State([2] * n, dtype=qa.C64)
        qa.Program(n, gates, dtype=qa.C64).run(s32)
        r = s64.compare(s32)
        prec[f"{n}q"] = {"infidelity": r["infidelity"], "trace_distance": r["trace_distance"],
                         "max_abs": r["max_abs"]}
    # outcome agreement with the c128 batch on the same seeds
    sub = min(ntraj, 1024)
    ref_idx, _ = qa.Batch(circs, dtype=qa.C128, low_bits=args.traj_low_bits,
                          max_ops=args.traj_max_ops).run(cof[:sub], seeds[:sub], p)
    agree = float(np.mean(ref_idx == idx[:sub]))
    # oracle QNoisyEngine on a few 12-qubit trajectories (same seeds)
    import oracle
    i12 = [j for j in range(ntraj) if cof[j] == 1][: args.oracle_traj]
    t_c = time.perf_counter()
    om = sum(1 for j in i12
             if tuple(oracle.engine_run(circs[1], 1, p, int(seeds[j]))[1]) == b.dits(1, idx[j]))
    t12 = time.perf_counter() - t_c
    # the oracle's noisy engine (qpp's loop, restated) on 16-qubit trajectories
    # of the batch: outcome parity and the CPU baseline
    i16 = [j for j in range(ntraj) if cof[j] == 0][: args.oracle_traj16]
    t_c = time.perf_counter()
    om16 = sum(1 for j in i16
               if tuple(oracle.engine_run(circs[0], 1, p, int(seeds[j]))[1]) == b.dits(0, idx[j]))
    t16 = time.perf_counter() - t_c
    cpu = None
    if i16:
        cpu = {"value": round(len(i16) / t16, 3), "unit": "trajectories/s (16 qubits)",
               "cores": oracle_threads(), "kind": "port",
               "sample": f"{len(i16)} 16-qubit and {len(i12)} 12-qubit trajectories of the same "
                         f"batch (same seeds) through the oracle's QNoisyEngine (qpp's execute "
                         f"loop and qubit kernels restated): {t16:.2f} s and {t12:.2f} s; "
                         f"outcomes equal {om16}/{len(i16)} and {om}/{len(i12)}"}
    achieved = algo_bytes / (np.mean(dev_ms) / 1e3) / 1e9
    return {
        "metric": "batched noisy trajectories/s (mixed 16/12/20/22 qubits)",
        "value": round(ntraj / wall, 2),
        "unit": "trajectories/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(wall * 1e3, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "complex64" if dt == qa.C64 else "complex128",
        "data": "synthetic",
        "config": {"workload": f"{ntraj} QNoisyEngine trajectories "
                               f"({', '.join(f'{max(1, int(c * scale))}x{n}q' for n, c in sizes)}), "
                               f"depth {depth}, depolarizing p={p}, measure all",
                   "passes_per_circuit": dict(zip([f"{n}q" for n, _ in sizes], passes)),
                   "fuse_opts": {"low_bits": args.traj_low_bits, "max_ops": args.traj_max_ops},
                   "circuit_sha256": {f"{n}q": workloads.gates_digest(
                       workloads.ladder_layers(n, depth, 300 + n)) for n, _ in sizes}},
        "plain_gpu_trajectories_per_s": round(ntraj / plain_sec, 3),
        "speedup_vs_plain": round(plain_sec / wall, 2),
        "host_plan_ms": round(inf["host_ms"], 2),
        "kernel_ms": round(float(np.mean(dev_ms)), 3),
        "exact_replays": inf["exact_replays"],
        "precision_noiseless_c64_vs_c128": prec,
        "outcomes_equal_c128_batch": agree,
        "oracle_outcomes_matched": f"{om}/{len(i12)} (12-qubit), {om16}/{len(i16)} (16-qubit)",
        "roofline": {"bound": "hbm",
                     "kernel": "k_traj_pass (20- and 22-qubit members, pass by pass, a workgroup per "
                               "tile) + k_traj (persistent: 12- and 16-qubit members, every measurement)",
                     "achieved": round(achieved, 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                     "avg_launch_ms": round(float(np.mean(dev_ms)), 3),
                     "timing": "HIP events around every launch of one batch run"},
        "cpu_baseline": cpu,
    }


def bench_serve(args, qa):
    """Serving: Poisson arrivals of random circuits (n ~ U{10..26}, depth ~
    U{10..100}, layers as the plain workload, one measurement of all qubits),
    1024 shots each (QEngine.execute(1024)), dtype complex128/complex64 50/50,
    admitted by HBM footprint; rate = 0.8 x the capacity measured by a
    closed-loop calibration batch."""
    from qpp_amd import workloads
    rng = np.random.default_rng(args.seed)
    nreq = args.requests

    def make(k, r):
        q = workloads.serve_request(k, r)
        q["dtype"] = qa.C64 if q["c64"] else qa.C128
        return q

    reqs = [make(k, rng) for k in range(nreq)]
    circs = [workloads.measured_circuit(qa, q["n"], q["depth"], q["gseed"]) for q in reqs]
    srv = qa.Server(lanes=args.lanes)
    # closed-loop calibration on a separate sample of the same distribution
    crng = np.random.default_rng(args.seed + 1)
    cal = [make(10000 + k, crng) for k in range(args.cal_requests)]
    ccircs = [workloads.measured_circuit(qa, q["n"], q["depth"], q["gseed"]) for q in cal]
    t0 = time.perf_counter()
    ts = [srv.submit(c, q["dtype"], 1024, q["seed"]) for c, q in zip(ccircs, cal)]
    for t in ts:
        srv.wait(t)
        srv.release(t)
    cap = len(cal) / (time.perf_counter() - t0)
    lam = 0.8 * cap
    log(f"serve: capacity {cap:.1f} circuits/s (closed loop, {len(cal)} requests); rate {lam:.1f}/s")
    gaps = rng.exponential(1.0 / lam, size=nreq)
    arrivals = np.cumsum(gaps)
    tickets = []
    start = time.perf_counter()
    for k, (c, q) in enumerate(zip(circs, reqs)):
        now = time.perf_counter() - start
        if arrivals[k] > now:
            time.sleep(arrivals[k] - now)
        tickets.append(srv.submit(c, q["dtype"], 1024, q["seed"]))
        if (k + 1) % 50 == 0:
            log(f"serve: submitted {k + 1}/{nreq}, {srv.info()}")
    lat, run_ms, stats, bsz = [], [], [], []
    for t in tickets:
        w = srv.wait(t)
        lat.append(w[2])
        run_ms.append(w[1])
        stats.append(srv.stats(t))
        bsz.append(srv.batch_size(t))
    wall = time.perf_counter() - start
    info = srv.info()
    # a burst of small requests (12-14 qubits) submitted at once: they queue
    # together and run batched
    brng = np.random.default_rng(args.seed + 2)
    burst = []
    for k in range(32):
        q = make(20000 + k, brng)
        q["n"] = int(brng.integers(12, 15))
        burst.append((workloads.measured_circuit(qa, q["n"], q["depth"], q["gseed"]), q))
    tb = time.perf_counter()
    bt = [srv.submit(c, q["dtype"], 1024, q["seed"]) for c, q in burst]
    blat, bsizes = [], []
    for t in bt:
        blat.append(srv.wait(t)[2])
        bsizes.append(srv.batch_size(t))
    bwall = time.perf_counter() - tb
    srv.close()
    burst_res = {"requests": len(burst), "circuits_per_s": round(len(burst) / bwall, 2),
                 "p50_ms": round(float(np.percentile(blat, 50)), 2),
                 "p99_ms": round(float(np.percentile(blat, 99)), 2),
                 "batch_sizes": sorted(set(bsizes)),
                 "launches": int(round(sum(1.0 / b for b in bsizes)))}
    log(f"serve burst: {burst_res}")
    lat = np.array(lat)
    log(f"serve: {nreq} requests in {wall:.2f} s, p50 {np.percentile(lat, 50):.1f} ms, "
        f"p99 {np.percentile(lat, 99):.1f} ms, {info}")
    # each result alone == in the mix: a few requests (the batched ones
    # first) served again, each on a server of its own
    same = 0
    batched = [k for k in range(nreq) if bsz[k] > 1]
    chk = (batched[:3] + [k for k in range(0, nreq, max(1, nreq // 6)) if k not in batched])[:6]
    for k in chk:
        s1 = qa.Server(lanes=1)
        t1 = s1.submit(circs[k], reqs[k]["dtype"], 1024, reqs[k]["seed"])
        s1.wait(t1)
        same += int(s1.stats(t1) == stats[k])
        s1.close()
    # precision of the complex64 requests: the gate part in c64 vs c128
    ok64, n64, worst = 0, 0, 0.0
    pass_bytes = 0.0
    for q in reqs:
        gates = workloads.random_layers(q["n"], q["depth"], q["gseed"])
        es = 8 if q["dtype"] == qa.C64 else 16
        if q["n"] >= 12:  # tile passes (what the engine runs)
            p = qa.Program(q["n"], gates, dtype=q["dtype"])
            pass_bytes += p.num_passes() * 2 * es * 2 ** q["n"]
            run = lambda st, dt: qa.Program(q["n"], gates, dtype=dt).run(st)  # noqa: E731
        else:  # below one tile the engine applies gate by gate
            pass_bytes += sum(workloads.gate_bytes(g, q["n"], es) for g in gates)
            run = lambda st, dt: workloads.apply_gates(st, gates)  # noqa: E731
        if q["dtype"] != qa.C64:
            continue
        s32 = qa.State([2] * q["n"], dtype=qa.C64)
        run(s32, qa.C64)
        s64 = qa.State([2] * q["n"])
        run(s64, qa.C128)
        inf = s64.compare(s32)["infidelity"]
        worst = max(worst, inf)
        n64 += 1
        ok64 += int(inf <= 1e-5)
    frac = (ok64 + (nreq - n64)) / nreq  # complex128 requests: exact to rounding
    # CPU baseline: small complex128 requests of the same stream through the
    # oracle's QEngine (qpp's execute loop and kernels restated), same seeds;
    # their statistics must equal the served ones
    cpu = None
    small = [k for k, q in enumerate(reqs) if q["n"] <= 14 and q["dtype"] == qa.C128][: args.cpu_requests]
    if small:
        import oracle
        t_c = time.perf_counter()
        eq = sum(int(oracle.engine_run(circs[k], 1024, -1.0, reqs[k]["seed"])[0] == stats[k][0])
                 for k in small)
        t_o = time.perf_counter() - t_c
        g_ms = sum(run_ms[k] for k in small)
        cpu = {"value": round(len(small) / t_o, 3), "unit": "circuits/s (small requests only)",
               "cores": oracle_threads(), "kind": "port",
               "sample": f"{len(small)} complex128 requests of the stream with n <= 14 "
                         f"(n = {[reqs[k]['n'] for k in small]}, depth = "
                         f"{[reqs[k]['depth'] for k in small]}) through the oracle's "
                         f"QEngine.execute(1024): {t_o * 1e3:.1f} ms (the GPU server ran the "
                         f"same requests in {g_ms:.1f} ms in the mix); statistics equal to the "
                         f"served ones {eq}/{len(small)}; requests up to 26 qubits and depth "
                         f"100 are out of reach for a bounded CPU sample"}
    return {
        "metric": "served circuits/s + p99 latency (1024 shots each)",
        "value": round(nreq / wall, 3),
        "unit": "circuits/s",
        "n_gpus": 1, "steps": 1, "warmup": 0,
        "ms_per_step": round(wall * 1e3 / nreq, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "complex128+complex64", "data": "synthetic",
        "config": {"workload": f"{nreq} Poisson-arriving random circuits, n~U{{10..26}}, "
                               f"depth~U{{10..100}}, 1024 shots, c128/c64 50/50, "
                               f"{args.lanes} lanes",
                   "arrival_rate_per_s": round(lam, 3), "capacity_per_s": round(cap, 3),
                   "requests_sha256": workloads.requests_digest(reqs)},
        "latency_ms": {"p50": round(float(np.percentile(lat, 50)), 2),
                       "p99": round(float(np.percentile(lat, 99)), 2),
                       "mean": round(float(lat.mean()), 2), "max": round(float(lat.max()), 2)},
        "fraction_at_infidelity_le_1e-5": round(frac, 4),
        "batched": {"requests": len(batched),
                    "launches": int(sum(1.0 / bsz[k] for k in batched) + 0.5),
                    "max_batch": int(max(bsz)),
                    "rule": "12-14 qubits, gates then one measurement of every qubit: the gate "
                            "steps of every such queued request of one dtype run in one launch"},
        "burst_small": burst_res,
        "c64_requests": n64, "c64_worst_infidelity": worst,
        "alone_equals_mix": f"{same}/{len(chk)}",
        "server": info,
        "roofline": {"bound": "hbm", "kernel": "k_tile (all requests)",
                     "achieved": round(pass_bytes / wall / 1e9, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(pass_bytes / wall / 1e9 / HBM_PEAK_GBS, 4),
                     "traffic": None},
        "cpu_baseline": cpu,
    }


def bench_rho(args, qa):
    """Density matrix: 12 qubits (4096 x 4096 complex128, 256 MiB), rho =
    0.9 |psi><psi| + 0.1 I/4096; 10 layers of Haar 1q gates + CZ on pairs
    (2i, 2i+1), each gate followed by the depolarizing channel (p = 1e-3)
    on the qubits it touched; then ptrace over qubits 6..11, ptranspose on
    {0, 1, 2} and syspermute by a seeded permutation of 12."""
    from qpp_amd import workloads
    import oracle
    n, depth = 12, 10
    rho0, ops, tr, ptt, perm = workloads.dm_pipeline(n, depth, seed=args.seed)
    bp = workloads.dm_program(qa, n, ops).compile()
    npass = bp.info()["passes"]
    native_s = None
    if getattr(args, "native", 1):
        tn = time.perf_counter()
        try:
            bp.compile_native()
            native_s = time.perf_counter() - tn
            log(f"rho pipeline: specialised pass kernels compiled in {native_s:.1f} s")
        except qa.QppError as e:
            log(f"rho pipeline: specialised pass kernels not built: {e}")
    src = qa.State.from_array(rho0, [2] * n)  # input resident in HBM
    st = src.clone()

    def pipeline():
        s = src.clone()  # device-to-device copy of the input
        bp.run(s)
        a = s.ptrace(tr)
        b = s.ptranspose(ptt, qa.AS_BUILT)
        c = s.syspermute(perm, qa.AS_BUILT)
        return a, b, c

    for _ in range(args.warmup):
        pipeline()
    qa.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        outs = pipeline()
    qa.synchronize()
    wall = (time.perf_counter() - t0) / args.steps
    fused_ms = profile_region(qa, lambda: bp.run(st))
    # gate by gate (each op its own pass, as qpp applies them)
    s2 = qa.State.from_array(rho0, [2] * n)

    def one_by_one():
        for kind, q, M in ops:
            if kind == "u":
                s2.apply(M, q)
            elif kind == "cz":
                s2.apply_ctrl(M, [q[0]], [q[1]])
            else:
                s2.apply_kraus(M, q)

    gbg_ms = profile_region(qa, one_by_one)
    st.upload(rho0)
    bp.run(st)
    got = st.download()
    diff = float(np.max(np.abs(got - s2.download())))
    # maps of the device rho against the oracle on the same input
    exact = {
        "ptrace_maxabs": float(np.max(np.abs(st.ptrace(tr).download() - oracle.ptrace(got, tr, [2] * n)))),
        "ptranspose_bit_exact": bool(np.array_equal(st.ptranspose(ptt, qa.AS_BUILT).download(),
                                                    oracle.ptranspose(got, ptt, [2] * n, semantics=1))),
        "syspermute_bit_exact": bool(np.array_equal(st.syspermute(perm, qa.AS_BUILT).download(),
                                                    oracle.syspermute(got, perm, [2] * n, semantics=1))),
    }
    map_ms = profile_region(qa, lambda: (st.ptrace(tr), st.ptranspose(ptt, qa.AS_BUILT),
                                         st.syspermute(perm, qa.AS_BUILT)))
    s = 16
    algo = workloads.dm_bytes(n, ops) + (s * 2 ** (2 * n - len(tr)) + s * 4 ** (n - len(tr))) \
        + 2 * (2 * s * 4 ** n)
    pass_bytes = npass * 2 * s * 4 ** n
    # CPU baseline: the oracle (qpp's rho kernels restated, out-of-place, as
    # qpp applies them) on the first ops of the same pipeline, then its maps
    cpu = None
    if args.cpu_ops > 0:
        sub = ops[: args.cpu_ops]
        r = rho0
        t_c = time.perf_counter()
        for kind, q, M in sub:
            if kind == "u":
                r = oracle.apply(r, M, q, [2] * n)
            elif kind == "cz":
                r = oracle.applyCTRL(r, M, [q[0]], [q[1]], [2] * n)
            else:
                r = oracle.apply_kraus(r, M, q, [2] * n)
        t_ops = time.perf_counter() - t_c
        t_c = time.perf_counter()
        oracle.ptrace(r, tr, [2] * n)
        oracle.ptranspose(r, ptt, [2] * n, semantics=1)
        oracle.syspermute(r, perm, [2] * n, semantics=1)
        t_maps = time.perf_counter() - t_c
        cpu_bytes = workloads.dm_bytes(n, sub) + (algo - workloads.dm_bytes(n, ops))
        cpu = {"value": round(cpu_bytes / (t_ops + t_maps) / 1e9, 3), "unit": "GB/s",
               "cores": oracle_threads(), "kind": "port",
               "sample": f"first {len(sub)} of the {len(ops)} gates/channels ({t_ops:.2f} s, "
                         f"{t_ops / len(sub) * 1e3:.1f} ms each) and the three maps "
                         f"({t_maps:.2f} s) through the oracle (qpp's rho kernels restated)"}
    # HBM bytes the pipeline streams: every fused pass reads and writes rho
    # once, ptrace reads the entries it sums and writes the result, and
    # ptranspose / syspermute read and write rho once each
    streamed = pass_bytes + (algo - workloads.dm_bytes(n, ops))
    return {
        "metric": "12-qubit density-matrix pipeline GB/s (HBM bytes streamed)",
        "value": round(streamed / wall / 1e9, 2),
        "value_definition": "bytes the fused passes and the three maps stream / pipeline wall "
                            "time (the input copy excluded)",
        "gate_apply_equiv_GBps": round(algo / wall / 1e9, 2),
        "unit": "GB/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(wall * 1e3, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "complex128", "data": "synthetic",
        "config": {"workload": f"rho 4096x4096, {depth} layers: {len(ops)} gates+channels fused "
                               f"into {npass} passes, ptrace {tr}, ptranspose {ptt}, syspermute",
                   "ops": len(ops), "passes": npass,
                   "workload_sha256": workloads.dm_digest(rho0, ops, tr, ptt, perm),
                   "pass_kernels": (f"specialised per pass (hipRTC), compiled in {native_s:.1f} s "
                                    "outside the timed region") if native_s is not None
                                   else "general tile kernel k_tile"},
        "fused_ops_ms": round(fused_ms, 3),
        "gate_by_gate_ms": round(gbg_ms, 3),
        "maps_ms": round(map_ms, 3),
        "fused_vs_gate_by_gate_maxabs": diff,
        "maps_vs_oracle": exact,
        "roofline": {"bound": "hbm",
                     "kernel": "qs_pass_c128_<i>" if native_s is not None else "k_tile<double,16>",
                     "achieved": round(pass_bytes / (fused_ms / 1e3) / 1e9, 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(pass_bytes / (fused_ms / 1e3) / 1e9 / HBM_PEAK_GBS, 4),
                     "traffic": None, "avg_launch_ms": round(fused_ms / npass, 4)},
        "cpu_baseline": cpu,
    }


def bench_block(args, qa, n=30, reps=3):
    """Dense 4- and 5-qubit gates (16x16 and 32x32 unitaries, the widest
    blocks qpp::apply takes on the matrix cores here) on an n-qubit
    complex64 ket, every amplitude read and written once per launch: the
    FP32 MFMA kernel (k_gatek_mfma) against the VALU kernel (k_gatek, forced
    with QPP_AMD_GATEK_VALU=1), plus the complex128 VALU kernel.  HBM
    roofline: 2 s 2^n bytes per launch; arithmetic: 4 x 2^k real FMA per
    amplitude on either unit (the complex product as a real GEMM)."""
    import os
    from qpp_amd.workloads import haar
    rng = np.random.default_rng(args.seed)
    out = []
    for k in (4, 5):
        U = haar(2 ** k, rng)
        tg = [0, 7, 13, 21, 29][:k]  # low and high target bits
        row = {"qubits_in_gate": k, "targets": tg}
        for label, dt, valu in (("c64_mfma", qa.C64, False), ("c64_valu", qa.C64, True),
                                ("c128_valu", qa.C128, True)):
            st = qa.State([2] * n, dtype=dt)
            if valu:
                os.environ["QPP_AMD_GATEK_VALU"] = "1"
            try:
                st.apply(U, tg)  # warm-up
                qa.synchronize()
                ms = profile_region(qa, lambda: [st.apply(U, tg) for _ in range(reps)]) / reps
            finally:
                os.environ.pop("QPP_AMD_GATEK_VALU", None)
            es = 8 if dt == qa.C64 else 16
            gbs = 2 * es * 2 ** n / (ms / 1e3) / 1e9
            fl = 8 * 2 ** k * 2 ** n  # 4 * 2^k real FMA per amplitude on either unit
            row[label] = {"ms": round(ms, 3), "GBps": round(gbs, 1),
                          "hbm_frac": round(gbs / HBM_PEAK_GBS, 4),
                          "TFLOPs": round(fl / (ms / 1e3) / 1e12, 2)}
            del st
        row["c64_mfma"]["mfma_frac"] = round(row["c64_mfma"]["TFLOPs"] / 157.3, 4)
        row["mfma_speedup_vs_valu"] = round(row["c64_valu"]["ms"] / row["c64_mfma"]["ms"], 3)
        log(f"block k={k}: {row}")
        out.append(row)
    # the 5-qubit gate is the MFMA-heavy case: its roofline against the FP32
    # matrix peak, beside its HBM rate
    k5 = out[-1]["c64_mfma"]
    roof = {"bound": "mfma", "kernel": "k_gatek_mfma<5>", "achieved": k5["TFLOPs"], "peak": 157.3,
            "unit": "TFLOP/s", "frac": round(k5["TFLOPs"] / 157.3, 4), "traffic": None,
            "avg_launch_ms": k5["ms"], "hbm_GBps": k5["GBps"], "hbm_frac": k5["hbm_frac"]}
    # CPU: the oracle's qpp apply (dense k-qubit kernel restated, OpenMP) of
    # the same 32x32 gate on a 24-qubit complex128 ket, scaled by 2^6
    cpu = None
    try:
        import oracle
        m = 24
        psi = np.zeros(2 ** m, dtype=np.complex128)
        psi[0] = 1
        U5 = haar(32, rng)
        t0 = time.perf_counter()
        oracle.apply(psi, U5, [0, 7, 13, 21, 23], 2)
        dt_cpu = time.perf_counter() - t0
        cpu = {"value": round(2 * 16 * 2 ** m / dt_cpu / 1e9, 3), "unit": "GB/s", "cores": oracle_threads(),
               "kind": "port",
               "sample": f"one 32x32 gate on a {m}-qubit complex128 ket through the oracle's apply "
                         f"(qpp's apply_psi_kq restated, OpenMP): {dt_cpu:.2f} s; at {n} qubits about "
                         f"{dt_cpu * 2 ** (n - m):.0f} s (EXTRAPOLATED x 2^{n - m})"}
    except Exception as e:  # the oracle is test infrastructure; report its absence
        cpu = {"value": None, "unit": "GB/s", "cores": 0, "kind": "port", "sample": f"not run: {e}"}
    return {
        "metric": "dense k-qubit gate on a 30-qubit complex64 ket: MFMA vs VALU",
        "roofline": roof,
        "cpu_baseline": cpu,
        "workload": f"{n}-qubit ket, one dense 2^k x 2^k unitary per launch (k = 4, 5), "
                    f"{reps} launches timed with HIP events (matrix upload included)",
        "flops_definition": "8 x 2^k real flops per amplitude (4 x 2^k FMA: the complex "
                            "product, or the same count as a real GEMM on the matrix cores); "
                            "peaks 157.3 TFLOP/s FP32 matrix, 78.6 TFLOP/s FP32 / FP64 vector",
        "blocks": out,
    }


def bench_sample(args, qa, sizes=(20, 24, 26), shots=1024, reps=5):
    """qpp::sample of every qubit (1024 shots) on resident random complex128
    kets of 2^20..2^26 amplitudes: the error-bounded parallel search, and the
    exact libstdc++ table (binade scans) forced for every shot with
    QPP_AMD_FORCE_EXACT=1.  Host wall time per call (the call returns the
    indices to the host)."""
    import os
    rng = np.random.default_rng(args.seed)
    rows = []
    for n in sizes:
        D = 2 ** n
        psi = rng.standard_normal(D) + 1j * rng.standard_normal(D)
        psi /= np.linalg.norm(psi)
        st = qa.State.from_array(psi, [2] * n)
        u = rng.random(shots)
        row = {"qubits": n}
        for label, force in (("bounded_search_ms", "0"), ("exact_table_ms", "1")):
            os.environ["QPP_AMD_FORCE_EXACT"] = force
            try:
                st.sample(list(range(n)), u)
                qa.synchronize()
                t = time.perf_counter()
                for _ in range(reps):
                    st.sample(list(range(n)), u)
                qa.synchronize()
                row[label] = round((time.perf_counter() - t) / reps * 1e3, 3)
            finally:
                os.environ.pop("QPP_AMD_FORCE_EXACT", None)
        # exact table traffic: p read 4 times (sum, q scan, table, search
        # neighbourhood) and two tables written: about 5 x 8 D bytes
        row["exact_table_GBps"] = round(5 * 8 * D / (row["exact_table_ms"] / 1e3) / 1e9, 1)
        log(f"sample n={n}: {row}")
        rows.append(row)
        del st
    return {"metric": "qpp::sample of 1024 shots, ms per call", "shots": shots, "rows": rows}


FP64_PROBE_SRC = r"""
typedef double d4 __attribute__((ext_vector_type(4)));
// 16 complex rotations per thread per iteration: the compiled loop is 32
// v_mul_f64, 32 v_fmac_f64 and 32 v_add_f64
extern "C" __global__ __launch_bounds__(256) void p_valu(const double* __restrict__ in,
        const double* __restrict__ cs, double* __restrict__ out, int iters) {
    const unsigned i = blockIdx.x * 256u + threadIdx.x;
    double x[16], y[16];
    for (int k = 0; k < 16; ++k) { x[k] = in[i * 32u + 2u * k]; y[k] = in[i * 32u + 2u * k + 1u]; }
    const double c = cs[2u * i], s = cs[2u * i + 1u];
    for (int it = 0; it < iters; ++it) {
#pragma unroll
        for (int k = 0; k < 16; ++k) {
            const double nx = fma(c, x[k], -(s * y[k]));
            const double ny = fma(s, x[k], c * y[k]);
            x[k] = nx; y[k] = ny;
        }
    }
    double acc = 0;
    for (int k = 0; k < 16; ++k) acc += x[k] + y[k];
    out[i] = acc;
}
// four independent v_mfma_f64_16x16x4f64 chains per wave
extern "C" __global__ __launch_bounds__(256) void p_mfma(const double* __restrict__ in,
        const double* __restrict__ cs, double* __restrict__ out, int iters) {
    const unsigned i = blockIdx.x * 256u + threadIdx.x;
    const double a = in[i * 32u], b0 = in[i * 32u + 1u], b1 = in[i * 32u + 2u];
    d4 q0 = {0, 0, 0, 0}, q1 = q0, q2 = q0, q3 = q0;
    for (int it = 0; it < iters; ++it) {
        const double b = (it & 1) ? b1 : b0;
        q0 = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, q0, 0, 0, 0);
        q1 = __builtin_amdgcn_mfma_f64_16x16x4f64(b, a, q1, 0, 0, 0);
        q2 = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, q2, 0, 0, 0);
        q3 = __builtin_amdgcn_mfma_f64_16x16x4f64(b, a, q3, 0, 0, 0);
    }
    out[i] = q0[0] + q1[1] + q2[2] + q3[3];
}
// both in one wave: the same rotations and MFMA chains interleaved
extern "C" __global__ __launch_bounds__(256) void p_both(const double* __restrict__ in,
        const double* __restrict__ cs, double* __restrict__ out, int iters) {
    const unsigned i = blockIdx.x * 256u + threadIdx.x;
    double x[16], y[16];
    for (int k = 0; k < 16; ++k) { x[k] = in[i * 32u + 2u * k]; y[k] = in[i * 32u + 2u * k + 1u]; }
    const double c = cs[2u * i], s = cs[2u * i + 1u];
    const double a = in[i * 32u], b0 = in[i * 32u + 1u], b1 = in[i * 32u + 2u];
    d4 q0 = {0, 0, 0, 0}, q1 = q0, q2 = q0, q3 = q0;
    for (int it = 0; it < iters; ++it) {
        const double b = (it & 1) ? b1 : b0;
        q0 = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, q0, 0, 0, 0);
        q1 = __builtin_amdgcn_mfma_f64_16x16x4f64(b, a, q1, 0, 0, 0);
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const double nx = fma(c, x[k], -(s * y[k]));
            const double ny = fma(s, x[k], c * y[k]);
            x[k] = nx; y[k] = ny;
        }
        q2 = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, q2, 0, 0, 0);
        q3 = __builtin_amdgcn_mfma_f64_16x16x4f64(b, a, q3, 0, 0, 0);
#pragma unroll
        for (int k = 8; k < 16; ++k) {
            const double nx = fma(c, x[k], -(s * y[k]));
            const double ny = fma(s, x[k], c * y[k]);
            x[k] = nx; y[k] = ny;
        }
    }
    double acc = q0[0] + q1[1] + q2[2] + q3[3];
    for (int k = 0; k < 16; ++k) acc += x[k] + y[k];
    out[i] = acc;
}
"""


def bench_fp64(args, qa):
    """FP64 on the vector ALUs against FP64 on the matrix cores, with zero
    and with random operands: whether complex128 gate arithmetic could move
    to v_mfma_f64_16x16x4f64, and how much the clock depends on the data.
    Kernels compiled here with hipRTC; 2048 workgroups of 256 threads."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    rtc = ctypes.CDLL("libhiprtc.so")
    vp = ctypes.c_void_p
    prog = vp()
    assert rtc.hiprtcCreateProgram(ctypes.byref(prog), FP64_PROBE_SRC.encode(), b"probe.hip", 0,
                                   None, None) == 0
    opts = [b"--offload-arch=gfx950", b"-O3", b"-Xclang", b"-target-feature", b"-Xclang",
            b"-packed-fp32-ops"]
    if rtc.hiprtcCompileProgram(prog, len(opts), (ctypes.c_char_p * len(opts))(*opts)) != 0:
        raise RuntimeError("fp64 probe: hipRTC compile failed")
    size = ctypes.c_size_t()
    rtc.hiprtcGetCodeSize(prog, ctypes.byref(size))
    code = ctypes.create_string_buffer(size.value)
    rtc.hiprtcGetCode(prog, code)
    mod = vp()
    assert hip.hipModuleLoadData(ctypes.byref(mod), code) == 0
    G, iters = 2048, 20000
    T = G * 256
    bufs = {}
    for name, n in (("in", 32 * T), ("cs", 2 * T), ("out", T)):
        bufs[name] = vp()
        assert hip.hipMalloc(ctypes.byref(bufs[name]), ctypes.c_size_t(8 * n)) == 0
    ev0, ev1 = vp(), vp()
    hip.hipEventCreate(ctypes.byref(ev0))
    hip.hipEventCreate(ctypes.byref(ev1))
    rng = np.random.default_rng(args.seed)
    # per thread-iteration: VALU FP64 instructions (p_valu's compiled loop)
    # and MFMA flops (4 x 16x16x4 x 2 / 64 lanes)
    cases = [("p_valu", 96.0, 0.0), ("p_mfma", 0.0, 128.0), ("p_both", 96.0, 128.0)]
    rows = []
    for data in ("zeros", "random"):
        if data == "zeros":
            hin, hcs = np.zeros(32 * T), np.zeros(2 * T)
        else:
            hin = rng.standard_normal(32 * T)
            th = rng.random(T) * 2 * np.pi
            hcs = np.stack([np.cos(th), np.sin(th)], 1).reshape(-1).copy()
        for name, arr in (("in", hin), ("cs", hcs)):
            assert hip.hipMemcpy(bufs[name], arr.ctypes.data_as(vp), ctypes.c_size_t(arr.nbytes), 1) == 0
        for kname, vi, mf in cases:
            f = vp()
            assert hip.hipModuleGetFunction(ctypes.byref(f), mod, kname.encode()) == 0
            it = ctypes.c_int(iters)
            argv = (vp * 4)(*[ctypes.cast(ctypes.byref(x), vp) for x in (bufs["in"], bufs["cs"], bufs["out"], it)])
            best = float("inf")
            for _ in range(4):  # the first is a warm-up
                hip.hipEventRecord(ev0, None)
                assert hip.hipModuleLaunchKernel(f, G, 1, 1, 256, 1, 1, 0, None, argv, None) == 0
                hip.hipEventRecord(ev1, None)
                hip.hipEventSynchronize(ev1)
                ms = ctypes.c_float()
                hip.hipEventElapsedTime(ctypes.byref(ms), ev0, ev1)
                best = min(best, ms.value)
            sec = best / 1e3
            out = np.empty(T)
            hip.hipMemcpy(out.ctypes.data_as(vp), bufs["out"], ctypes.c_size_t(out.nbytes), 2)
            # FMA-rate issue: 16 FP64 lanes per clock per SIMD at 2.4 GHz
            valu = vi * T * iters / sec / (256 * 4 * 16 * 2.4e9)
            mtf = mf * T * iters / sec / 1e12
            rows.append({"data": data, "kernel": kname, "ms": round(best, 3),
                         "valu_fp64_issue_vs_fma_peak": round(valu, 3),
                         "mfma_f64_TFLOPs": round(mtf, 1), "mfma_frac": round(mtf / 78.6, 3),
                         "finite": bool(np.isfinite(out).all())})
            log(f"fp64 probe {data} {kname}: {rows[-1]}")
    for b in bufs.values():
        hip.hipFree(b)
    return {"metric": "FP64 vector vs matrix-core throughput (probe)", "unit": "ms",
            "higher_is_better": False, "n_gpus": 1, "data": "synthetic",
            "config": {"workload": f"{G} workgroups x 256 threads, {iters} iterations; "
                                   "p_valu: 16 complex rotations per thread per iteration; "
                                   "p_mfma: 4 v_mfma_f64_16x16x4f64 chains per wave; p_both: both "
                                   "interleaved in one wave"},
            "rows": rows}


def bench_copy(args, qa):
    """Reference streaming kernel: a 30-qubit complex128 state (16 GiB) copied
    into another with qs_state_copy (each byte read once, written once).
    Used to calibrate memory-traffic counters and as the practical HBM
    ceiling."""
    import ctypes
    n = args.qubits
    L = qa.lib()
    L.qs_state_copy.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
    src = qa.State([2] * n)
    dst = qa.State([2] * n)
    for _ in range(max(1, args.warmup)):
        qa._check(L.qs_state_copy(dst.handle, src.handle))
    reps = max(1, args.steps)
    ms = profile_region(qa, lambda: [qa._check(L.qs_state_copy(dst.handle, src.handle))
                                     for _ in range(reps)]) / reps
    nbytes = 2 * 16 * 2 ** n
    gbs = nbytes / (ms / 1e3) / 1e9
    return {
        "metric": "HBM copy GB/s (reference streaming kernel)",
        "value": round(gbs, 1), "unit": "GB/s", "n_gpus": 1, "steps": reps,
        "warmup": args.warmup, "ms_per_step": round(ms, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "complex128", "data": "synthetic",
        "config": {"workload": f"copy of a {n}-qubit complex128 state ({nbytes / 2**30:.0f} GiB moved)",
                   "bytes_per_launch": nbytes},
        "roofline": {"bound": "hbm", "kernel": "k_stream_copy", "achieved": round(gbs, 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(gbs / HBM_PEAK_GBS, 4),
                     "traffic": None, "avg_launch_ms": round(ms, 4)},
        "cpu_baseline": None,
    }


def bench_circuit(args, qa, n, depth, primary):
    """The random circuit (n qubits, `depth` layers) on a resident complex128
    state: one step applies the whole circuit.  The primary run (30 qubits,
    depth 40) also tunes the fused options, splits off the complex64 run and
    times the streaming copy ceiling."""
    from qpp_amd import workloads
    gates = workloads.random_layers(n, depth, args.seed)
    n1q = sum(1 for g in gates if g[0] == "u")
    ncx = len(gates) - n1q
    st = qa.State([2] * n)

    mode = args.mode
    if args.sweep:
        sub = gates[: (n + n // 2) * 4]
        for rb in (3, 4):
            for low in (5, 6):
                for mo in (24, 48):
                    for gc in (512, 1024, 2048, 8192):
                        p = qa.Program(n, sub, reg_bits=rb, low_bits=low, max_ops=mo, grid_cap=gc)
                        p.run(st)
                        ms = profile_region(qa, lambda: p.run(st))
                        npas = p.num_passes()
                        gbs = npas * 2 * 16 * 2 ** n / (ms / 1e3) / 1e9
                        log(f"sweep reg_bits={rb} low_bits={low} max_ops={mo} grid={gc}: "
                            f"passes={npas} phases={p.info()['phases']} {ms:.1f} ms "
                            f"({ms / npas:.2f} ms/pass, {gbs:.0f} GB/s per pass, "
                            f"{len(sub) / ms * 1e3:.0f} gates/s)")
    tuned = None
    if mode == "fused" and args.autotune and primary:
        from qpp_amd import autotune
        cands = [dict(reg_bits=4, low_bits=lb, max_ops=mo, grid_cap=gc)
                 for lb in (3, 4, 5) for mo in (48, 64, 96) for gc in (16384, 32768, 65536)]
        tt = time.perf_counter()
        tuned = autotune.tune(n, gates, candidates=cands, time_budget_s=90,
                              prefix_gates=12 * (n + n // 2), native=bool(args.native),
                              verbose=True)
        log(f"autotune: {tuned} in {time.perf_counter() - tt:.1f} s")
        args.reg_bits, args.low_bits = tuned["reg_bits"], tuned["low_bits"]
        args.max_ops, args.grid_cap = tuned["max_ops"], tuned["grid_cap"]
    native_s = None
    if mode == "fused":
        prog = qa.Program(n, gates, reg_bits=args.reg_bits, low_bits=args.low_bits,
                          max_ops=args.max_ops, grid_cap=args.grid_cap)
        if args.native and args.reg_bits == 4:
            tn = time.perf_counter()
            try:
                prog.compile_native()
                native_s = time.perf_counter() - tn
                log(f"specialised pass kernels compiled in {native_s:.1f} s")
            except qa.QppError as e:  # stays on the general tile kernel
                log(f"specialised pass kernels not built: {e}")
        run = lambda: prog.run(st)  # noqa: E731
        npasses = prog.num_passes()
        pass_bytes = npasses * 2 * 16 * 2 ** n
        kernel_name = prog.kernel_name()
    else:
        run = lambda: run_gates(qa, st, gates)  # noqa: E731
        npasses = len(gates)
        pass_bytes = sum(workloads.gate_bytes(g, n) for g in gates)
        kernel_name = "k_gate1_hi<double,4>"
    # gate-apply bytes: every gate application at its own algorithmic bytes
    # (2*16*2^n per dense gate, 2*16*2^(n-1) per CNOT), whether the engine
    # fuses the gates into passes or not
    algo_bytes = sum(workloads.gate_bytes(g, n) for g in gates)

    for w in range(args.warmup):
        run()
        qa.synchronize()
        log(f"warmup {w + 1}/{args.warmup} done")
    qa.synchronize()
    t0 = time.perf_counter()
    ev_ms = profile_region(qa, lambda: [run() for _ in range(args.steps)])
    qa.synchronize()
    wall = time.perf_counter() - t0
    log(f"timed {args.steps} steps: wall {wall:.3f} s, events {ev_ms / 1e3:.3f} s")
    sec = wall / args.steps
    # value: the HBM rate the circuit's passes stream (every pass reads and
    # writes the whole state once: 2*16*2^n B), over circuit wall time
    value = pass_bytes / sec / 1e9
    # the same circuit counted gate by gate: what an unfused engine would
    # have to stream to match this time (not an HBM rate)
    equiv = algo_bytes / sec / 1e9

    # dominant kernel: every pass timed with HIP events recorded between the
    # pass launches on the state's stream
    pass_ms = None
    if mode == "fused":
        pass_ms = np.concatenate([prog.run_timed(st) for _ in range(3)])
        avg_ms = float(pass_ms.mean())
        k_bytes = 2 * 16 * 2 ** n
    else:
        U = gates[0][3]
        hi = [q for q in range(n) if n - 1 - q >= 6]
        reps = 2
        ms = profile_region(qa, lambda: [st.apply(U, [q]) for _ in range(reps) for q in hi])
        avg_ms = ms / (reps * len(hi))
        k_bytes = 2 * 16 * 2 ** n
    achieved = k_bytes / (avg_ms / 1e3) / 1e9
    log(f"dominant kernel {kernel_name}: {avg_ms:.3f} ms/launch, {achieved:.1f} GB/s")

    # precision split: the same circuit from |0..0> in complex64 (FP32
    # amplitudes, FP64 reductions) against complex128, compared on the device
    # in FP64
    precision = None
    if mode == "fused" and args.precision and primary:
        st.set_basis(0)
        prog.run(st)
        s64 = qa.State([2] * n, dtype=qa.C64)
        p64 = qa.Program(n, gates, dtype=qa.C64, reg_bits=args.reg_bits, low_bits=args.low_bits,
                         max_ops=args.max_ops, grid_cap=args.grid_cap)
        if native_s is not None:
            p64.compile_native()
        p64.run(s64)
        cmp = st.compare(s64)
        s64.set_basis(0)
        ms64 = profile_region(qa, lambda: p64.run(s64))
        precision = {"complex64_ms_per_circuit": round(ms64, 2),
                     "complex64_circuits_per_s": round(1e3 / ms64, 4),
                     "complex64_passes": p64.num_passes(),
                     "infidelity_vs_complex128": cmp["infidelity"],
                     "trace_distance_vs_complex128": cmp["trace_distance"],
                     "max_abs_vs_complex128": cmp["max_abs"],
                     "complex128_norm2": cmp["norm2_a"], "complex64_norm2": cmp["norm2_b"]}
        log(f"precision: c64 {ms64:.1f} ms/circuit, infidelity {cmp['infidelity']:.3e}, "
            f"max-abs {cmp['max_abs']:.3e}")
        del s64, p64

    out = {
        "metric": METRIC,
        "value": round(value, 2),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(sec * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "complex128",
        "data": "synthetic",
        "config": {
            "workload": f"{n}-qubit complex128 random circuit, depth {depth} "
                        f"({n1q} Haar 1q + {ncx} CNOT), {'tile-fused passes' if mode == 'fused' else 'gate by gate'}",
            "qubits": n, "depth": depth, "gates": len(gates), "mode": mode,
            "passes_per_step": npasses, "seed": args.seed,
            "circuit_sha256": workloads.gates_digest(gates),
            "fuse_opts": {"reg_bits": args.reg_bits, "low_bits": args.low_bits,
                          "max_ops": args.max_ops, "grid_cap": args.grid_cap,
                          "autotuned": tuned is not None},
            "pass_kernels": ("specialised per pass (generated HIP, hipRTC), compiled in "
                             f"{native_s:.1f} s outside the timed region") if native_s is not None
                            else "general tile kernel k_tile",
        },
        "hbm_frac": round(value / HBM_PEAK_GBS, 4),
        "value_definition": "HBM bytes the circuit's passes stream (2*16*2^n per pass: the state "
                            "read and written once) / circuit wall time",
        "gate_apply_equiv_GBps": round(equiv, 1),
        "gate_apply_equiv_definition": "every gate application at its own algorithmic bytes "
                                       "(2*16*2^n per dense 1q gate, 2*16*2^(n-1) per CNOT) / "
                                       "circuit wall time: the rate an unfused engine would need "
                                       "to match this circuit time; not an HBM rate",
        "gate_apps_per_s": round(len(gates) / sec, 2),
        "circuits_per_s": round(1.0 / sec, 4),
        # the arithmetic side of the same circuit: a dense complex 2x2 gate
        # costs 14 FP64 flops per amplitude (two complex products and a sum
        # per output), a CNOT none; rate over the circuit's wall time against
        # MI355X's dense FP64 peak
        "fp64_compute": {
            "flops_per_circuit": n1q * 14 * 2 ** n,
            "achieved_TFLOPs": round(n1q * 14 * 2 ** n / sec / 1e12, 2),
            "peak_TFLOPs": FP64_PEAK_TFLOPS,
            "frac": round(n1q * 14 * 2 ** n / sec / 1e12 / FP64_PEAK_TFLOPS, 4),
        },
        "precision": precision,
        "roofline": {
            "bound": "hbm",
            "kernel": kernel_name,
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": None,
            "avg_launch_ms": round(avg_ms, 4),
            "launch_ms_min_max": ([round(float(pass_ms.min()), 3), round(float(pass_ms.max()), 3)]
                                  if pass_ms is not None else None),
            "timing": ("HIP events between the pass launches on the state's stream, "
                       f"{len(pass_ms)} launches" if pass_ms is not None else
                       "HIP events around the launches on the library stream"),
        },
        "cpu_baseline": None,
    }
    if primary and args.copy_ceiling:
        cc = bench_copy(argparse.Namespace(qubits=n, steps=3, warmup=1), qa)
        out["roofline"]["copy_ceiling_GBps"] = cc["value"]
        out["roofline"]["frac_of_copy_ceiling"] = round(achieved / cc["value"], 4)
        log(f"copy ceiling: {cc['value']} GB/s")
    # HBM bytes per launch of this kernel at this size, measured with
    # rocprofv3 FETCH_SIZE / WRITE_SIZE and scaled by the copy-kernel
    # calibration (profiles/traffic.json; the summaries beside it hold the
    # commands and counter values)
    try:
        with open(os.path.join(ROOT, "profiles", "traffic.json")) as f:
            tr = json.load(f).get(kernel_name)
        if tr and tr.get("qubits") == n and mode == "fused":
            out["roofline"]["traffic"] = tr["bytes_per_launch"]
    except (OSError, ValueError):
        pass
    if args.cpu_gates != 0:
        if n <= 22:
            picks = list(range(len(gates)))
        elif args.cpu_gates > 0:  # the circuit's first N gates (45: one whole layer at 30 qubits)
            picks = list(range(min(args.cpu_gates, len(gates))))
        else:  # dense gates on high and low qubits, and CNOTs, of layer 0
            picks = [0, 1, n - 2, n - 1, n, n + 1]
        log(f"cpu baseline: {len(picks)} gates at {n} qubits ...")
        out["cpu_baseline"] = cpu_baseline(gates, n, picks)
    return out


def secondary_configs(args, qa):
    """The other benchmark configurations, each at a bounded size so that the
    default run stays within minutes; every object carries its own roofline
    and CPU baseline.  Sizes are stated in each object's config."""
    out = []
    base = vars(args).copy()

    def ns(**kw):
        d = dict(base)
        d.update(kw)
        return argparse.Namespace(**d)

    t = time.perf_counter()
    # 20-qubit complex128 random circuit, depth 20 (the CPU-runnable plain
    # method), fused with fixed options; the oracle runs the whole circuit
    a0 = ns(qubits=20, depth=20, steps=20, warmup=3, reg_bits=4, low_bits=3, max_ops=48,
            grid_cap=65536, autotune=0, precision=0, cpu_gates=-1, sweep=False)
    r0 = bench_circuit(a0, qa, 20, 20, primary=False)
    r0["config_id"] = 0
    r0["metric"] = "20q random circuit GB/s of the fused passes"
    out.append(r0)
    log(f"secondary 20-qubit circuit: {time.perf_counter() - t:.1f} s")
    t = time.perf_counter()
    r1 = bench_rho(ns(steps=3, warmup=1, cpu_ops=8), qa)
    r1["config_id"] = 1
    out.append(r1)
    log(f"secondary 12-qubit rho pipeline: {time.perf_counter() - t:.1f} s")
    t = time.perf_counter()
    r3 = bench_batch(ns(steps=2, warmup=1, traj_scale=0.25, dtype="c64", oracle_traj=6,
                        oracle_traj16=2), qa)
    r3["config_id"] = 3
    out.append(r3)
    log(f"secondary batched trajectories: {time.perf_counter() - t:.1f} s")
    t = time.perf_counter()
    r4 = bench_serve(ns(requests=60, cal_requests=12, cpu_requests=4), qa)
    r4["config_id"] = 4
    out.append(r4)
    log(f"secondary serving: {time.perf_counter() - t:.1f} s")
    t = time.perf_counter()
    rb = bench_block(ns(), qa, n=30)
    rb["config_id"] = 2
    out.append(rb)
    log(f"secondary dense blocks on the matrix cores: {time.perf_counter() - t:.1f} s")
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--qubits", type=int, default=30)
    ap.add_argument("--depth", type=int, default=40)
    ap.add_argument("--seed", type=int, default=30)
    ap.add_argument("--mode", choices=["fused", "gates", "batch", "serve", "rho", "copy", "block", "sample",
                                       "fp64"],
                    default="fused")
    ap.add_argument("--requests", type=int, default=200, help="serve mode: requests")
    ap.add_argument("--cal-requests", type=int, default=40,
                    help="serve mode: closed-loop calibration requests")
    ap.add_argument("--lanes", type=int, default=4, help="serve mode: concurrent lanes (the device has 4 hardware queues)")
    ap.add_argument("--traj-low-bits", type=int, default=5, help="batch mode: fused low_bits")
    ap.add_argument("--traj-max-ops", type=int, default=48, help="batch mode: fused max_ops")
    ap.add_argument("--dtype", choices=["c128", "c64"], default="c64",
                    help="batch mode: complex dtype of the trajectories")
    ap.add_argument("--traj-scale", type=float, default=1.0,
                    help="batch mode: fraction of the 4096/512/512/512 trajectories")
    ap.add_argument("--traj-depth", type=int, default=30)
    ap.add_argument("--oracle-traj", type=int, default=16,
                    help="batch mode: 12-qubit trajectories checked against the oracle engine")
    ap.add_argument("--oracle-traj16", type=int, default=6,
                    help="batch mode: 16-qubit trajectories run on the oracle engine (parity + CPU baseline)")
    ap.add_argument("--cpu-ops", type=int, default=24,
                    help="rho mode: gates/channels of the pipeline timed on the oracle (CPU baseline)")
    ap.add_argument("--cpu-requests", type=int, default=4,
                    help="serve mode: small requests run on the oracle engine (parity + CPU baseline)")
    ap.add_argument("--secondary", type=int, default=1,
                    help="fused mode: also run the other configurations (20-qubit circuit, "
                         "12-qubit rho pipeline, batched trajectories, serving) at bounded "
                         "sizes and report them under 'secondary'")
    ap.add_argument("--copy-ceiling", type=int, default=1,
                    help="time the streaming copy of the same state as the HBM ceiling")
    ap.add_argument("--cpu-gates", type=int, default=-1,
                    help="0 skips the CPU baseline; otherwise it runs the whole circuit up "
                         "to 22 qubits, else (-1) 4 dense gates and 2 CNOTs of the first "
                         "layer or (N > 0) the first N gates (45: one whole layer at 30 "
                         "qubits, about 3 minutes), extrapolated per gate kind")
    ap.add_argument("--precision", type=int, default=1,
                    help="fused mode: also run the circuit in complex64 and compare with complex128 (1/0)")
    ap.add_argument("--reg-bits", type=int, default=4)
    ap.add_argument("--low-bits", type=int, default=5)
    ap.add_argument("--max-ops", type=int, default=64)
    ap.add_argument("--grid-cap", type=int, default=0)
    ap.add_argument("--native", type=int, default=1,
                    help="fused mode: run the passes as specialised kernels (compiled once per program)")
    ap.add_argument("--autotune", type=int, default=1,
                    help="fused mode: pick tile options with the autotuner first (1/0)")
    ap.add_argument("--sweep", action="store_true",
                    help="time fused options on the first layers and print them")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1",
                          "n_gpus": args.gpus}))
        sys.exit(2)

    import qpp_amd as qa
    from qpp_amd import workloads

    info = qa.device_info()
    log("device:", info)
    if args.mode == "batch":
        print(json.dumps(bench_batch(args, qa)), flush=True)
        return
    if args.mode == "serve":
        print(json.dumps(bench_serve(args, qa)), flush=True)
        return
    if args.mode == "rho":
        print(json.dumps(bench_rho(args, qa)), flush=True)
        return
    if args.mode == "copy":
        print(json.dumps(bench_copy(args, qa)), flush=True)
        return
    if args.mode == "block":
        print(json.dumps(bench_block(args, qa, n=args.qubits)), flush=True)
        return
    if args.mode == "sample":
        print(json.dumps(bench_sample(args, qa)), flush=True)
        return
    if args.mode == "fp64":
        print(json.dumps(bench_fp64(args, qa)), flush=True)
        return
    out = bench_circuit(args, qa, args.qubits, args.depth, primary=True)
    if args.mode == "fused" and args.secondary:
        out["secondary"] = secondary_configs(args, qa)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
