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


def cpu_baseline(gates, n, sample_gates):
    """Oracle (qpp qubit kernels restated, OpenMP) on the first
    `sample_gates` gates of the same 30-qubit circuit."""
    import ctypes
    from ctypes import POINTER, c_double, c_uint32
    import oracle
    L = oracle.lib()
    L.orc_time_gates.argtypes = [c_uint32, c_uint32, POINTER(c_uint32), POINTER(c_uint32),
                                 POINTER(c_uint32), POINTER(c_double), POINTER(c_double)]
    L.orc_time_gates.restype = ctypes.c_int
    g = gates[:sample_gates]
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
    return {
        "value": round(nbytes / total / 1e9, 3),
        "unit": "GB/s",
        "cores": int(threads),
        "kind": "port",
        "sample": f"first {len(g)} gates of the same 30-qubit circuit "
                  f"({int((kind == 0).sum())} dense 1q + {int((kind == 1).sum())} CNOT), "
                  f"{total:.2f} s, {len(g) / total:.3f} gate-apps/s; qpp 7.0.3 qubit kernels "
                  f"restated (apply_psi_1q / apply_ctrl_psi_1q, out-of-place)",
    }


def run_gates(qa, st, gates):
    from qpp_amd.workloads import apply_gates
    apply_gates(st, gates)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--qubits", type=int, default=30)
    ap.add_argument("--depth", type=int, default=40)
    ap.add_argument("--seed", type=int, default=30)
    ap.add_argument("--mode", choices=["fused", "gates"], default="fused")
    ap.add_argument("--cpu-gates", type=int, default=6,
                    help="gates in the CPU baseline sample (0 = skip)")
    ap.add_argument("--reg-bits", type=int, default=4)
    ap.add_argument("--low-bits", type=int, default=5)
    ap.add_argument("--max-ops", type=int, default=64)
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
    n, depth = args.qubits, args.depth
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
    if mode == "fused":
        prog = qa.Program(n, gates, reg_bits=args.reg_bits, low_bits=args.low_bits,
                          max_ops=args.max_ops)
        run = lambda: prog.run(st)  # noqa: E731
        npasses = prog.num_passes()
        pass_bytes = npasses * 2 * 16 * 2 ** n
        algo_bytes = pass_bytes
        kernel_name = prog.kernel_name()
    else:
        run = lambda: run_gates(qa, st, gates)  # noqa: E731
        npasses = len(gates)
        algo_bytes = sum(workloads.gate_bytes(g, n) for g in gates)
        kernel_name = "k_gate1_hi<double,4>"

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
    value = algo_bytes / sec / 1e9

    # dominant kernel, measured alone with events on the library stream
    if mode == "fused":
        reps = 3
        ms = profile_region(qa, lambda: [prog.run(st) for _ in range(reps)])
        avg_ms = ms / (reps * npasses)
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
        },
        "hbm_frac": round(value / HBM_PEAK_GBS, 4),
        "gate_apps_per_s": round(len(gates) / sec, 2),
        "circuits_per_s": round(1.0 / sec, 4),
        "roofline": {
            "bound": "hbm",
            "kernel": kernel_name,
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": None,
            "avg_launch_ms": round(avg_ms, 4),
        },
        "cpu_baseline": None,
    }
    if args.cpu_gates > 0:
        log(f"cpu baseline: {args.cpu_gates} gates at {n} qubits ...")
        out["cpu_baseline"] = cpu_baseline(gates, n, args.cpu_gates)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
