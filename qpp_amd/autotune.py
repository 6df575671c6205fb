"""Autotuner for the tile-fused pass engine.

The fused engine has four knobs (qs_fuse_opts): register bits per thread
(tile = 2^(8 + reg_bits) amplitudes: 2^11 or 2^12), low index bits every
tile keeps for coalescing, the fusion width (ops per pass, which trades
VALU work per byte against the number of passes over HBM) and the
persistent grid size.  Which setting is fastest depends on the qubit count,
the dtype and the gate mix, so the tuner times candidates on the device over
a prefix of the actual circuit (HIP events on the library stream, one
warm-up run each) and keeps the fastest per (qubits, dtype, kind, gate
mix) key.  Results are cached in memory and, when `path` is given, in a JSON
file.

    opts = autotune.tune(n, gates)                 # dict of qs_fuse_opts fields
    prog = qpp_amd.Program(n, gates, **opts)
"""
import json
import os
import sys
import time

import numpy as np

_CACHE = {}

CANDIDATES = [
    dict(reg_bits=rb, low_bits=lb, max_ops=mo, grid_cap=gc)
    for rb in (3, 4)
    for lb in (3, 4, 5, 6)
    for mo in (16, 32, 48, 64, 96)
    for gc in (2048, 4096, 8192, 16384, 32768, 65536)
]


def _events_ms(qa, fn):
    import ctypes
    L = qa.lib()
    L.qs_profile_begin.restype = ctypes.c_int
    L.qs_profile_end.argtypes = [ctypes.POINTER(ctypes.c_double)]
    qa._check(L.qs_profile_begin())
    fn()
    ms = ctypes.c_double()
    qa._check(L.qs_profile_end(ctypes.byref(ms)))
    return ms.value


def _key(n, dtype, kind, gates):
    n1 = sum(1 for g in gates if not g[2])
    mix = round(n1 / max(1, len(gates)), 1)
    return f"n={n},dtype={dtype},kind={kind},dense_frac={mix}"


def tune(n, gates, dtype=None, kind=None, prefix_gates=None, candidates=None,
         time_budget_s=60.0, path=None, verbose=False, native=False, native_top=3):
    """Fastest qs_fuse_opts for this circuit shape (dict usable as Program
    keyword arguments).  `gates`: a workloads gate list.  native: the
    program will run as specialised pass kernels (Program.compile_native),
    so the `native_top` fastest distinct reg_bits-4 candidates on the
    general kernel are timed again as specialised kernels (seconds of
    compilation each) and the fastest of those is kept."""
    import qpp_amd as qa
    dtype = qa.C128 if dtype is None else dtype
    kind = qa.KET if kind is None else kind
    key = _key(n, dtype, kind, gates) + (",native" if native else "")
    if path and not _CACHE and os.path.exists(path):
        with open(path) as f:
            _CACHE.update(json.load(f))
    if key in _CACHE:
        return dict(_CACHE[key]["opts"])
    sub = gates[: prefix_gates or min(len(gates), 6 * (n + n // 2))]
    st = qa.State([2] * n, kind, dtype)
    best, best_t = None, float("inf")
    t_end = time.perf_counter() + time_budget_s
    tried = []
    for opt in candidates or CANDIDATES:
        if time.perf_counter() > t_end:
            break
        try:
            p = qa.Program(n, sub, kind=kind, dtype=dtype, **opt)
        except qa.QppError:
            continue
        p.run(st)  # warm-up (first launch of this shape)
        ms = _events_ms(qa, lambda: p.run(st))
        tried.append((ms, opt))
        if verbose:
            print(f"autotune {key} {opt}: {ms:.3f} ms ({p.num_passes()} passes)", file=sys.stderr)
        if ms < best_t:
            best, best_t = opt, ms
    if native and tried:
        tried.sort(key=lambda x: x[0])
        picks, seen = [], set()
        for _, opt in tried:
            k = (opt["low_bits"], opt["max_ops"])
            if opt["reg_bits"] != 4 or k in seen:
                continue
            seen.add(k)
            picks.append(opt)
            if len(picks) == native_top:
                break
        nbest, nbest_t = None, float("inf")
        for opt in picks:
            try:
                p = qa.Program(n, sub, kind=kind, dtype=dtype, **opt).compile_native()
            except qa.QppError:
                continue
            p.run(st)
            ms = _events_ms(qa, lambda: p.run(st))
            if verbose:
                print(f"autotune {key} native {opt}: {ms:.3f} ms ({p.num_passes()} passes)",
                      file=sys.stderr)
            if ms < nbest_t:
                nbest, nbest_t = opt, ms
        if nbest is not None:
            best, best_t = nbest, nbest_t
    if best is None:
        best = dict(reg_bits=4, low_bits=5, max_ops=48, grid_cap=2048)
    _CACHE[key] = {"opts": best, "ms": best_t, "gates": len(sub),
                   "tried": len(tried)}
    if path:
        with open(path, "w") as f:
            json.dump(_CACHE, f, indent=1, sort_keys=True)
    return dict(best)


def tune_batch(circuits, dtype=None, circuit_of=None, seeds=None, noise_p=1e-3,
               candidates=None, verbose=False):
    """Fastest (reg_bits, low_bits, max_ops) for a trajectory batch, timed
    on the batch itself (kernel time of one run per candidate)."""
    import qpp_amd as qa
    dtype = qa.C64 if dtype is None else dtype
    if circuit_of is None:
        circuit_of = np.arange(len(circuits), dtype=np.uint32).repeat(64)
    if seeds is None:
        seeds = np.arange(len(circuit_of), dtype=np.uint64) + 1
    best, best_t = None, float("inf")
    for opt in candidates or [dict(reg_bits=rb, low_bits=lb, max_ops=mo)
                              for rb in (3, 4) for lb in (5,) for mo in (24, 48, 64)]:
        try:
            b = qa.Batch(circuits, dtype=dtype, **opt)
        except qa.QppError:
            continue
        b.run(circuit_of, seeds, noise_p)
        ms = b.info(0)["device_ms"]
        if verbose:
            print(f"autotune batch {opt}: {ms:.2f} ms", file=sys.stderr)
        if ms < best_t:
            best, best_t = opt, ms
    return dict(best or dict(reg_bits=4, low_bits=5, max_ops=48))
