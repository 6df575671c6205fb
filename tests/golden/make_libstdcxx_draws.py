"""Writes tests/golden/libstdcxx_draws.json: libstdc++'s own
std::discrete_distribution / std::generate_canonical<double, 53> outputs
(random.tcc:2656-2712, random.h:164-192) for seeded weight tables, produced
by compiling libstdcxx_draws.cpp with the system g++ and running it.

    python3 tests/golden/make_libstdcxx_draws.py
"""
import json
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def main():
    with tempfile.TemporaryDirectory() as tmp:
        exe = os.path.join(tmp, "draws")
        subprocess.run(["g++", "-O2", "-std=c++17", "-o", exe,
                        os.path.join(HERE, "libstdcxx_draws.cpp")], check=True)
        gxx = subprocess.run(["g++", "--version"], capture_output=True, text=True).stdout.split("\n")[0]
        rng = np.random.default_rng(12)
        cases = []
        for N, seed, K in ((7, 1, 300), (300, 99, 300), (1024, 12345, 300), (2, 7, 50)):
            p = rng.random(N) ** 3
            p[rng.integers(N, size=N // 5)] = 0.0
            if N == 2:
                p = np.array([1.0, 3.0])
            inp = "\n".join(float(v).hex() for v in p)
            out = subprocess.run([exe, str(seed), str(K)], input=inp, capture_output=True,
                                 text=True, check=True).stdout.split()
            cases.append({
                "seed": seed,
                "weights_hex": [float(v).hex() for v in p],
                "draws": [int(v) for v in out[:K]],
                "canonical_hex": out[K:K + 4],
                "probabilities_hex": out[K + 4:],
            })
    doc = {"about": "libstdc++ std::discrete_distribution<size_t>(mt19937(seed)) index streams, the "
                    "first 4 generate_canonical<double,53>(mt19937(seed)) uniforms and "
                    "discrete_distribution::probabilities(), for seeded weights; hex floats are exact.",
           "generator": "tests/golden/make_libstdcxx_draws.py + libstdcxx_draws.cpp",
           "compiler": gxx,
           "cases": cases}
    path = os.path.join(HERE, "libstdcxx_draws.json")
    with open(path, "w") as f:
        json.dump(doc, f, separators=(",", ":"))
    print(path, sum(len(c["draws"]) for c in cases), "draws")


if __name__ == "__main__":
    sys.exit(main())
