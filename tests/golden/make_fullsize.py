"""Writes tests/golden/fullsize_oracle.json: the FP64 oracle's transmission eigenvalues of the
full-size guides of tests/fullsize_cases.py (the seeded potentials of the oracle's generator,
which the device generator reproduces to a few ulp).

    python tests/golden/make_fullsize.py [--threads 8]

The oracle (oracle/rg_oracle.c) runs one realization per thread; the c3 guide at its full length
(N = 256, 2000 slices) takes about 90 s a realization on one core, everything together a few
minutes on 8 cores.  The fixture stores, per case, the parameters, the seeds and T (descending).
"""
import argparse
import concurrent.futures as cf
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(HERE))

from fullsize_cases import UNIFORM, uneven_sample  # noqa: E402
from oracle import oracle as O  # noqa: E402

OUT = os.path.join(HERE, "fullsize_oracle.json")


def jobs():
    out = []
    for name, (p, seeds) in UNIFORM.items():
        out.extend((name, p, s) for s in seeds)
    ps, seeds = uneven_sample()
    out.extend(("c5", p, s) for p, s in zip(ps, seeds))
    return out


def run(job):
    name, p, seed = job
    t0 = time.time()
    T = O.Waveguide(**p).tdistrib_seed(seed)
    return name, p, seed, T.tolist(), time.time() - t0


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--threads", type=int, default=os.cpu_count())
    args = ap.parse_args()
    O.build()
    todo = sorted(jobs(), key=lambda j: -(j[1]["nxdiso"] * (2 * j[1]["wol"]) ** 3))  # longest first
    res = {}
    with cf.ThreadPoolExecutor(args.threads) as ex:  # ctypes releases the GIL during oracle calls
        for name, p, seed, T, dt in ex.map(run, todo):
            print(f"{name} seed {seed}: N={int(2 * p['wol'])} nx={p['nxdiso']} {dt:.1f} s", flush=True)
            res.setdefault(name, []).append({"params": p, "seed": seed, "T": T})
    for v in res.values():
        v.sort(key=lambda e: (e["params"]["wol"], e["seed"]))
    with open(OUT, "w") as f:
        json.dump(res, f, indent=0)
    print(f"wrote {OUT}")


if __name__ == "__main__":
    main()
