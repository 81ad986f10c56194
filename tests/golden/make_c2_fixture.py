"""Writes the golden fixture of the 134,217,728-particle plain-FoF workload:
SHA-256 digests of the input and of the oracle's partition, in_halo and
halo_id, plus the halo count and member total.

The input is py_fof_amd.synth.clumpy_nfw(n=134217728, L=512, seed=2) (NFW
clumps 50 % + uniform, b 0.2, np_halo_min 10, one periodic region). The
oracle links it with every core (oracle_set_link_threads); the union always
hooks the larger root under the smaller, so the partition does not depend on
the thread count. Run from the repository root on a host with about 24 GB of
free memory:

    python tests/golden/make_c2_fixture.py [--n N] [--out PATH]

The digests are of little-endian arrays: roots and halo_id int32, in_halo
int32 (0/1), the input (3, N) float64 in Fortran order.
"""
import argparse
import hashlib
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)


def digest_input(xp):
    h = hashlib.sha256()
    flat = xp.reshape(-1, order="F")  # (3, N) Fortran order is x0,y0,z0,x1,...
    step = 1 << 24
    for s in range(0, flat.size, step):
        h.update(np.ascontiguousarray(flat[s:s + step]).tobytes())
    return h.hexdigest()


def digest(a):
    return hashlib.sha256(np.ascontiguousarray(a, dtype=np.int32).tobytes()).hexdigest()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--n", type=int, default=134_217_728)
    ap.add_argument("--seed", type=int, default=2)
    ap.add_argument("--threads", type=int, default=os.cpu_count() or 1)
    ap.add_argument("--out", default=os.path.join(HERE, "c2_nfw134M.json"))
    args = ap.parse_args()
    from oracle import oracle_fof, oracle_set_link_threads
    from py_fof_amd.synth import clumpy_nfw
    n = args.n
    L = 512.0 * (n / 134_217_728) ** (1.0 / 3.0)
    t0 = time.time()
    xp = clumpy_nfw(n=n, L=L, seed=args.seed)
    print(f"generated {n} particles in {time.time() - t0:.1f} s", flush=True)
    d_in = digest_input(xp)
    oracle_set_link_threads(args.threads)
    t0 = time.time()
    in_h, hid, _, roots = oracle_fof(xp, [n], L, 0.2, 10, with_labels=True)
    dt = time.time() - t0
    oracle_set_link_threads(0)
    print(f"oracle FoF in {dt:.1f} s on {args.threads} threads", flush=True)
    out = {
        "workload": f"clumpy_nfw(n={n}, L={L:g}, seed={args.seed}), b 0.2, np_halo_min 10, "
                    "one periodic region, bound_check False",
        "n": n, "L_Box": L, "seed": args.seed, "linking_length": 0.2, "np_halo_min": 10,
        "input_sha256": d_in,
        "roots_sha256": digest(roots),
        "halo_id_sha256": digest(hid),
        "in_halo_sha256": digest(in_h),
        "nhalo": int(hid.max()),
        "members": int(in_h.sum()),
        "components": int(np.count_nonzero(roots == np.arange(n))),
        "oracle_seconds": round(dt, 1),
        "oracle_threads": args.threads,
    }
    with open(args.out, "w") as f:
        json.dump(out, f, indent=1)
        f.write("\n")
    print(json.dumps(out, indent=1))


if __name__ == "__main__":
    main()
