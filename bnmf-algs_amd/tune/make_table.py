"""Writes the tile table the library loads by default (gfx950.txt in this
directory): rows per pass tile for each (dtype, z class, size class) the
benchmark configurations touch, chosen by timing on the GPU this runs on.

    python bnmf-algs_amd/tune/make_table.py [--out PATH]

1. FP32, the 256-instance batched workload (x, y ~ U{32..512}, z ~ U{4..32}):
   whole-batch iterations, one height for all its classes then each class
   in turn (autotune_workload), heights 16, 32, 64.
2. FP32, single large instances 2048 x 2048 x 64, 2048 x 2048 x 32 and
   1024 x 1024 x 16 (one instance each, the instance alone on the GPU),
   heights 32, 64, 128, 256.
3. FP32, the serving workload's requests (x, y log-uniform in [64, 2048],
   z in [8, 64]), each size class not set by steps 1-2 timed on its own
   requests batched together, heights 32, 64, 128.
4. FP64, 1024 x 1024 x 32 with exact digamma (the plain form), heights 12
   to 32 in steps of 4, 64 and 128.
Classes no step reaches keep the built-in rule.
"""
import argparse
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
for p in (ROOT, os.path.join(ROOT, "bnmf-algs_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=os.path.join(HERE, "gfx950.txt"))
    args = ap.parse_args()
    # tune from the built-in rule, not from a table already in place
    os.environ["BNMF_TUNE_FILE"] = os.devnull
    import bench
    import bnmf_hip as hip

    hip.init(0)
    set_by = {}

    log("[tune] 1: the batched workload (FP32)")
    c3 = [bench.as_dtype(i, hip, np.float32) for i in bench.make_c3(hip)]
    res = hip.autotune_workload(np.float32, c3, candidates=(16, 32, 64), reps=10, trials=2,
                                refine=True, log=log)
    for k in res["rows"]:
        set_by[("f32",) + tuple(k)] = "batched workload"

    log("[tune] 2: single large instances (FP32)")
    for (x, y, z) in ((2048, 2048, 64), (2048, 2048, 32), (1024, 1024, 16)):
        inst = bench.as_dtype(bench.single(hip, x, y, z, 40.0, 0xB1D2), hip, np.float32)
        r = hip.autotune(np.float32, instances=[inst], candidates=(32, 64, 128, 256), reps=10,
                         trials=3, log=log)
        for k in r:
            set_by[("f32",) + tuple(k)] = f"{x}x{y}x{z} alone"

    log("[tune] 3: the serving workload's size classes (FP32)")
    pool = bench.make_pool(hip, 48)
    rest = [i for i in pool
            if ("f32",) + tuple(hip._tune_class(i.z, i.X.shape[0] * i.X.shape[1])) not in set_by]
    if rest:
        r = hip.autotune(np.float32, instances=rest, candidates=(32, 64, 128), reps=10, trials=2,
                         log=log)
        for k in r:
            set_by[("f32",) + tuple(k)] = "serving requests"

    log("[tune] 4: 1024x1024x32 (FP64)")
    c1 = bench.single(hip, 1024, 1024, 32, 40.0, 0xB1D1)
    r = hip.autotune(np.float64, instances=[c1], candidates=(12, 16, 20, 24, 28, 32, 64, 128), reps=5,
                     trials=2, use_psi_appr=False, log=log)
    for k in r:
        set_by[("f64",) + tuple(k)] = "1024x1024x32 FP64 alone"

    hip.tune_save(args.out)
    with open(args.out, "a") as fh:
        fh.write("# chosen on " + ", ".join(sorted(set(set_by.values()))) + "\n")
    log(f"[tune] wrote {args.out}: {len(set_by)} classes")


if __name__ == "__main__":
    main()
