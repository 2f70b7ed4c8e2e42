"""Writes config_digests.json: the oracle's partition of the benchmark
graphs that are too large to run the oracle inside a test.

For each graph the digest holds the SHA-256 of part[] (int32, little
endian), the cut, the hierarchy's level sizes and the Jet iterations per
level, from oracle_dga_partition (the serial restatement of the pipeline,
oracle/oracle_dga.c).  tests/test_gpu_configs.py compares the GPU pipeline's
partition of the same generated graph with it.

Graphs: R-MAT 2^22 vertices, 2^26 unique undirected edges
(a, b, c = 0.57, 0.19, 0.19), seed 1, k = 256, ub 1.03, seed 1 (the
headline configuration).  The oracle needs a few minutes on one core.

Run: python tests/golden/make_config_digests.py  (writes next to this file).
"""
import hashlib
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)


def digest(part, st):
    return {
        "part_sha256": hashlib.sha256(np.ascontiguousarray(part, dtype=np.int32).tobytes()).hexdigest(),
        "cut": int(st["cut"]),
        "levels": int(st["levels"]),
        "level_n": [int(v) for v in st["level_n"]],
        "level_m": [int(v) for v in st["level_m"]],
        "level_iters": [int(v) for v in st["level_iters"]],
    }


def main():
    import hunyuansc25_amd as H   # host-side graph generator only (no device call)
    import oracle as O
    out = {}
    t0 = time.time()
    x, a, vw, aw = H.gen_rmat(22, 1 << 26, seed=1)
    part, st = O.dga_partition(x, a, 256, vw, aw, seed=1)
    d = digest(part, st)
    d["oracle_seconds"] = round(time.time() - t0, 1)
    out["rmat_s22_e26_k256_seed1"] = d
    print("rmat_s22_e26_k256_seed1", d["cut"], d["levels"], d["oracle_seconds"], "s", flush=True)
    with open(os.path.join(HERE, "config_digests.json"), "w") as f:
        json.dump(out, f, indent=1)


if __name__ == "__main__":
    main()
