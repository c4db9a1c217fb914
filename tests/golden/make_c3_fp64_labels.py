"""Writes the FP64 oracle's accurate ascending labels of the 256^3 labelling
workload (200 Gaussians, sigma 4-16, plus 0.05 sin sin sin, seed 303; err 1e-3
(squared-error test), grad threshold 1e-6, 4096 steps, no presimplification):

  tests/golden/c3/labels_fp64_256.npz   (oracle mode 1: the reference pipeline,
      ondemandaccurate.cxx:728-1061, with every advection quantity in FP64,
      path compression in the reference's one-thread visiting order)
  tests/golden/c3/labels_exact_256.npz  (oracle mode 2: every vertex's own FP64
      destination, each non-certain vertex integrated alone to a certain
      terminal; with --mode 2)

These labels are what the GPU labels are measured against at full size (label
mismatch rate).  The oracle's reference-arithmetic mode is pinned bit-exact to
reference runs by tests/test_oracle_ondemand.py.

Runs in the development container (a minute or two on 8 cores):
  make -C oracle liboracle.so && python tests/golden/make_c3_fp64_labels.py [--mode 2]
Stored: shape, input SHA-256, parameters, the sorted table of region labels
(int32) and the per-vertex index into it (uint16, LZMA).
"""
import lzma
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [ROOT, os.path.join(ROOT, "tests")]

import oracle_lib  # noqa: E402
from msceer_amd import synth  # noqa: E402

PARAMS = (1e-3, 1e-6, 4096, 0.0)


def main():
    mode = 2 if "--mode" in sys.argv and sys.argv[sys.argv.index("--mode") + 1] == "2" else 1
    f = synth.labels_field(303, 256)
    oracle_lib.set_threads(os.cpu_count() or 1)
    t0 = time.perf_counter()
    _, lab, _, st = oracle_lib.ondemand_asc(f, *PARAMS, mode=mode)
    dt = time.perf_counter() - t0
    table, idx = np.unique(lab, return_inverse=True)
    assert table.size < 65536
    out = os.path.join(HERE, "c3", "labels_fp64_256.npz" if mode == 1 else "labels_exact_256.npz")
    os.makedirs(os.path.dirname(out), exist_ok=True)
    np.savez(out, shape=np.array(f.shape), input_sha256=synth.sha256(f), params=np.array(PARAMS),
             table=table.astype(np.int32),
             idx_lzma=np.frombuffer(lzma.compress(idx.astype(np.uint16).tobytes(), preset=9), np.uint8),
             oracle_stats=st, oracle_seconds=dt)
    print(f"{out}: {table.size} regions, oracle {dt:.1f} s, stats {st}")


if __name__ == "__main__":
    main()
