"""Writes tests/golden/ers/*.npz: the reference's own
NumericIntegratorExpandingRegionStop (AdaptiveEulerAdvector3D<F, -1>,
IndexCompareLessThan) on the small labelling volumes, run with one OpenMP
thread, and how far its 8-thread runs are from that (the class's pass 1 is
schedule-dependent).

Runs in the development container only:
  make -C oracle ref && python tests/golden/make_ers_golden.py
The reference driver is oracle/refdrv/ref_ers.cpp (oracle/_ref/ref_ers).
Stored per case: shape, input SHA-256, parameters, the one-thread labels
(int32 vertex ids of the minima, zlib) and the fractions of vertices on which
two 8-thread runs differ from it.
"""
import os
import subprocess
import sys
import tempfile
import zlib

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, HERE)

from cases import ONDEMAND_CASES  # noqa: E402
from msceer_amd import synth  # noqa: E402

REF = os.path.join(ROOT, "oracle", "_ref", "ref_ers")


def run(raw, shape, params, pre, threads):
    Z, Y, X = shape
    err, gthr, iters = params
    r = subprocess.run([REF, str(X), str(Y), str(Z), raw, repr(err), repr(gthr), str(iters), pre, str(threads)],
                       check=True, capture_output=True, text=True)
    ext = int([l for l in r.stdout.splitlines() if l.startswith("EXTREMA")][0].split()[1])
    return np.fromfile(pre + ".labels", np.int64).astype(np.int32), ext


def main():
    out_dir = os.path.join(HERE, "ers")
    os.makedirs(out_dir, exist_ok=True)
    with tempfile.TemporaryDirectory() as tmp:
        for name, (shape, gen, err, gthr, iters, presimp) in ONDEMAND_CASES.items():
            f = np.ascontiguousarray(gen(), dtype=np.float32)
            raw = os.path.join(tmp, name + ".raw")
            f.tofile(raw)
            pre = os.path.join(tmp, name)
            lab1, ext = run(raw, f.shape, (err, gthr, iters), pre, 1)
            spread = [float((run(raw, f.shape, (err, gthr, iters), pre, 8)[0] != lab1).mean()) for _ in range(2)]
            np.savez(os.path.join(out_dir, name + ".npz"), shape=np.array(f.shape, np.int64),
                     input_sha256=np.array(synth.sha256(f)), params=np.array([err, gthr, iters], np.float64),
                     extrema=np.array(ext, np.int64),
                     labels_z=np.frombuffer(zlib.compress(lab1.tobytes(), 9), np.uint8),
                     threads8_mismatch=np.array(spread, np.float64))
            print(f"{name}: extrema {ext}, 8-thread runs differ on {[round(100 * x, 3) for x in spread]} %",
                  flush=True)


if __name__ == "__main__":
    main()
