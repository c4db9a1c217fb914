"""Writes tests/golden/lstar/*.npz: reference .grad bytes for the lower-star cases.

Runs in the development container only, where the reference tree exists:
  make -C oracle ref && python tests/golden/make_lstar_golden.py
The reference driver (oracle/refdrv/ref_steepest.cpp) instantiates the
reference's own classes compiled from /root/reference; only its outputs are
committed.  Each fixture holds: shape (Z, Y, X), the input SHA-256, the .grad
bytes (zlib-compressed) and the thread count used.
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

from cases import LSTAR_CASES  # noqa: E402
from msceer_amd import synth  # noqa: E402

REF = os.path.join(ROOT, "oracle", "_ref", "ref_steepest")


def main():
    out_dir = os.path.join(HERE, "lstar")
    os.makedirs(out_dir, exist_ok=True)
    with tempfile.TemporaryDirectory() as tmp:
        for name, (shape, gen) in LSTAR_CASES.items():
            f = np.ascontiguousarray(gen(), dtype=np.float32)
            assert f.shape == tuple(shape), (name, f.shape, shape)
            Z, Y, X = f.shape
            raw = os.path.join(tmp, name + ".raw")
            f.tofile(raw)
            threads = 4
            subprocess.run([REF, str(X), str(Y), str(Z), raw, os.path.join(tmp, name), str(threads)],
                           check=True, capture_output=True)
            grad = np.fromfile(os.path.join(tmp, name) + ".grad", np.uint8)
            assert grad.size == (2 * X - 1) * (2 * Y - 1) * (2 * Z - 1)
            np.savez(os.path.join(out_dir, name + ".npz"),
                     shape=np.array(f.shape, np.int64),
                     input_sha256=np.array(synth.sha256(f)),
                     grad_z=np.frombuffer(zlib.compress(grad.tobytes(), 9), np.uint8),
                     threads=np.array(threads))
            ncrit = int((((grad >> 2) & 7) == 7).sum())
            print(f"{name}: {X}x{Y}x{Z} cells={grad.size} critical={ncrit}")


if __name__ == "__main__":
    main()
