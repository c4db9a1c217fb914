"""Writes tests/golden/msc/*.npz: the reference's Morse-Smale complex read of
a .grad before any simplification (what extractmsc computes from a .grad:
the nodes, the base basins and the base mountains) on the lower-star golden
cases' reference .grad bytes.

Runs in the development container only:
  make -C oracle ref && python tests/golden/make_msc_golden.py
The reference driver (oracle/refdrv/ref_msc.cpp built as oracle/_ref/ref_msc)
runs MorseSmaleComplexBasic::ComputeFromGrad and fillGeometry of every
minimum (ascending) and maximum (descending).  Stored per case: shape, the
input SHA-256, node counts by dimension, and zlib-compressed int32 basins and
mountains (node ids per vertex, -1 where none).
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

from golden import load_lstar  # noqa: E402
from msceer_amd import synth  # noqa: E402

REF = os.path.join(ROOT, "oracle", "_ref", "ref_msc")
CASES = ["noise_48x97x31", "smooth_60x50x40", "plateau_33x29x21", "ties_10x11x12", "slab_64x64x2", "tiny_5x2x3",
         "tiny_2x2x2", "c1_gauss64"]


def main():
    out_dir = os.path.join(HERE, "msc")
    os.makedirs(out_dir, exist_ok=True)
    with tempfile.TemporaryDirectory() as tmp:
        for name in CASES:
            f, grad = load_lstar(name)
            Z, Y, X = f.shape
            raw, gfile, pre = (os.path.join(tmp, name + s) for s in (".raw", ".grad", ""))
            f.tofile(raw)
            grad.tofile(gfile)
            subprocess.run([REF, str(X), str(Y), str(Z), raw, gfile, pre], check=True, capture_output=True)
            nodes = np.fromfile(pre + ".nodes", np.int64).reshape(-1, 2)
            basins = np.fromfile(pre + ".basins", np.int32)
            mounts = np.fromfile(pre + ".mounts", np.int32)
            counts = np.bincount(nodes[:, 1], minlength=4)
            np.savez(os.path.join(out_dir, name + ".npz"), shape=np.array(f.shape, np.int64),
                     input_sha256=np.array(synth.sha256(f)), counts=counts.astype(np.int64),
                     basins_z=np.frombuffer(zlib.compress(basins.tobytes(), 9), np.uint8),
                     mounts_z=np.frombuffer(zlib.compress(mounts.tobytes(), 9), np.uint8))
            print(f"{name}: {X}x{Y}x{Z} nodes by dimension {counts.tolist()}", flush=True)


if __name__ == "__main__":
    main()
