"""Writes tests/golden/dsc2/*.npz: the reference's needDSC2 branch
(descending 3-manifolds, ondemandaccurate.cxx:913-991), alone and after
needASC2, on the small labelling volumes.

Runs in the development container only:
  make -C oracle ref && python tests/golden/make_dsc2_golden.py
The reference driver is oracle/refdrv/ref_ondemand.cpp (oracle/_ref/ref_ondemand)
with one OpenMP thread.  Stored per case and flag combination: shape, input
SHA-256, parameters, flags, the descending certain / integrated / cleaned
labels (int32, zlib; the reference's region ids), the restriction labels
(int8 per cell, zlib: 2 and 3 carry the MAX boundary) and the SHA-256 of the
conforming .grad.
"""
import hashlib
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

REF = os.path.join(ROOT, "oracle", "_ref", "ref_ondemand")
FLAGS = [(0, 0, 0, 1), (0, 0, 1, 1)]  # asc1 dsc1 asc2 dsc2


def z(a):
    return np.frombuffer(zlib.compress(np.ascontiguousarray(a).tobytes(), 9), np.uint8)


def main():
    out_dir = os.path.join(HERE, "dsc2")
    os.makedirs(out_dir, exist_ok=True)
    with tempfile.TemporaryDirectory() as tmp:
        for name, (shape, gen, err, gthr, iters, presimp) in ONDEMAND_CASES.items():
            f = np.ascontiguousarray(gen(), dtype=np.float32)
            Z, Y, X = f.shape
            raw = os.path.join(tmp, name + ".raw")
            f.tofile(raw)
            for flags in FLAGS:
                tag = "%s_%d%d%d%d" % ((name,) + flags)
                pre = os.path.join(tmp, tag)
                subprocess.run([REF, str(X), str(Y), str(Z), raw, pre, repr(err), repr(gthr), str(iters),
                                repr(presimp), "1", *map(str, flags)], check=True, capture_output=True, text=True)
                rd = {k: np.fromfile(pre + "." + k, np.int32) for k in ("certain_dsc", "integrated_dsc", "cleaned_dsc")}
                res = np.fromfile(pre + ".restriction", np.int8)
                grad = np.fromfile(pre + ".grad", np.uint8)
                np.savez(os.path.join(out_dir, tag + ".npz"), shape=np.array(f.shape, np.int64),
                         input_sha256=np.array(synth.sha256(f)),
                         params=np.array([err, gthr, iters, presimp], np.float64), flags=np.array(flags, np.int64),
                         certain_z=z(rd["certain_dsc"]), integrated_z=z(rd["integrated_dsc"]),
                         cleaned_z=z(rd["cleaned_dsc"]), restriction_z=z(res),
                         grad_sha256=np.array(hashlib.sha256(grad.tobytes()).hexdigest()))
                print(f"{tag}: regions {len(np.unique(rd['cleaned_dsc']))}, MAX-boundary cells "
                      f"{int((res >= 2).sum())}", flush=True)


if __name__ == "__main__":
    main()
