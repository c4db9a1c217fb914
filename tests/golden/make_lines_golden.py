"""Writes tests/golden/lines/*.npz: the reference's accurate 1-manifolds
(ondemandaccurate with needASC1 / needDSC1, alone and with needASC2) on the
small labelling volumes.

Runs in the development container only:
  make -C oracle ref && python tests/golden/make_lines_golden.py
The reference driver (oracle/refdrv/ref_ondemand.cpp built as
oracle/_ref/ref_lines) runs the reference's digitizing streamline
integrators with one OpenMP thread, in the order of its saddle loops.
Stored per case and flag combination: shape, input SHA-256, parameters,
flags, the restriction labels (int8 per cell, zlib: 4 on the ascending
1-manifolds, 5 on the descending ones, 1 on the MIN boundary), the SHA-256 of
the conforming .grad, and the reference's REFTIME line.
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

REF = os.path.join(ROOT, "oracle", "_ref", "ref_lines")
FLAGS = [(1, 0, 0), (0, 1, 0), (1, 1, 0), (1, 1, 1)]


def main():
    out_dir = os.path.join(HERE, "lines")
    os.makedirs(out_dir, exist_ok=True)
    with tempfile.TemporaryDirectory() as tmp:
        for name, (shape, gen, err, gthr, iters, presimp) in ONDEMAND_CASES.items():
            f = np.ascontiguousarray(gen(), dtype=np.float32)
            assert f.shape == tuple(shape), (name, f.shape)
            Z, Y, X = f.shape
            raw = os.path.join(tmp, name + ".raw")
            f.tofile(raw)
            for flags in FLAGS:
                tag = "%s_%d%d%d" % ((name,) + flags)
                pre = os.path.join(tmp, tag)
                r = subprocess.run([REF, str(X), str(Y), str(Z), raw, pre, repr(err), repr(gthr), str(iters),
                                    repr(presimp), "1", *map(str, flags)], check=True, capture_output=True, text=True)
                line = [l for l in r.stdout.splitlines() if l.startswith("REFTIME")][0]
                res = np.fromfile(pre + ".restriction", np.int8)
                grad = np.fromfile(pre + ".grad", np.uint8)
                np.savez(os.path.join(out_dir, tag + ".npz"), shape=np.array(f.shape, np.int64),
                         input_sha256=np.array(synth.sha256(f)),
                         params=np.array([err, gthr, iters, presimp], np.float64),
                         flags=np.array(flags, np.int64),
                         restriction_z=np.frombuffer(zlib.compress(res.tobytes(), 9), np.uint8),
                         grad_sha256=np.array(hashlib.sha256(grad.tobytes()).hexdigest()),
                         reftime=np.array(line))
                print(f"{tag}: {X}x{Y}x{Z} cells(4)={int((res == 4).sum())} cells(5)={int((res == 5).sum())} {line}",
                      flush=True)


if __name__ == "__main__":
    main()
