"""Writes tests/golden/ondemand/*.npz: stage outputs of the reference's accurate
ascending labelling (ondemandaccurate, needASC2) for small seeded volumes.

Runs in the development container only:
  make -C oracle ref && python tests/golden/make_ondemand_golden.py
The reference driver (oracle/refdrv/ref_ondemand.cpp) runs the reference's own
classes with one OpenMP thread (its first integration pass depends on the
visiting order).  Region labels are stored canonically: each region is named
by its lowest vertex.  Stored per case: shape, input SHA-256, parameters,
certain / integrated / cleaned labels (int32, zlib), restriction (int8, zlib),
.grad (zlib), and the reference's REFTIME line.
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

REF = os.path.join(ROOT, "oracle", "_ref", "ref_ondemand")


def canonical(lab, f):
    """Rename each label by its lowest vertex (value, then vertex id)."""
    flat = f.ravel()
    order = np.lexsort((np.arange(flat.size), flat))
    out = np.full_like(lab, -1)
    seen = {}
    for v in order:
        l = int(lab[v])
        if l != -1 and l not in seen:
            seen[l] = int(v)
    for l, v in seen.items():
        out[lab == l] = v
    return out


def z(a):
    return np.frombuffer(zlib.compress(np.ascontiguousarray(a).tobytes(), 9), np.uint8)


def main():
    out_dir = os.path.join(HERE, "ondemand")
    os.makedirs(out_dir, exist_ok=True)
    with tempfile.TemporaryDirectory() as tmp:
        for name, (shape, gen, err, gthr, iters, presimp) in ONDEMAND_CASES.items():
            f = np.ascontiguousarray(gen(), dtype=np.float32)
            assert f.shape == tuple(shape), (name, f.shape)
            Z, Y, X = f.shape
            raw = os.path.join(tmp, name + ".raw")
            f.tofile(raw)
            pre = os.path.join(tmp, name)
            r = subprocess.run([REF, str(X), str(Y), str(Z), raw, pre, repr(err), repr(gthr), str(iters),
                                repr(presimp), "1"], check=True, capture_output=True, text=True)
            line = [l for l in r.stdout.splitlines() if l.startswith("REFTIME")][0]
            lab = {s: canonical(np.fromfile(pre + "." + s, np.int32), f) for s in ("certain", "integrated", "cleaned")}
            res = np.fromfile(pre + ".restriction", np.int8)
            grad = np.fromfile(pre + ".grad", np.uint8)
            np.savez(os.path.join(out_dir, name + ".npz"), shape=np.array(f.shape, np.int64),
                     input_sha256=np.array(synth.sha256(f)),
                     params=np.array([err, gthr, iters, presimp], np.float64),
                     certain_z=z(lab["certain"]), integrated_z=z(lab["integrated"]), cleaned_z=z(lab["cleaned"]),
                     restriction_z=z(res), grad_z=z(grad), reftime=np.array(line))
            print(f"{name}: {X}x{Y}x{Z} regions={len(np.unique(lab['cleaned']))} {line}")


if __name__ == "__main__":
    main()
