"""Loader for the committed golden vectors."""
import os
import zlib

import numpy as np

from cases import LSTAR_CASES
from msceer_amd import synth

HERE = os.path.dirname(os.path.abspath(__file__))


def lstar_case_names():
    return sorted(LSTAR_CASES)


def load_lstar(name):
    """(values (Z,Y,X) float32, reference .grad bytes) of a lower-star golden case."""
    d = np.load(os.path.join(HERE, "lstar", name + ".npz"))
    shape, gen = LSTAR_CASES[name]
    f = np.ascontiguousarray(gen(), dtype=np.float32)
    assert tuple(d["shape"]) == tuple(shape) == f.shape
    assert synth.sha256(f) == str(d["input_sha256"]), f"{name}: input generator drifted"
    grad = np.frombuffer(zlib.decompress(d["grad_z"].tobytes()), np.uint8)
    return f, grad
