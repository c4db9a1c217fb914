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


def ondemand_case_names():
    from cases import ONDEMAND_CASES
    return sorted(ONDEMAND_CASES)


def load_ondemand(name):
    """(values, params(err, gthr, iters, presimp), dict of reference stage outputs)."""
    from cases import ONDEMAND_CASES
    d = np.load(os.path.join(HERE, "ondemand", name + ".npz"))
    shape, gen, err, gthr, iters, presimp = ONDEMAND_CASES[name]
    f = np.ascontiguousarray(gen(), dtype=np.float32)
    assert tuple(d["shape"]) == tuple(shape) == f.shape
    assert synth.sha256(f) == str(d["input_sha256"]), f"{name}: input generator drifted"
    un = lambda k, dt: np.frombuffer(zlib.decompress(d[k].tobytes()), dt)
    out = {"certain": un("certain_z", np.int32), "integrated": un("integrated_z", np.int32),
           "cleaned": un("cleaned_z", np.int32), "restriction": un("restriction_z", np.int8),
           "grad": un("grad_z", np.uint8)}
    return f, (err, gthr, iters, presimp), out


LINE_FLAGS = [(1, 0, 0), (0, 1, 0), (1, 1, 0), (1, 1, 1)]


def lines_case_names():
    """(case, flags) pairs of the 1-manifold golden vectors."""
    from cases import ONDEMAND_CASES
    return [(n, fl) for n in ONDEMAND_CASES for fl in LINE_FLAGS]


def load_lines(name, flags):
    """(f, params, restriction (cells int8), SHA-256 of the conforming .grad)
    of the reference's run with needASC1, needDSC1, needASC2 = flags."""
    from cases import ONDEMAND_CASES
    shape, gen, *_ = ONDEMAND_CASES[name]
    z = np.load(os.path.join(HERE, "lines", "%s_%d%d%d.npz" % ((name,) + tuple(flags))), allow_pickle=False)
    f = np.ascontiguousarray(gen(), np.float32)
    assert synth.sha256(f) == str(z["input_sha256"]), "input generator drifted from the fixture"
    assert tuple(int(v) for v in z["flags"]) == tuple(flags)
    err, gthr, iters, presimp = (float(v) for v in z["params"])
    res = np.frombuffer(zlib.decompress(z["restriction_z"].tobytes()), np.int8)
    return f, (err, gthr, int(iters), presimp), res, str(z["grad_sha256"])


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


C3_SHAPE = (256, 256, 256)
C3_PARAMS = (1e-3, 1e-6, 4096, 0.0)  # err_thr, grad_thr, max_iters, presimp


def load_c3_fp64(exact=False):
    """(labels (V,) int32 of the FP64 oracle on the 256^3 workload, input SHA-256);
    exact: every vertex's own FP64 destination (oracle mode 2) instead of the
    path-compressing reference pipeline in FP64 (mode 1)."""
    import lzma
    d = np.load(os.path.join(HERE, "c3", "labels_exact_256.npz" if exact else "labels_fp64_256.npz"))
    assert tuple(d["shape"]) == C3_SHAPE
    idx = np.frombuffer(lzma.decompress(d["idx_lzma"].tobytes()), np.uint16)
    return d["table"][idx].astype(np.int32), str(d["input_sha256"])


def msc_case_names():
    return sorted(os.path.splitext(f)[0] for f in os.listdir(os.path.join(HERE, "msc")) if f.endswith(".npz"))


def load_msc(name):
    """(values, reference .grad, basins, mounts, node counts by dimension) of
    the reference's Morse-Smale complex read of a lower-star golden .grad."""
    f, grad = load_lstar(name)
    d = np.load(os.path.join(HERE, "msc", name + ".npz"))
    assert tuple(d["shape"]) == f.shape and str(d["input_sha256"]) == synth.sha256(f)
    basins = np.frombuffer(zlib.decompress(d["basins_z"].tobytes()), np.int32).reshape(f.shape)
    mounts = np.frombuffer(zlib.decompress(d["mounts_z"].tobytes()), np.int32).reshape(f.shape)
    return f, grad, basins, mounts, d["counts"]


def load_ers(name):
    """(values, (err, gthr, iters), one-thread reference labels, extrema,
    8-thread mismatch fractions) of NumericIntegratorExpandingRegionStop on a
    labelling case."""
    from cases import ONDEMAND_CASES
    shape, gen, err, gthr, iters, presimp = ONDEMAND_CASES[name]
    f = np.ascontiguousarray(gen(), dtype=np.float32)
    d = np.load(os.path.join(HERE, "ers", name + ".npz"))
    assert tuple(d["shape"]) == f.shape and str(d["input_sha256"]) == synth.sha256(f)
    lab = np.frombuffer(zlib.decompress(d["labels_z"].tobytes()), np.int32).reshape(f.shape)
    return f, (err, gthr, iters), lab, int(d["extrema"]), d["threads8_mismatch"]


def dsc2_case_names():
    return sorted(os.path.splitext(f)[0] for f in os.listdir(os.path.join(HERE, "dsc2")) if f.endswith(".npz"))


def load_dsc2(tag):
    """(values, (err, gthr, iters, presimp), flags (asc1, dsc1, asc2, dsc2), dict of
    the reference's descending certain / integrated / cleaned labels and
    restriction, .grad SHA-256) of a needDSC2 case."""
    d = np.load(os.path.join(HERE, "dsc2", tag + ".npz"))
    from cases import ONDEMAND_CASES
    name = tag.rsplit("_", 1)[0]
    shape, gen, err, gthr, iters, presimp = ONDEMAND_CASES[name]
    f = np.ascontiguousarray(gen(), dtype=np.float32)
    assert tuple(d["shape"]) == f.shape and str(d["input_sha256"]) == synth.sha256(f)
    out = {k: np.frombuffer(zlib.decompress(d[k + "_z"].tobytes()), np.int32).reshape(f.shape)
           for k in ("certain", "integrated", "cleaned")}
    out["restriction"] = np.frombuffer(zlib.decompress(d["restriction_z"].tobytes()), np.int8)
    return f, (err, gthr, iters, presimp), tuple(int(x) for x in d["flags"]), out, str(d["grad_sha256"])


def same_partition(a, b):
    """Do two label arrays split the vertices the same way (ids may differ)?"""
    a = np.asarray(a).ravel().astype(np.int64)
    b = np.asarray(b).ravel().astype(np.int64)
    pairs = np.unique(np.stack([a, b], 1), axis=0)
    return len(pairs) == len(np.unique(a)) == len(np.unique(b))
