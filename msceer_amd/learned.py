"""Learned integration mask: a near-separatrix classifier (msceer_cnn_* C ABI)
trained on labels the GPU integrator itself produces.

Target of a vertex: 1 if its accurate label (the full integrator's) differs
from its combinatorial label (the minimum its discrete-gradient V-path
reaches), or if one of its All6 neighbours carries another accurate label.
Those are the vertices that must be integrated numerically; every other vertex
may take its combinatorial label.  The mask used for labelling is the network's
flag (probability > threshold) dilated by a margin of vertices.
"""
from __future__ import annotations

import ctypes
import os
import time

import numpy as np

from . import labels, native, synth

WEIGHTS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "weights", "sepnet.npz")


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


class SeparatrixNet:
    """One network on one msceer context (device memory, Adam state)."""

    def __init__(self, ctx: native.Context, seed: int = 0):
        self.ctx = ctx
        self._h = ctypes.c_void_p()
        native.check(native.lib().msceer_cnn_create(ctx.handle, seed, ctypes.byref(self._h)), "msceer_cnn_create")
        self.num_params = native.lib().msceer_cnn_num_params()
        self.threshold = 0.5
        self.dilation = 2

    def close(self):
        if self._h:
            native.lib().msceer_cnn_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def get_params(self) -> np.ndarray:
        p = np.empty(self.num_params, np.float32)
        native.check(native.lib().msceer_cnn_get_params(self.ctx.handle, self._h, _ptr(p)), "msceer_cnn_get_params")
        return p

    def set_params(self, p: np.ndarray):
        p = np.ascontiguousarray(p, np.float32)
        assert p.size == self.num_params
        native.check(native.lib().msceer_cnn_set_params(self.ctx.handle, self._h, _ptr(p)), "msceer_cnn_set_params")

    def predict(self, values: np.ndarray) -> np.ndarray:
        v = np.ascontiguousarray(values, np.float32)
        Z, Y, X = v.shape
        out = np.empty(v.shape, np.float32)
        native.check(native.lib().msceer_cnn_predict(self.ctx.handle, self._h, _ptr(v), X, Y, Z, _ptr(out), 0),
                     "msceer_cnn_predict")
        return out

    def gradient(self, values: np.ndarray, target: np.ndarray, pos_weight: float = 1.0, want_grad: bool = True):
        v = np.ascontiguousarray(values, np.float32)
        t = np.ascontiguousarray(target, np.uint8)
        Z, Y, X = v.shape
        loss = ctypes.c_float()
        g = np.empty(self.num_params, np.float32) if want_grad else None
        native.check(native.lib().msceer_cnn_gradient(self.ctx.handle, self._h, _ptr(v), X, Y, Z, _ptr(t),
                                                      float(pos_weight), ctypes.byref(loss),
                                                      _ptr(g) if want_grad else None, 0), "msceer_cnn_gradient")
        return float(loss.value), g

    def adam_step(self, lr: float):
        native.check(native.lib().msceer_cnn_adam_step(self.ctx.handle, self._h, float(lr)), "msceer_cnn_adam_step")

    def mask(self, values: np.ndarray, threshold: float | None = None, dilation: int | None = None):
        v = np.ascontiguousarray(values, np.float32)
        Z, Y, X = v.shape
        out = np.empty(v.shape, np.uint8)
        cnt = ctypes.c_int64()
        thr = self.threshold if threshold is None else threshold
        dil = self.dilation if dilation is None else dilation
        native.check(native.lib().msceer_cnn_mask(self.ctx.handle, self._h, _ptr(v), X, Y, Z, float(thr), int(dil),
                                                  _ptr(out), ctypes.byref(cnt), 0), "msceer_cnn_mask")
        return out, int(cnt.value)

    def save(self, path: str = WEIGHTS, **meta):
        os.makedirs(os.path.dirname(path), exist_ok=True)
        np.savez(path, params=self.get_params(), threshold=np.float32(self.threshold),
                 dilation=np.int32(self.dilation), **{k: np.asarray(v) for k, v in meta.items()})

    def load(self, path: str = WEIGHTS):
        d = np.load(path)
        self.set_params(d["params"])
        self.threshold = float(d["threshold"])
        self.dilation = int(d["dilation"])
        return d


def separatrix_target(accurate: np.ndarray, combinatorial: np.ndarray) -> np.ndarray:
    """1 where the accurate label differs from the combinatorial one or from an All6 neighbour's."""
    a = np.asarray(accurate)
    t = a != np.asarray(combinatorial)
    for ax in range(3):
        d = np.diff(a, axis=ax) != 0
        lo = [slice(None)] * 3
        hi = [slice(None)] * 3
        lo[ax] = slice(0, -1)
        hi[ax] = slice(1, None)
        t[tuple(lo)] |= d
        t[tuple(hi)] |= d
    return t.astype(np.uint8)


def training_volume(i: int, n: int = 128) -> np.ndarray:
    """Volume i of the training families (seed 5000 + i): even i smoothed noise
    (sigma 1.5-3), odd i Gaussian mixtures with the sine term."""
    seed = 5000 + i
    s = synth.Stream(seed ^ 0x5EED)
    if i % 2 == 0:
        sigma = 1.5 + 1.5 * float(s.uniform(1)[0])
        return synth.smoothed_noise((n, n, n), seed, sigma)
    # the 256^3 labelling family at its own scale: sigma 4-16, 200 Gaussians per 256^3
    return synth.gaussian_mixture((n, n, n), seed, max(8, 200 * n ** 3 // 256 ** 3), 4.0, 16.0, -1.0, 1.0, sine=0.05)


def label_pair(values, ctx, err=1e-3, gthr=1e-6, iters=4096):
    """(accurate labels, combinatorial labels) of a volume from the GPU integrator."""
    _, acc, _ = labels.ondemand_asc(values, err, gthr, iters, 0.0, want_grad=False, ctx=ctx)
    zero = np.zeros(values.shape, np.uint8)
    _, comb, _ = labels.ondemand_asc(values, err, gthr, iters, 0.0, mask=zero, want_grad=False, ctx=ctx)
    return acc, comb


def train(net: SeparatrixNet, volumes, targets, epochs: int = 4, lr: float = 2e-3, pos_weight: float = 4.0,
          log=None):
    """Adam over the volumes, one volume per step; returns the loss history."""
    hist = []
    t0 = time.perf_counter()
    for ep in range(epochs):
        for k, (v, t) in enumerate(zip(volumes, targets)):
            loss, _ = net.gradient(v, t, pos_weight, want_grad=False)
            net.adam_step(lr)
            hist.append(loss)
        if log:
            log(f"epoch {ep}: mean loss {np.mean(hist[-len(volumes):]):.4f} ({time.perf_counter() - t0:.1f} s)")
    return hist


def learned_labels(values, net: SeparatrixNet, ctx, err=1e-3, gthr=1e-6, iters=4096, presimp=0.0, want_grad=True):
    """Accurate labels with the learned mask: (grad, labels, stats, mask fraction)."""
    m, cnt = net.mask(values)
    grad, lab, st = labels.ondemand_asc(values, err, gthr, iters, presimp, mask=m, want_grad=want_grad, ctx=ctx)
    return grad, lab, st, cnt / m.size
