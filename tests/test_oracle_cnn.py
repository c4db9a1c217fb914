"""The CPU oracle of the near-separatrix classifier, pinned before it judges
the GPU kernels: the windowed normalisation against a direct computation, the
analytic gradient against central finite differences (unrounded mode) along
random directions in every parameter block, and the BF16-rounded mode against
the unrounded one."""
import numpy as np
import pytest

import oracle_lib
from msceer_amd import synth

BLOCKS = {  # name -> (offset, size), the flat parameter layout
    "W1": (0, 432), "b1": (432, 16), "W2": (448, 13824), "b2": (14272, 32),
    "W3": (14304, 13824), "b3": (28128, 16), "wh": (28144, 16), "bh": (28160, 1),
}


def _params(seed):
    s = synth.Stream(seed)
    p = s.normal(oracle_lib.cnn_num_params()).astype(np.float32)
    for name, fan in (("W1", 27), ("W2", 27 * 16), ("W3", 27 * 32), ("wh", 16)):
        o, n = BLOCKS[name]
        p[o:o + n] *= np.float32(np.sqrt(2.0 / fan))
    for name in ("b1", "b2", "b3"):
        o, n = BLOCKS[name]
        p[o:o + n] = np.float32(0.05)
    p[28160] = np.float32(-0.5)
    return p


def _volume(shape, seed):
    return synth.smoothed_noise(shape, seed, 1.5).astype(np.float32)


def test_normalisation_matches_direct_window_statistics():
    f = _volume((7, 9, 11), 3)
    n = oracle_lib.cnn_normalize(f)
    Z, Y, X = f.shape
    for (z, y, x) in [(0, 0, 0), (3, 4, 5), (6, 8, 10), (2, 0, 9)]:
        w = f[max(z - 4, 0):z + 5, max(y - 4, 0):y + 5, max(x - 4, 0):x + 5].astype(np.float64)
        expect = (float(f[z, y, x]) - w.mean()) / (w.std() + 1e-6)
        assert abs(n[z, y, x] - expect) < 1e-5 * max(1.0, abs(expect))


@pytest.mark.parametrize("block", list(BLOCKS))
def test_gradient_matches_finite_differences(block):
    f = _volume((5, 6, 7), 11)
    rng = np.random.default_rng(7)
    target = (rng.random(f.shape) < 0.3).astype(np.uint8)
    p = _params(5)
    loss, g = oracle_lib.cnn_gradient(p, f, target, pos_w=2.5, quant=0)
    o, n = BLOCKS[block]
    d = np.zeros_like(p)
    d[o:o + n] = rng.standard_normal(n).astype(np.float32)
    scale = 1e-3 * max(float(np.abs(p[o:o + n]).mean()), 1e-2) / float(np.abs(d[o:o + n]).mean())
    d *= np.float32(scale)
    lp, _ = oracle_lib.cnn_gradient(p + d, f, target, pos_w=2.5, quant=0)
    lm, _ = oracle_lib.cnn_gradient(p - d, f, target, pos_w=2.5, quant=0)
    fd = (lp - lm) / 2.0
    an = float(np.dot(g, (p + d).astype(np.float64) - p.astype(np.float64)))
    # 1%: float steps and the odd ReLU kink crossed by the step; a wrong tap,
    # transpose or gate gives errors of order one
    assert abs(fd - an) <= 1e-2 * abs(an) + 1e-9, (block, fd, an)


def test_rounded_mode_close_to_unrounded():
    f = _volume((6, 7, 8), 2)
    target = (np.random.default_rng(1).random(f.shape) < 0.2).astype(np.uint8)
    p = _params(9)
    l0, g0 = oracle_lib.cnn_gradient(p, f, target, 2.0, quant=0)
    l1, g1 = oracle_lib.cnn_gradient(p, f, target, 2.0, quant=1)
    assert abs(l0 - l1) < 1e-2 * l0
    cos = float(np.dot(g0, g1) / np.linalg.norm(g0) / np.linalg.norm(g1))
    assert cos > 0.995


def test_prediction_is_the_sigmoid_of_the_loss_model():
    # an all-zero target with unit weight: loss = mean(-log(1 - p))
    f = _volume((4, 5, 6), 4)
    p = _params(3)
    prob = oracle_lib.cnn_predict(p, f, quant=0)
    loss, _ = oracle_lib.cnn_gradient(p, f, np.zeros(f.shape, np.uint8), 1.0, quant=0)
    assert abs(loss - float(np.mean(-np.log1p(-prob)))) < 1e-9


def test_forward_matches_torch_conv3d():
    # an independent statement of the same network through torch's conv3d
    # (CPU, float64): catches a transposed weight layout, a wrong tap order or
    # a missing bias/ReLU in the oracle's own loops
    import torch
    import torch.nn.functional as F
    f = _volume((6, 7, 9), 21)
    p = _params(13).astype(np.float64)
    prob = oracle_lib.cnn_predict(p.astype(np.float32), f, quant=0).reshape(f.shape)
    n = torch.from_numpy(oracle_lib.cnn_normalize(f).astype(np.float64)).reshape(1, 1, *f.shape)

    def weight(name, cout, cin):
        # flat index (t * cout + co) * cin + ci, tap t = kx + 3 ky + 9 kz -> torch [co, ci, kz, ky, kx]
        o, size = BLOCKS[name]
        w = torch.from_numpy(p[o:o + size].copy()).reshape(3, 3, 3, cout, cin)
        return w.permute(3, 4, 0, 1, 2).contiguous()

    def bias(name):
        o, size = BLOCKS[name]
        return torch.from_numpy(p[o:o + size].copy())

    a1 = F.relu(F.conv3d(n, weight("W1", 16, 1), bias("b1"), padding=1))
    a2 = F.relu(F.conv3d(a1, weight("W2", 32, 16), bias("b2"), padding=1))
    a3 = F.relu(F.conv3d(a2, weight("W3", 16, 32), bias("b3"), padding=1))
    o, _ = BLOCKS["wh"]
    logit = torch.einsum("c,bczyx->bzyx", torch.from_numpy(p[o:o + 16].copy()), a3) + float(p[28160])
    expect = torch.sigmoid(logit)[0].numpy()
    assert np.abs(prob - expect).max() < 1e-9, float(np.abs(prob - expect).max())
