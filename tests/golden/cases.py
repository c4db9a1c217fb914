"""Input fields of the lower-star golden vectors (shared by the fixture writer and the tests).

Each case is regenerated bit-for-bit from msceer_amd.synth; the fixture stores
the SHA-256 of the input so a drifting generator is caught before any compare.
"""
import numpy as np

from msceer_amd import synth


def _normal(shape, seed):
    s = synth.Stream(seed)
    return s.normal(int(np.prod(shape))).reshape(shape).astype(np.float32)


def _ties(shape, seed, levels=3):
    s = synth.Stream(seed)
    return np.floor(s.uniform(int(np.prod(shape))) * levels).reshape(shape).astype(np.float32)


def _signed_zeros(shape, seed):
    # +0.0 / -0.0 / 1.0 mixture: +0 and -0 compare equal, so ties fall back to the index
    s = synth.Stream(seed)
    u = s.uniform(int(np.prod(shape))).reshape(shape)
    f = np.where(u < 0.4, np.float32(0.0), np.where(u < 0.8, np.float32(-0.0), np.float32(1.0)))
    return f.astype(np.float32)


def _ramp(shape, reverse=False):
    f = np.arange(int(np.prod(shape)), dtype=np.float32).reshape(shape)
    return (-f if reverse else f).astype(np.float32)


def _plateau(shape, seed):
    f = synth.smoothed_noise(shape, seed, 2.0)
    return (np.round(f * 8.0) / 8.0).astype(np.float32)


# name -> (shape_zyx, generator)
LSTAR_CASES = {
    "c1_gauss64": ((64, 64, 64), lambda: synth.lstar_small(101)),
    "noise_17x9x5": ((5, 9, 17), lambda: _normal((5, 9, 17), 11)),
    "noise_48x97x31": ((31, 97, 48), lambda: _normal((31, 97, 48), 12)),
    "smooth_60x50x40": ((40, 50, 60), lambda: synth.smoothed_noise((40, 50, 60), 13, 2.0)),
    "smooth_96": ((96, 96, 96), lambda: synth.smoothed_noise((96, 96, 96), 202, 2.0)),
    "const_8x7x6": ((6, 7, 8), lambda: np.zeros((6, 7, 8), np.float32)),
    "ties_10x11x12": ((12, 11, 10), lambda: _ties((12, 11, 10), 14)),
    "signed_zeros_9x8x7": ((7, 8, 9), lambda: _signed_zeros((7, 8, 9), 15)),
    "ramp_13x6x5": ((5, 6, 13), lambda: _ramp((5, 6, 13))),
    "ramp_rev_13x6x5": ((5, 6, 13), lambda: _ramp((5, 6, 13), True)),
    "plateau_33x29x21": ((21, 29, 33), lambda: _plateau((21, 29, 33), 16)),
    "tiny_2x2x2": ((2, 2, 2), lambda: _normal((2, 2, 2), 17)),
    "tiny_5x2x3": ((3, 2, 5), lambda: _normal((3, 2, 5), 18)),
    "slab_64x64x2": ((2, 64, 64), lambda: _normal((2, 64, 64), 19)),
}


# ondemandaccurate (ascending 3-manifolds) golden cases:
# name -> (shape_zyx, generator, err_thr, grad_thr, max_iters, presimp)
ONDEMAND_CASES = {
    "labels_48": ((48, 48, 48), lambda: synth.labels_field(303, 48), 1e-3, 1e-6, 4096, 0.0),
    "gauss_44x36x40": ((40, 36, 44), lambda: synth.gaussian_mixture((40, 36, 44), 7, 20, 3.0, 8.0), 1e-3, 1e-6, 4096, 0.0),
    "smooth_34x33x32": ((32, 33, 34), lambda: synth.smoothed_noise((32, 33, 34), 8, 2.0), 1e-3, 1e-6, 4096, 0.0),
    "smooth_presimp_40": ((40, 40, 40), lambda: synth.smoothed_noise((40, 40, 40), 9, 2.5), 1e-3, 1e-6, 4096, 0.05),
    # a tiny step budget forces OVER_MAX_ITERATIONS stops
    "budget_24x30x20": ((20, 30, 24), lambda: synth.gaussian_mixture((20, 30, 24), 31, 12, 4.0, 9.0, sine=0.05),
                        1e-3, 1e-6, 60, 0.0),
}
