"""Region boundary kinds on the GPU (msceer_region_boundary_kind) against the
oracle, and the Accurate3DMorseLabeling command-line drop-in: its `_int_`
labels are the accurate labelling's, its `_edges_` file their kinds; with
AscOrDsc = 1 it labels the negated volume, as the reference does."""
import os
import subprocess

import numpy as np
import pytest

import oracle_lib
from golden import load_ondemand
from msceer_amd import labels, native, synth

pytestmark = pytest.mark.gpu

CLI = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "msceer_amd", "cli")


@pytest.fixture(scope="module")
def ctx():
    c = native.Context(0)
    yield c
    c.close()


@pytest.mark.parametrize("seed", [1, 2])
def test_kinds_match_the_oracle_on_random_and_real_labels(ctx, seed):
    rng = np.random.default_rng(seed)
    shape = (9, 13, 17)
    lab = rng.integers(-1, 4, shape).astype(np.int32)  # many regions, some unlabelled vertices
    k = labels.region_boundary_kind(lab, ctx=ctx)
    assert (k.ravel() == oracle_lib.region_boundary_kind(lab, shape)).all()
    f = synth.smoothed_noise((40, 44, 48), 70 + seed, 2.0)
    _, real, _ = labels.ondemand_asc(f, want_grad=False, ctx=ctx)
    k = labels.region_boundary_kind(real, ctx=ctx)
    expect = oracle_lib.region_boundary_kind(real, f.shape)
    assert (k.ravel() == expect).all(), int((k.ravel() != expect).sum())
    assert (k == 3).any() or (k == 2).any()


@pytest.mark.parametrize("asc_or_dsc", [0, 1])
def test_accurate3dmorselabeling_cli(tmp_path, asc_or_dsc):
    exe = os.path.join(CLI, "accurate3dmorselabeling")
    if not os.path.exists(exe):
        pytest.fail(f"{exe} is not built (python -m msceer_amd.build)")
    f, (err, gthr, iters, presimp), _ = load_ondemand("gauss_44x36x40")
    Z, Y, X = f.shape
    raw = tmp_path / "vol.raw"
    f.tofile(raw)
    r = subprocess.run([exe, str(X), str(Y), str(Z), str(raw), repr(err), repr(gthr), str(iters), repr(presimp),
                        str(asc_or_dsc), "4"], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    lab = np.fromfile(f"{raw}._int_{X}x{Y}x{Z}.raw", np.int32)
    kinds = np.fromfile(f"{raw}._edges_{2 * X - 1}x{2 * Y - 1}x{2 * Z - 1}.raw", np.int8)
    g = -f if asc_or_dsc else f
    _, lab_api, _ = labels.ondemand_asc(g, err, gthr, iters, presimp, want_grad=False)
    assert (lab == lab_api.ravel()).all()
    assert (kinds == oracle_lib.region_boundary_kind(lab, f.shape)).all()
    assert open(f"{raw}.004.gtime.txt").read().count("\n") == 4
