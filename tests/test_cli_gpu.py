"""The steepest_lstar / ondemandaccurate command-line drop-ins on the GPU.

They take the reference programs' arguments (steepest_lstar.cxx:52-80,
ondemandaccurate.cxx:704-760) and write the same files."""
import os
import subprocess

import numpy as np
import pytest

from golden import load_lstar, load_ondemand
from msceer_amd import labels

CLI = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "msceer_amd", "cli")

pytestmark = pytest.mark.gpu


def _exe(name):
    p = os.path.join(CLI, name)
    if not os.path.exists(p):
        pytest.fail(f"{p} is not built (python -m msceer_amd.build)")
    return p


def _run(args, tmp_path):
    r = subprocess.run(args, cwd=tmp_path, capture_output=True, text=True, timeout=120)
    # the reference programs return 1 on success
    assert r.returncode == 1, r.stdout + r.stderr
    return r.stdout


def test_steepest_lstar_cli_writes_reference_grad(tmp_path):
    f, ref = load_lstar("noise_48x97x31")
    Z, Y, X = f.shape
    raw = tmp_path / "vol.raw"
    f.tofile(raw)
    out = _run([_exe("steepest_lstar"), str(X), str(Y), str(Z), str(raw), "0", "0", "0", "4"], tmp_path)
    g = np.fromfile(str(raw) + ".grad", np.uint8)
    assert g.size == ref.size and (g == ref).all()
    assert "index-0=" in out
    line = open(str(raw) + ".steep.gtime.txt").read().split()
    assert line[:3] == ["Total", "Overall", "0"]


def test_ondemandaccurate_cli_matches_library(tmp_path):
    f, (err, gthr, iters, presimp), _ = load_ondemand("smooth_34x33x32")
    Z, Y, X = f.shape
    raw = tmp_path / "vol.raw"
    f.tofile(raw)
    _run([_exe("ondemandaccurate"), str(X), str(Y), str(Z), str(raw), repr(err), repr(gthr), str(iters),
          "0", "0", "1", "0", repr(presimp), "8", "1"], tmp_path)
    g = np.fromfile(str(raw) + ".grad", np.uint8)
    lab = np.fromfile(f"{raw}._int_{X}x{Y}x{Z}.raw", np.int32)
    grad, lab_api, _ = labels.ondemand_asc(f, err, gthr, iters, presimp, want_grad=True)
    assert (g == grad.ravel()).all()
    assert (lab == lab_api.ravel()).all()
    tasks = open(str(raw) + ".008.gtime.txt").read().splitlines()
    assert len(tasks) == 5


def test_ondemandaccurate_cli_without_accuracy_writes_lstar_grad(tmp_path):
    f, ref = load_lstar("noise_48x97x31")
    Z, Y, X = f.shape
    raw = tmp_path / "vol.raw"
    f.tofile(raw)
    _run([_exe("ondemandaccurate"), str(X), str(Y), str(Z), str(raw), "1e-3", "1e-6", "100",
          "0", "0", "0", "0", "0"], tmp_path)
    g = np.fromfile(str(raw) + ".grad", np.uint8)
    assert (g == ref).all()


@pytest.mark.parametrize("tag", ["smooth_34x33x32_0001", "gauss_44x36x40_0011"])
def test_ondemandaccurate_cli_descending_matches_library_and_reference(tmp_path, tag):
    # needDSC2 (argv[11], ondemandaccurate.cxx:170): the .grad and the
    # restriction labels equal the library's; the descending labels split the
    # vertices as the reference's run does, up to the integration's
    # visiting-order mismatch
    import hashlib
    from golden import load_dsc2, same_partition
    f, (err, gthr, iters, presimp), flags, ref, grad_sha = load_dsc2(tag)
    Z, Y, X = f.shape
    raw = tmp_path / "vol.raw"
    f.tofile(raw)
    out = _run([_exe("ondemandaccurate"), str(X), str(Y), str(Z), str(raw), repr(err), repr(gthr), str(iters),
                *map(str, flags), repr(presimp), "8", "1"], tmp_path)
    assert "built maxgraph" in out
    g = np.fromfile(str(raw) + ".grad", np.uint8)
    r = np.fromfile(tmp_path / f"boundaries_{2 * X - 1}x{2 * Y - 1}x{2 * Z - 1}.raw", np.int8)
    ld = np.fromfile(f"{raw}.dsc._int_{X}x{Y}x{Z}.raw", np.int32)
    grad, _, restr, _, ld_api = labels.ondemand(f, *map(bool, flags[:3]), err, gthr, iters, presimp,
                                                need_dsc2=True, with_dsc=True)
    assert (g == grad).all() and (r == restr).all() and (ld == ld_api.ravel()).all()
    if tag.startswith("smooth_34"):  # no visiting-order mismatch on this case
        assert same_partition(ld, ref["cleaned"])
        assert (r == ref["restriction"]).all()
        assert hashlib.sha256(g.tobytes()).hexdigest() == grad_sha
    tasks = open(str(raw) + ".008.gtime.txt").read().splitlines()
    assert any(t.startswith("NumericalTracing NewIntegrationDSC") for t in tasks)


@pytest.mark.parametrize("flags", [(1, 0, 0), (1, 1, 1)])
def test_ondemandaccurate_cli_one_manifolds_match_library(tmp_path, flags):
    # needASC1 / needDSC1 argv positions (ondemandaccurate.cxx:167-168): the
    # .grad, the restriction labels (boundaries_MxMyMz.raw with outputdebug)
    # and the task lines of the branches that ran
    from golden import load_lines
    f, (err, gthr, iters, presimp), ref_r, _ = load_lines("gauss_44x36x40", flags)
    Z, Y, X = f.shape
    raw = tmp_path / "vol.raw"
    f.tofile(raw)
    out = _run([_exe("ondemandaccurate"), str(X), str(Y), str(Z), str(raw), repr(err), repr(gthr), str(iters),
                str(flags[0]), str(flags[1]), str(flags[2]), "0", repr(presimp), "8", "1"], tmp_path)
    assert "1-manifold streamlines" in out
    g = np.fromfile(str(raw) + ".grad", np.uint8)
    r = np.fromfile(tmp_path / f"boundaries_{2 * X - 1}x{2 * Y - 1}x{2 * Z - 1}.raw", np.int8)
    grad, _, restr, _ = labels.ondemand(f, *map(bool, flags), err, gthr, iters, presimp)
    assert (g == grad).all() and (r == restr).all()
    tasks = open(str(raw) + ".008.gtime.txt").read().splitlines()
    assert len(tasks) == (6 if flags[2] else 3)
    assert any(t.startswith("Numerical SaddleInt") for t in tasks)
