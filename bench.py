#!/usr/bin/env python3
"""Benchmark: steepest_lstar .grad (bit-exact lower-star gradient) on one MI355X.

Workload: 512x512x512 float32, N(0,1) white noise smoothed by a Gaussian
(sigma 2 voxels, seed 202).  One step = one full .grad (lower-star pairing +
ldir passes) computed from values already resident in HBM into a device
buffer.  Prints one JSON line.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--size N] [--no-cpu]
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Mvert/s: lower-star grad (bit-exact); accurate labels plain vs learned"
HBM_PEAK_GBS = 8000.0


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def read_traffic(name: str):
    """HBM bytes per launch from a committed rocprofv3 PMC summary (profiles/<name>)."""
    path = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch="):
            return float(line.split("=", 1)[1])
    return None


REF_DIR = os.path.join(ROOT, "oracle", "_ref")


def _run_reference(prog, args, cores):
    """Run a reference driver (the reference's own sources built by
    oracle/Makefile) and return its REFTIME fields, or None if it is absent."""
    import subprocess
    exe = os.path.join(REF_DIR, prog)
    if not os.path.exists(exe):
        return None
    env = dict(os.environ, OMP_NUM_THREADS=str(cores))
    r = subprocess.run([exe, *map(str, args), str(cores)], capture_output=True, text=True, env=env, timeout=600)
    for line in r.stdout.splitlines():
        if line.startswith("REFTIME"):
            return {k: float(v) for k, v in (kv.split("=") for kv in line.split()[1:])}
    return None


def cpu_baseline(f: np.ndarray, slab: int):
    """The reference's own lower-star gradient (maxv + pairing + ldir, its
    OpenMP code built from its sources) on a slab of the same volume, all host
    cores; the oracle restatement where the reference build is absent."""
    import tempfile
    cores = _cores()
    sample = np.ascontiguousarray(f[:slab])
    Z, Y, X = sample.shape
    with tempfile.TemporaryDirectory() as tmp:
        raw = os.path.join(tmp, "slab.raw")
        sample.tofile(raw)
        t = _run_reference("ref_steepest", [X, Y, Z, raw, os.path.join(tmp, "out")], cores)
    if t is not None:
        dt = t["total_ms"] * 1e-3
        return {"value": round(X * Y * Z / dt / 1e6, 3), "unit": "Mvert/s", "cores": cores, "kind": "reference",
                "sample": f"{X}x{Y}x{Z} slab (first {Z} z-planes) of the benchmark volume; the reference's "
                          f"RegularGridMaxMinVertexLabeling3D + SlidingWindowRobinsNoalloc + "
                          f"setAscendingManifoldDimensions (OpenMP, built from its sources), {dt:.2f} s "
                          f"(maxv {t['maxv_ms']:.0f} ms, pairing {t['pair_ms']:.0f} ms, ldir {t['ldir_ms']:.0f} ms)"}
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib  # noqa: E402  (checker / baseline only)

    oracle_lib.set_threads(cores)
    oracle_lib.steepest_grad(sample[:8])  # warm the thread pool
    t0 = time.perf_counter()
    oracle_lib.steepest_grad(sample)
    dt = time.perf_counter() - t0
    Z, Y, X = sample.shape
    return {"value": round(X * Y * Z / dt / 1e6, 3), "unit": "Mvert/s", "cores": cores, "kind": "port",
            "sample": f"{X}x{Y}x{Z} slab (first {Z} z-planes) of the benchmark volume, oracle "
                      f"max/min labelling + lower-star pairing + ldir + flags, {dt:.2f} s"}


def labels_cpu_baseline(f: np.ndarray, slab: int, params):
    """The reference's own accurate labelling (its OpenMP ondemandaccurate
    needASC2 pipeline, built from its sources) on a z-slab of the labelling
    volume, all host cores; the oracle restatement where the reference build
    is absent."""
    import tempfile
    cores = _cores()
    sample = np.ascontiguousarray(f[:slab])
    Z, Y, X = sample.shape
    err, gthr, iters, presimp = params
    with tempfile.TemporaryDirectory() as tmp:
        raw = os.path.join(tmp, "slab.raw")
        sample.tofile(raw)
        t = _run_reference("ref_ondemand", [X, Y, Z, raw, os.path.join(tmp, "out"), err, gthr, iters, presimp], cores)
    if t is not None:
        dt = t["total_ms"] * 1e-3
        return {"value": round(X * Y * Z / dt / 1e6, 3), "unit": "Mvert/s", "cores": cores, "kind": "reference",
                "sample": f"{X}x{Y}x{Z} slab (first {Z} z-planes) of the labelling volume; the reference's "
                          f"needASC2 pipeline (OpenMP, built from its sources), {dt:.2f} s "
                          f"(integration {t['integrate_ms']:.0f} ms)"}
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib  # noqa: E402  (checker / baseline only)

    oracle_lib.set_threads(cores)
    t0 = time.perf_counter()
    oracle_lib.ondemand_asc(sample, *params, mode=0)
    dt = time.perf_counter() - t0
    Z, Y, X = sample.shape
    return {"value": round(X * Y * Z / dt / 1e6, 3), "unit": "Mvert/s", "cores": cores, "kind": "port",
            "sample": f"{X}x{Y}x{Z} slab (first {Z} z-planes) of the labelling volume, oracle accurate "
                      f"ascending pipeline in the reference's mixed precision, {dt:.2f} s"}


def _cores():
    cores = os.cpu_count() or 1
    try:
        cores = len(os.sched_getaffinity(0))
    except Exception:
        pass
    return cores


def bench_labels(args):
    """Accurate ascending labels + conforming .grad of the 256^3 labelling
    workload, values resident in HBM; label mismatch against the committed FP64
    oracle labels."""
    sys.path.insert(0, os.path.join(ROOT, "tests", "golden"))
    from golden import C3_PARAMS, load_c3_fp64  # committed fixture (data)
    from msceer_amd import hipmem, native, synth

    log("generating 256^3 labelling field (200 Gaussians + sine, seed 303)")
    f = synth.labels_field(303, 256)
    ref, sha = load_c3_fp64()
    exact, sha2 = load_c3_fp64(exact=True)
    assert synth.sha256(f) == sha == sha2, "labelling field generator drifted from the fixtures"
    Z, Y, X = f.shape
    V = X * Y * Z
    cells = (2 * X - 1) * (2 * Y - 1) * (2 * Z - 1)
    err, gthr, iters, presimp = C3_PARAMS
    ctx = native.Context(0)
    lib = native.lib()
    dv = hipmem.DeviceBuffer(f.nbytes)
    dv.upload(f)
    dg = hipmem.DeviceBuffer(cells)
    dl = hipmem.DeviceBuffer(4 * V)
    st = native.LabelStats()

    def run(mask_addr=None, extra=0):
        native.check(lib.msceer_ondemand_asc(ctx.handle, ctypes.c_void_p(dv.addr), X, Y, Z, err, gthr, iters,
                                             presimp, ctypes.c_void_p(mask_addr) if mask_addr else None,
                                             ctypes.c_void_p(dg.addr), ctypes.c_void_p(dl.addr),
                                             ctypes.byref(st), native.DEVICE_PTRS | extra), "msceer_ondemand_asc")

    for _ in range(args.label_warmup):
        run()
    hipmem.synchronize()
    stages = []
    t0 = time.perf_counter()
    for _ in range(args.label_steps):
        run()
        stages.append(st.as_dict())
    hipmem.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / args.label_steps
    lab = dl.download(np.int32, V)
    mism = float((lab != ref).mean())
    last = stages[-1]
    stage = {k[3:]: round(float(np.mean([d[k] for d in stages])), 3)
             for k in ("ms_lstar", "ms_regions", "ms_integrate", "ms_pass1", "ms_pass3", "ms_cleanup", "ms_grad",
                       "ms_total")}
    # algorithmic bytes of the labelling: 4 B read + 4 B label written per vertex
    integ_gbs = 8.0 * V / (stage["integrate"] * 1e-3) / 1e9
    plain = {"value": round(V / (ms * 1e-3) / 1e6, 2), "unit": "Mvert/s", "ms_per_step": round(ms, 3),
             "steps": args.label_steps, "mismatch_vs_fp64_oracle": mism,
             "mismatch_vs_exact_fp64_oracle": float((lab != exact).mean()),
             "oracle_fp64_vs_exact": float((ref != exact).mean()),
             "roofline": {"bound": "hbm", "achieved": round(integ_gbs, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                          "frac": round(integ_gbs / HBM_PEAK_GBS, 6),
                          "kernel": "integration (pass1_kernel levels + pass2 + pass3_kernel<false>)",
                          "algorithmic_bytes": 8.0 * V, "kernel_ms": stage["integrate"],
                          "pass3_ms": stage["pass3"],
                          "certain_rounds_share_of_total": round(stage["regions"] / stage["total"], 4),
                          "note": "latency-bound: dependent streamline steps and certain-region levels, "
                                  "not bytes"},
             "stage_ms": stage,
             "counts": {k: int(last[k]) for k in ("minima", "regions", "traced_fp32", "over_budget_fp32",
                                                  "reintegrated_pass3", "escalated_fp64", "reintegrated_fp64",
                                                  "over_budget_fp64", "certain_sweeps", "pass3_rounds",
                                                  "steps_fp32", "steps_fp64", "ahead_fp64")},
             "fp64_share_of_steps": round(float(last["steps_fp64"]) /
                                          max(1.0, float(last["steps_fp32"] + last["steps_fp64"])), 4)}
    # the FP32-first precision policy on the same volume (MSCEER_FP32_FIRST)
    run(extra=native.FP32_FIRST)
    hipmem.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.label_steps):
        run(extra=native.FP32_FIRST)
    hipmem.synchronize()
    ms32 = (time.perf_counter() - t0) * 1e3 / args.label_steps
    lab32 = dl.download(np.int32, V)
    s32 = st.as_dict()
    plain["fp32_first"] = {
        "ms_per_step": round(ms32, 3), "value": round(V / (ms32 * 1e-3) / 1e6, 2), "unit": "Mvert/s",
        "policy": "pass 3 in FP32 (voxel-local coordinates); FP64 re-trace of the streamlines whose FP32 trace meets "
                  "a near-tie decision (2^-16 voxel of a voxel face or half-integer; step error or gradient within "
                  "2^-14 relative of its threshold) or uses up its step budget",
        "mismatch_vs_fp64_oracle": float((lab32 != ref).mean()),
        "mismatch_vs_exact_fp64_oracle": float((lab32 != exact).mean()),
        "pass3_ms": round(float(s32["ms_pass3"]), 3), "escalated_fp64": int(s32["escalated_fp64"]),
        "reintegrated_pass3": int(s32["reintegrated_pass3"]),
        "steps_fp32": int(s32["steps_fp32"]), "steps_fp64": int(s32["steps_fp64"]),
        "fp64_share_of_steps": round(float(s32["steps_fp64"]) /
                                     max(1.0, float(s32["steps_fp32"] + s32["steps_fp64"])), 4)}
    if args.ideal_mask_probe:
        from scipy import ndimage
        zero = hipmem.DeviceBuffer(V)
        zero.upload(np.zeros(V, np.uint8))
        run(mask_addr=zero.addr)
        comb = dl.download(np.int32, V).reshape(Z, Y, X)
        pl = lab.reshape(Z, Y, X)
        bnd = np.zeros_like(pl, bool)
        for ax in range(3):
            d = np.diff(pl, axis=ax) != 0
            sl0 = [slice(None)] * 3
            sl1 = [slice(None)] * 3
            sl0[ax] = slice(0, -1)
            sl1[ax] = slice(1, None)
            bnd[tuple(sl0)] |= d
            bnd[tuple(sl1)] |= d
        core = bnd | (comb != pl)
        for margin in (1, 2):
            m = ndimage.binary_dilation(core, structure=np.ones((3, 3, 3), bool), iterations=margin)
            dm = hipmem.DeviceBuffer(V)
            dm.upload(m.astype(np.uint8).ravel())
            run(mask_addr=dm.addr)
            hipmem.synchronize()
            t0 = time.perf_counter()
            run(mask_addr=dm.addr)
            hipmem.synchronize()
            ms_m = (time.perf_counter() - t0) * 1e3
            labm = dl.download(np.int32, V)
            plain[f"ideal_mask_probe_{margin}"] = {
                "ms": round(ms_m, 3), "mask_fraction": float(m.mean()), "core_fraction": float(core.mean()),
                "comb_mismatch_vs_plain": float((comb != pl).mean()),
                "mismatch_vs_fp64_oracle": float((labm != ref).mean()),
                "stage_ms": {k[3:]: round(float(st.as_dict()[k]), 3) for k in
                             ("ms_lstar", "ms_regions", "ms_integrate", "ms_cleanup", "ms_grad")},
                "fp64": int(st.reintegrated_fp64)}
    lines_out = None if args.no_lines else bench_lines(args, ctx, lib, dv, dg, f, C3_PARAMS)
    dsc_out = None if args.no_dsc else bench_dsc(args, ctx, lib, dv, dg, f, C3_PARAMS)
    learned_out = None
    if not args.no_learned:
        learned_out = bench_learned(args, ctx, lib, dv, dg, dl, st, X, Y, Z, ref, plain, run)
    base = None
    if not args.no_cpu:
        log("timing the labelling CPU baseline")
        base = labels_cpu_baseline(f, args.label_cpu_slab, C3_PARAMS)
    ctx.close()
    return {"workload": f"ondemandaccurate needASC2 + conforming .grad, {X}x{Y}x{Z} float32, 200 Gaussians "
                        f"(sigma 4-16) + 0.05 sin sin sin, seed 303; err 1e-3 (squared error), grad 1e-6, "
                        f"4096 steps, presimp 0; values resident in HBM",
            "plain": plain, "lines": lines_out, "descending": dsc_out, "learned": learned_out, "cpu_baseline": base}


CNN_FLOP_PER_VOXEL = 2 * (27 * 16 + 27 * 16 * 32 + 27 * 32 * 16 + 16)  # 56 288


def bench_lines(args, ctx, lib, dv, dg, f, params):
    """Accurate 1-manifolds on the labelling volume: OnDemandAccurate with
    needASC1 + needDSC1 (and with needASC2 as well), values resident in HBM;
    the 1-manifold stage timed apart.  CPU: the oracle's restatement of the
    reference's saddle loops (one thread, as the reference's order is one
    thread's; its own build cannot be optimised, see DESIGN), the loop alone."""
    from msceer_amd import hipmem, native

    Z, Y, X = f.shape
    V = X * Y * Z
    cells = (2 * X - 1) * (2 * Y - 1) * (2 * Z - 1)
    err, gthr, iters, presimp = params
    dr = hipmem.DeviceBuffer(cells)
    dl = hipmem.DeviceBuffer(4 * V)
    st = native.LabelStats()

    def run(need):
        native.check(lib.msceer_ondemand(ctx.handle, ctypes.c_void_p(dv.addr), X, Y, Z, err, gthr, iters, presimp,
                                         need, None, ctypes.c_void_p(dg.addr),
                                         ctypes.c_void_p(dl.addr) if need & native.NEED_ASC2 else None,
                                         ctypes.c_void_p(dr.addr), ctypes.byref(st), native.DEVICE_PTRS),
                     "msceer_ondemand")

    out = {}
    for key, need in (("asc1_dsc1", native.NEED_ASC1 | native.NEED_DSC1),
                      ("asc1_dsc1_asc2", native.NEED_ASC1 | native.NEED_DSC1 | native.NEED_ASC2)):
        run(need)
        hipmem.synchronize()
        t0 = time.perf_counter()
        stats = []
        for _ in range(args.label_steps):
            run(need)
            stats.append(st.as_dict())
        hipmem.synchronize()
        ms = (time.perf_counter() - t0) * 1e3 / args.label_steps
        d = stats[-1]
        out[key] = {"value": round(V / (ms * 1e-3) / 1e6, 2), "unit": "Mvert/s", "ms_per_step": round(ms, 3),
                    "lines_ms": round(float(np.mean([x["ms_lines"] for x in stats])), 3),
                    "streamlines": {"asc": int(d["lines_asc"]), "dsc": int(d["lines_dsc"]),
                                    "replayed": int(d["lines_replayed"]), "windows": int(d["line_windows"]),
                                    "replay_cuts": int(d["line_cuts"]), "overflow_cuts": int(d["line_overflows"]),
                                    "traced_alone": int(d["line_alone"])}}
    r = dr.download(np.int8, cells)
    out["cells_4"], out["cells_5"] = int((r == 4).sum()), int((r == 5).sum())
    out["roofline"] = {"bound": "hbm", "achieved": None, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": None,
                       "note": "latency-bound: streamlines of a few voxels each, in the saddles' order; "
                               "the stage is speculation, in-order replays of interacting lines and host "
                               "read-backs per window"}
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        import oracle_lib  # noqa: E402  (checker / baseline only)

        log("timing the 1-manifold CPU baseline (oracle saddle loops, one thread)")
        oracle_lib.set_threads(_cores())
        base = oracle_lib.lstar_pairs(f)
        us = 0
        for asc in (1, 0):
            _, ost = oracle_lib.saddle_lines(f, base, asc, err=err, gthr=gthr, iters=iters, mode=0)
            us += int(ost[7])
        out["cpu_baseline"] = {"value": round(V / (us * 1e-6) / 1e6, 3), "unit": "Mvert/s", "cores": 1,
                               "kind": "port", "seconds": round(us * 1e-6, 3),
                               "sample": f"the whole {X}x{Y}x{Z} volume: saddle gathering, sorting and the "
                                         f"digitizing streamlines of both families (ASC, DSC), reference "
                                         f"arithmetic, given the steepest pairs and the gradient field"}
        out["gpu_lines_vs_cpu_loop"] = round(us * 1e-3 / out["asc1_dsc1"]["lines_ms"], 2)
    return out


def bench_dsc(args, ctx, lib, dv, dg, f, params):
    """Descending 3-manifolds (needDSC2) on the labelling volume, alone and
    with needASC2 (the reference's BOTH mode: one extremum graph build, the
    gradient field shared), values resident in HBM; the descending branch
    timed apart.  CPU: the reference's own needASC2 + needDSC2 pipeline
    (OpenMP, built from its sources) on a z-slab."""
    from msceer_amd import hipmem, native

    Z, Y, X = f.shape
    V = X * Y * Z
    cells = (2 * X - 1) * (2 * Y - 1) * (2 * Z - 1)
    err, gthr, iters, presimp = params
    dr = hipmem.DeviceBuffer(cells)
    dla = hipmem.DeviceBuffer(4 * V)
    dld = hipmem.DeviceBuffer(4 * V)
    st = native.LabelStats()

    def run(need):
        native.check(lib.msceer_ondemand2(ctx.handle, ctypes.c_void_p(dv.addr), X, Y, Z, err, gthr, iters, presimp,
                                          need, None, ctypes.c_void_p(dg.addr),
                                          ctypes.c_void_p(dla.addr) if need & native.NEED_ASC2 else None,
                                          ctypes.c_void_p(dld.addr), ctypes.c_void_p(dr.addr), ctypes.byref(st),
                                          native.DEVICE_PTRS), "msceer_ondemand2")

    out = {}
    for key, need in (("dsc2", native.NEED_DSC2), ("asc2_dsc2", native.NEED_ASC2 | native.NEED_DSC2)):
        run(need)
        hipmem.synchronize()
        t0 = time.perf_counter()
        stats = []
        for _ in range(args.label_steps):
            run(need)
            stats.append(st.as_dict())
        hipmem.synchronize()
        ms = (time.perf_counter() - t0) * 1e3 / args.label_steps
        d = stats[-1]
        out[key] = {"value": round(V / (ms * 1e-3) / 1e6, 2), "unit": "Mvert/s", "ms_per_step": round(ms, 3),
                    "dsc_ms": round(float(np.mean([x["ms_dsc"] for x in stats])), 3),
                    "maxima": int(d["maxima"]), "saddles2": int(d["saddles2"]),
                    "regions_dsc": int(d["regions_dsc"]), "certain_sweeps_dsc": int(d["certain_sweeps_dsc"]),
                    "stage_ms": {k[3:]: round(float(np.mean([x[k] for x in stats])), 3) for k in
                                 ("ms_lstar", "ms_regions", "ms_integrate", "ms_cleanup", "ms_dsc", "ms_dsc_wait",
                                  "ms_grad", "ms_total")}}
    ld = dld.download(np.int32, V)
    out["unlabelled_fraction"] = float((ld < 0).mean())
    if not args.no_cpu:
        import tempfile
        cores = _cores()
        slab = max(8, args.label_cpu_slab // 2)
        sample = np.ascontiguousarray(f[:slab])
        sz, sy, sx = sample.shape
        log("timing the needASC2 + needDSC2 CPU baseline (reference pipeline on a slab)")
        with tempfile.TemporaryDirectory() as tmp:
            raw = os.path.join(tmp, "slab.raw")
            sample.tofile(raw)
            t = _run_reference("ref_ondemand", [sx, sy, sz, raw, os.path.join(tmp, "out"), err, gthr, iters, presimp,
                                                cores, 0, 0, 1, 1], cores)
        if t is not None:
            dt = t["total_ms"] * 1e-3
            out["cpu_baseline"] = {
                "value": round(sx * sy * sz / dt / 1e6, 3), "unit": "Mvert/s", "cores": cores, "kind": "reference",
                "sample": f"{sx}x{sy}x{sz} slab (first {sz} z-planes) of the labelling volume; the reference's "
                          f"needASC2 + needDSC2 pipeline (OpenMP, built from its sources), {dt:.2f} s "
                          f"(integrations {t['integrate_ms']:.0f} ms)"}
            out["gpu_vs_cpu"] = round(out["asc2_dsc2"]["value"] / out["cpu_baseline"]["value"], 1)
    return out


def bench_learned(args, ctx, lib, dv, dg, dl, st, X, Y, Z, ref, plain, run):
    """Learned mask (trained weights, msceer_amd/weights) + masked labelling."""
    from msceer_amd import hipmem, learned, native

    if not os.path.exists(learned.WEIGHTS):
        return {"error": f"{learned.WEIGHTS} missing (python -m msceer_amd.train_sepnet)"}
    V = X * Y * Z
    net = learned.SeparatrixNet(ctx)
    meta = net.load()
    dm = hipmem.DeviceBuffer(V)
    cnt = ctypes.c_int64()

    def mask():
        native.check(lib.msceer_cnn_mask(ctx.handle, net._h, ctypes.c_void_p(dv.addr), X, Y, Z,
                                         float(net.threshold), int(net.dilation), ctypes.c_void_p(dm.addr),
                                         ctypes.byref(cnt), native.DEVICE_PTRS), "msceer_cnn_mask")

    mask()
    run(mask_addr=dm.addr)
    hipmem.synchronize()
    cnn_ms, tot_ms = [], []
    for _ in range(args.label_steps):
        t0 = time.perf_counter()
        mask()
        t1 = time.perf_counter()
        run(mask_addr=dm.addr)
        hipmem.synchronize()
        t2 = time.perf_counter()
        cnn_ms.append((t1 - t0) * 1e3)
        tot_ms.append((t2 - t0) * 1e3)
    lab = dl.download(np.int32, V)
    ms = float(np.mean(tot_ms))
    c_ms = float(np.mean(cnn_ms))
    tflops = CNN_FLOP_PER_VOXEL * V / (c_ms * 1e-3) / 1e12
    net.close()
    return {"value": round(V / (ms * 1e-3) / 1e6, 2), "unit": "Mvert/s", "ms_per_step": round(ms, 3),
            "speedup_vs_plain": round(plain["ms_per_step"] / ms, 3),
            "mismatch_vs_fp64_oracle": float((lab != ref).mean()),
            "mask_fraction": cnt.value / V, "threshold": net.threshold, "dilation": net.dilation,
            "cnn_ms": round(c_ms, 3),
            "cnn_roofline": {"bound": "mfma", "achieved": round(tflops, 2), "peak": 2500.0, "unit": "TFLOP/s",
                             "frac": round(tflops / 2500.0, 5),
                             "note": "normalisation + 3 convolutions + head + threshold + dilation, "
                                     f"{CNN_FLOP_PER_VOXEL} flop/voxel"},
            "integrated_vertices": int(st.integrated_vertices),
            "training": {"volumes": int(meta["volumes"]), "size": int(meta["size"]), "epochs": int(meta["epochs"]),
                         "seconds_on_gpu": round(float(meta["train_seconds"]), 2)}}


def _batch_volume(i):
    from msceer_amd import synth
    return synth.batch_volume(i).astype(np.float32)


def _percentiles(ms):
    ms = np.asarray(ms, np.float64)
    return {"p50": round(float(np.percentile(ms, 50)), 3), "p90": round(float(np.percentile(ms, 90)), 3),
            "p99": round(float(np.percentile(ms, 99)), 3), "max": round(float(ms.max()), 3)}


def bench_batch(args):
    """512 independent volumes (sides 32..160, smoothed noise / Gaussian
    mixtures): the lower-star .grad of all of them through msceer_batch_grad
    (one launch per stage per group of volumes, replayed from a hipGraph;
    values resident in HBM, and with host buffers in and out), through the
    worker-thread batch (msceer_batch), and the accurate labels + conforming
    .grad of all through msceer_batch; aggregate vertices/s and per-volume
    time percentiles."""
    from concurrent.futures import ProcessPoolExecutor

    from msceer_amd import batch, hipmem, native

    n = args.batch_volumes
    log(f"generating {n} batch volumes")
    with ProcessPoolExecutor(max_workers=min(16, _cores())) as ex:
        vols = list(ex.map(_batch_volume, range(n), chunksize=8))
    V = int(sum(v.size for v in vols))
    ctx = native.Context(0)
    lib = native.lib()
    err, gthr, iters, presimp = 1e-3, 1e-6, 4096, 0.0
    out = {"workload": f"{n} volumes, sides uniform 32..160 per axis (seed 4000+i), 50/50 smoothed noise "
                       f"(sigma 1.5-3) / Gaussian mixtures (sigma 4-16, sine term)",
           "vertices": V}

    # batched graph path, device-resident values and .grad
    al = lambda b: (int(b) + 255) & ~255  # noqa: E731
    voff, goff, vo, go = [], [], 0, 0
    for v in vols:
        Z, Y, X = v.shape
        voff.append(vo)
        goff.append(go)
        vo += al(v.nbytes)
        go += al((2 * X - 1) * (2 * Y - 1) * (2 * Z - 1))
    dv, dg = hipmem.DeviceBuffer(vo), hipmem.DeviceBuffer(go)
    for v, o in zip(vols, voff):
        dv.upload_at(o, v)
    arr = batch.volume_table([v.shape for v in vols], [dv.addr + o for o in voff], [dg.addr + o for o in goff])

    def graph_call(flags, cap=0):
        native.check(lib.msceer_batch_grad(ctx.handle, arr, n, cap, native.DEVICE_PTRS | flags), "msceer_batch_grad")
        st = native.BatchStats()
        native.check(lib.msceer_last_batch_stats(ctx.handle, ctypes.byref(st)), "msceer_last_batch_stats")
        return st

    def graph_leg(cap, reps):
        t0 = time.perf_counter()
        st0 = graph_call(0, cap)  # tuning, capture
        first = time.perf_counter() - t0
        graph_call(0, cap)
        hipmem.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            st = graph_call(0, cap)
        hipmem.synchronize()
        wall = (time.perf_counter() - t0) / reps
        graph_call(native.TIMING, cap)  # capture of the timed variant
        stt = graph_call(native.TIMING, cap)
        ms = np.array([arr[i].ms for i in range(n)], np.float64)
        return {"value": round(V / wall / 1e6, 2), "unit": "Mvert/s", "ms_per_call": round(wall * 1e3, 3),
                "calls": reps, "first_call_s": round(first, 3), "groups": st.groups,
                "max_group_verts": cap, "replays_per_call": st.replays, "captures_in_timed_calls": st.captures,
                "eager_rounds_first_call": st0.eager_rounds, "eager_rounds": st.eager_rounds,
                "rounds_per_pass": st.rounds, "gpu_ms": round(st.ms_gpu, 3),
                "launch": {"brick": native.BRICKS[st.shape], "persistent_blocks": st.blocks, "chunk": st.chunk},
                "per_volume_ms": _percentiles(ms),
                "per_volume_ms_is": "device clock, from the first brick of the volume's group to its last "
                                    "finished tile (timed call)",
                "gpu_ms_timed_call": round(stt.ms_gpu, 3)}

    log("batched .grad, one group, graph replays")
    out["grad_graph"] = graph_leg(0, args.batch_calls)
    if args.batch_graph_only:
        ctx.close()
        return out
    log("batched .grad, groups of at most 2^25 vertices")
    out["grad_graph_groups"] = graph_leg(1 << 25, 3)
    dv.free()
    dg.free()

    log("batched .grad, host buffers")
    batch.run_batch_grad(vols, ctx=ctx)
    t0 = time.perf_counter()
    _, _, st = batch.run_batch_grad(vols, ctx=ctx)
    wall = time.perf_counter() - t0
    out["grad_graph_host_buffers"] = {"value": round(V / wall / 1e6, 2), "unit": "Mvert/s",
                                      "seconds": round(wall, 3), "gpu_ms": round(st.ms_gpu, 3),
                                      "note": "PCIe-inclusive: values copied in and .grad out per call"}

    def timed(labels, streams):
        t0 = time.perf_counter()
        _, _, ms = batch.run_batch(vols, labels=labels, streams=streams, err=err, gthr=gthr, iters=iters,
                                   presimp=presimp, ctx=ctx)
        return time.perf_counter() - t0, ms

    def summary(wall, ms, streams):
        return {"value": round(V / wall / 1e6, 2), "unit": "Mvert/s", "seconds": round(wall, 3), "streams": streams,
                "note": "host buffers in and out", "per_volume_ms": _percentiles(ms)}

    k = args.batch_streams
    log("worker-thread batch (.grad)")
    timed(False, k)  # brick tuning for every x extent, buffers grown
    out["grad_threads"] = summary(*timed(False, k), k)
    out["grad_one_stream"] = summary(*timed(False, 1), 1)
    log("worker-thread batch (labels)")
    timed(True, k)
    out["labels"] = summary(*timed(True, k), k)
    ctx.close()
    return out


def bench_learned_large(args):
    """The learned mask on a 1024x768x512 smoothed-noise volume (sigma 2.5,
    seed 5999) against the plain full GPU integrator: speed and label
    mismatch (learned vs plain)."""
    from msceer_amd import hipmem, learned, native, synth

    log("generating 1024x768x512 smoothed noise (sigma 2.5, seed 5999)")
    f = synth.smoothed_noise((512, 768, 1024), 5999, 2.5)
    Z, Y, X = f.shape
    V = X * Y * Z
    cells = (2 * X - 1) * (2 * Y - 1) * (2 * Z - 1)
    ctx = native.Context(0)
    lib = native.lib()
    dv = hipmem.DeviceBuffer(f.nbytes)
    dv.upload(f)
    del f
    dg = hipmem.DeviceBuffer(cells)
    dl = hipmem.DeviceBuffer(4 * V)
    dm = hipmem.DeviceBuffer(V)
    st = native.LabelStats()
    err, gthr, iters, presimp = 1e-3, 1e-6, 4096, 0.0

    def run(mask=None):
        native.check(lib.msceer_ondemand_asc(ctx.handle, ctypes.c_void_p(dv.addr), X, Y, Z, err, gthr, iters, presimp,
                                             ctypes.c_void_p(mask) if mask else None, ctypes.c_void_p(dg.addr),
                                             ctypes.c_void_p(dl.addr), ctypes.byref(st), native.DEVICE_PTRS),
                     "msceer_ondemand_asc")
        hipmem.synchronize()

    log("plain integrator (warm-up)")
    run()
    log("plain integrator")
    t0 = time.perf_counter()
    run()
    plain_s = time.perf_counter() - t0
    plain_stats = st.as_dict()
    plain = dl.download(np.int32, V)
    net = learned.SeparatrixNet(ctx)
    net.load()
    cnt = ctypes.c_int64()
    log("learned mask + masked integration")
    t0 = time.perf_counter()
    native.check(lib.msceer_cnn_mask(ctx.handle, net._h, ctypes.c_void_p(dv.addr), X, Y, Z, float(net.threshold),
                                     int(net.dilation), ctypes.c_void_p(dm.addr), ctypes.byref(cnt),
                                     native.DEVICE_PTRS), "msceer_cnn_mask")
    hipmem.synchronize()
    cnn_s = time.perf_counter() - t0
    run(dm.addr)
    learned_s = time.perf_counter() - t0
    lab = dl.download(np.int32, V)
    net.close()
    probe = None
    if args.ideal_mask_probe:
        probe = ideal_mask_probe(run, dl, dm, st, plain, (Z, Y, X))
    ctx.close()
    out = {"workload": f"{X}x{Y}x{Z} smoothed noise sigma 2.5 seed 5999; accurate labels + conforming .grad",
            "plain": {"seconds": round(plain_s, 3), "value": round(V / plain_s / 1e6, 2), "unit": "Mvert/s",
                      "stage_ms": {k[3:]: round(float(plain_stats[k]), 1) for k in
                                   ("ms_lstar", "ms_regions", "ms_integrate", "ms_cleanup", "ms_grad")}},
            "learned": {"seconds": round(learned_s, 3), "value": round(V / learned_s / 1e6, 2), "unit": "Mvert/s",
                        "cnn_seconds": round(cnn_s, 3), "mask_fraction": cnt.value / V,
                        "threshold": net.threshold, "dilation": net.dilation,
                        "mismatch_vs_plain": float((lab != plain).mean()),
                        "speedup_vs_plain": round(plain_s / learned_s, 3)}}
    if probe is not None:
        out["ideal_mask_probe"] = probe
    return out


def ideal_mask_probe(run, dl, dm, st, plain, shape):
    """The masked path's bound: labelling with the ideal mask (every vertex
    next to a basin boundary of the plain labels or whose combinatorial label
    is wrong, dilated by 1) instead of a learned one, stage by stage.  The
    masks are built on the host with numpy (the labels are downloaded
    anyway)."""
    Z, Y, X = shape
    V = X * Y * Z
    stages = ("ms_lstar", "ms_regions", "ms_integrate", "ms_pass1", "ms_pass3", "ms_cleanup", "ms_grad", "ms_total")
    dm.upload(np.zeros(V, np.uint8))
    run(dm.addr)
    pl = plain.reshape(Z, Y, X)
    core = dl.download(np.int32, V).reshape(Z, Y, X) != pl
    for ax in range(3):
        n = pl.shape[ax]
        d = np.take(pl, range(1, n), axis=ax) != np.take(pl, range(0, n - 1), axis=ax)
        lo = [slice(None)] * 3
        hi = [slice(None)] * 3
        lo[ax] = slice(0, n - 1)
        hi[ax] = slice(1, n)
        core[tuple(lo)] |= d
        core[tuple(hi)] |= d
        del d
    m = core.copy()
    for ax in range(3):  # 3x3x3 dilation, separable
        n = m.shape[ax]
        src = m.copy()
        lo = [slice(None)] * 3
        hi = [slice(None)] * 3
        lo[ax] = slice(0, n - 1)
        hi[ax] = slice(1, n)
        m[tuple(lo)] |= src[tuple(hi)]
        m[tuple(hi)] |= src[tuple(lo)]
        del src
    frac_core, frac = float(core.mean()), float(m.mean())
    del core
    dm.upload(m.astype(np.uint8).ravel())
    del m
    run(dm.addr)  # warm-up
    t0 = time.perf_counter()
    run(dm.addr)
    secs = time.perf_counter() - t0
    labm = dl.download(np.int32, V)
    return {"mask": "vertices next to a plain-label boundary or with a wrong combinatorial label, dilated by 1",
            "core_fraction": frac_core, "mask_fraction": frac, "seconds": round(secs, 3),
            "mismatch_vs_plain": float((labm != plain).mean()),
            "stage_ms": {k[3:]: round(float(st.as_dict()[k]), 2) for k in stages},
            "integrated_vertices": int(st.integrated_vertices)}


def calibrate_copy():
    import torch

    n = 1 << 28
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(3):
        torch.mul(a, 2.0, out=b)
    torch.cuda.synchronize()
    print(json.dumps({"calibration": "torch.mul float32", "bytes_read": 4 * n, "bytes_written": 4 * n}))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--size", type=int, default=512)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-slab", type=int, default=96)
    ap.add_argument("--no-labels", action="store_true", help="skip the accurate-labelling measurement")
    ap.add_argument("--labels-only", action="store_true", help="only the accurate-labelling measurement")
    ap.add_argument("--label-steps", type=int, default=3)
    ap.add_argument("--label-warmup", type=int, default=1)
    ap.add_argument("--label-cpu-slab", type=int, default=48)
    ap.add_argument("--ideal-mask-probe", action="store_true")
    ap.add_argument("--no-learned", action="store_true")
    ap.add_argument("--no-lines", action="store_true", help="skip the accurate 1-manifold measurement")
    ap.add_argument("--no-dsc", action="store_true", help="skip the descending 3-manifold measurement")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--batch-only", action="store_true")
    ap.add_argument("--batch-volumes", type=int, default=512)
    ap.add_argument("--batch-calls", type=int, default=5, help="timed calls of the batched .grad (one group)")
    ap.add_argument("--batch-graph-only", action="store_true",
                    help="with --batch-only: only the batched .grad as one group (for profiling)")
    ap.add_argument("--batch-streams", type=int, default=16, help="worker streams of the batched workload")
    ap.add_argument("--hw-queues", type=int, default=16,
                    help="GPU_MAX_HW_QUEUES for this process (HIP maps the batch workers' streams onto this many "
                         "hardware queues; 0 keeps the environment's value)")
    ap.add_argument("--learned-large", action="store_true",
                    help="only the learned-mask comparison on the 1024x768x512 volume")
    ap.add_argument("--no-learned-large", action="store_true",
                    help="skip the learned-mask comparison on the 1024x768x512 volume")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run a plain 1 GiB read + 1 GiB write streaming kernel (torch), the "
                         "reference point for rocprofv3 FETCH_SIZE / WRITE_SIZE")
    args = ap.parse_args()
    if args.hw_queues > 0:
        # read by the HIP runtime at its first call, which has not happened yet
        os.environ["GPU_MAX_HW_QUEUES"] = str(min(args.hw_queues, 32))
    if args.calibrate_copy:
        calibrate_copy()
        return
    if args.gpus != 1:
        print(f"bench.py: this implementation runs on one GPU only (asked for --gpus {args.gpus}); "
              "sharding across GPUs (z-slabs with halo exchange) is a design, not built", file=sys.stderr)
        sys.exit(2)

    if args.labels_only:
        print(json.dumps({"labels": bench_labels(args)}), flush=True)
        return
    if args.learned_large:
        print(json.dumps({"learned_large": bench_learned_large(args)}), flush=True)
        return
    if args.batch_only:
        print(json.dumps({"batch": bench_batch(args)}), flush=True)
        return
    from msceer_amd import hipmem, native, synth

    n = args.size
    log(f"generating {n}^3 smoothed noise (sigma 2, seed 202)")
    f = synth.lstar_large(202, n)
    Z, Y, X = f.shape
    cells = (2 * X - 1) * (2 * Y - 1) * (2 * Z - 1)
    verts = X * Y * Z

    ctx = native.Context(0)
    dv = hipmem.DeviceBuffer(f.nbytes)
    dv.upload(f)
    dg = hipmem.DeviceBuffer(cells)
    lib = native.lib()

    def step(timing):
        flags = native.DEVICE_PTRS | (native.TIMING if timing else 0)
        native.check(lib.msceer_lstar_grad(ctx.handle, ctypes.c_void_p(dv.addr), X, Y, Z,
                                           ctypes.c_void_p(dg.addr), flags), "msceer_lstar_grad")

    for _ in range(args.warmup):
        step(False)
    hipmem.synchronize()
    lstar_ms, ldir_ms, rounds = [], [], []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step(True)
        t = ctx.last_timing()
        lstar_ms.append(t.lstar_ms)
        ldir_ms.append(t.ldir_ms)
        rounds.append(t.ldir_rounds)
    hipmem.synchronize()
    elapsed = time.perf_counter() - t0
    crit = list(ctx.last_timing().critical)
    ms_per_step = elapsed * 1e3 / args.steps

    algo_bytes = 4.0 * verts + 1.0 * cells  # 4 B read per vertex + 1 B written per cell
    k_ms = float(np.mean(lstar_ms))
    achieved = algo_bytes / (k_ms * 1e-3) / 1e9
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4),
                "traffic": read_traffic(f"lstar_grad_{n}_traffic.txt"),
                "kernel": "lstar_kernel", "kernel_ms": round(k_ms, 4),
                "algorithmic_bytes": algo_bytes}

    # extractmsc's reading of the same .grad (nodes, base basins, base mountains)
    log("Morse-Smale nodes, basins and mountains of the .grad")
    db, dm = hipmem.DeviceBuffer(4 * verts), hipmem.DeviceBuffer(4 * verts)
    counts = np.zeros(4, np.int64)

    def msc_call():
        native.check(lib.msceer_msc_base(ctx.handle, ctypes.c_void_p(dg.addr), X, Y, Z, ctypes.c_void_p(db.addr),
                                         ctypes.c_void_p(dm.addr), counts.ctypes.data_as(ctypes.c_void_p),
                                         native.DEVICE_PTRS), "msceer_msc_base")

    msc_call()
    hipmem.synchronize()
    t0 = time.perf_counter()
    for _ in range(3):
        msc_call()
    hipmem.synchronize()
    msc_ms = (time.perf_counter() - t0) * 1e3 / 3
    db.free()
    dm.free()
    msc_leg = {"ms": round(msc_ms, 3), "value": round(verts / (msc_ms * 1e-3) / 1e6, 2), "unit": "Mvert/s",
               "nodes_by_dim": counts.tolist(),
               "what": "msceer_msc_base on the .grad above: node numbering, V-path basins, descending-manifold "
                       "mountains (extractmsc before simplification), device buffers"}

    base = None
    if not args.no_cpu:
        log("timing the CPU baseline")
        base = cpu_baseline(f, min(args.cpu_slab, Z))

    out = {
        "metric": METRIC,
        "value": round(verts / (ms_per_step * 1e-3) / 1e6, 2),
        "unit": "Mvert/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"steepest_lstar .grad, {X}x{Y}x{Z} float32, white noise smoothed "
                               f"sigma=2 vox, seed 202; values resident in HBM",
                   "parallelism": "single GPU"},
        "roofline": roofline,
        "cpu_baseline": base,
        "msc_base": msc_leg,
        "detail": {"lstar_kernel_ms_mean": round(k_ms, 4),
                   "ldir_ms_mean": round(float(np.mean(ldir_ms)), 4),
                   "ldir_rounds": int(np.max(rounds)), "critical_cells_by_dim": crit},
    }
    ctx.close()
    if not args.no_labels:
        out["labels"] = bench_labels(args)
    if not args.no_batch:
        out["batch"] = bench_batch(args)
    if not args.no_learned_large:
        out["learned_large"] = bench_learned_large(args)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
