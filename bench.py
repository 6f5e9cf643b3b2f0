#!/usr/bin/env python3
"""Benchmark: fused FP64 WENO5-JS reconstruction, 200^3 x 8, one MI355X.

One step = the reference's WENO5 reconstruction in `kernel` scope
(s_run_weno5_reconstruction with the specialised x/y/z kernels,
inputs/weno_standalone.f90:1713-1724): three sweeps (x, y, z), each writing
the core range of vL/vR (2*206*212*212*8 = 148,135,424 face values per sweep)
straight from v_vf with direction-aware loads.  Inputs are resident in HBM
before the timed region.  value = face values reconstructed per second.

Prints one JSON line.  Also reports the dominant kernel's roofline figure
(algorithmic bytes per launch / HIP-event launch time on the launch stream),
the metric's second half (the same step with the state stored in FP32 and in
BF16: speedup over FP64 and element-wise error against the FP64 path, on the
field with 1e3 jumps and exactly constant boxes) and a CPU baseline (the
reference's own OpenMP build on every host core, else the oracle's C port).
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import re
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0   # MI355X HBM3E
# measured HBM traffic of the dominant kernel (TCC FETCH_SIZE / WRITE_SIZE,
# scaled by the same counters read on a plain copy): written by
# `bench.py --record-traffic FETCH_CSV WRITE_CSV`, used only while the kernel
# sources hash the same as when it was measured
TRAFFIC_RECORD = os.path.join(ROOT, "profiles", "sweep_weno5_fp64_200cube_traffic.json")
KERNEL_SOURCES = ["weno_sweep5.hip", "weno_cells.h", "weno_kernels.hip", "weno_device.h", "weno_internal.h",
                  "weno_capi.hip", "weno_tune.hip"]
SWEEP_KERNEL = "sweep5_kernel<"


def kernel_source_hash() -> str:
    import hashlib
    h = hashlib.sha256()
    for name in KERNEL_SOURCES:
        with open(os.path.join(ROOT, "wenostandalone_amd", "csrc", name), "rb") as fh:
            h.update(fh.read())
    return h.hexdigest()


def record_traffic(fetch_csv: str, write_csv: str, m: int = 199, v: int = 8) -> dict:
    """Per-launch HBM bytes of the x/y/z WENO5 sweeps from two rocprofv3
    --pmc runs of `bench.py --steps 1 --warmup 0 --no-cpu --copy-yardstick`.
    The stream copy of one face array (known bytes read and written) gives
    the KiB-per-count scale of each counter."""
    import csv

    def rows(path):
        with open(path) as fh:
            return [(r["Kernel_Name"], float(r["Counter_Value"])) for r in csv.DictReader(fh)]

    nx = m + 1 + 2 * 6
    copy_kib = 8 * nx ** 3 * v / 1024
    fr, wr = rows(fetch_csv), rows(write_csv)
    cf = [x for k, x in fr if "stream_copy" in k]
    cw = [x for k, x in wr if "stream_copy" in k]
    # the timed FP64 x/y/z sweeps come first (the precision leg, if run,
    # launches the same kernels afterwards)
    sf = [x for k, x in fr if SWEEP_KERNEL + "double" in k][:3]
    sw = [x for k, x in wr if SWEEP_KERNEL + "double" in k][:3]
    if len(cf) != 1 or len(cw) != 1 or len(sf) != 3 or len(sw) != 3:
        raise SystemExit("expected one stream copy and three FP64 sweeps in each counter file")
    kf, kw = copy_kib / cf[0], copy_kib / cw[0]
    per = [(a * kf + b * kw) * 1024 for a, b in zip(sf, sw)]
    rec = {"kernel_source_sha256": kernel_source_hash(), "traffic_bytes_per_launch": float(np.mean(per)),
           "per_sweep_bytes": dict(zip("xyz", per)), "fetch_KiB": dict(zip("xyz", sf)),
           "write_KiB": dict(zip("xyz", sw)), "copy_fetch_KiB": cf[0], "copy_write_KiB": cw[0],
           "fetch_scale": kf, "write_scale": kw}
    with open(TRAFFIC_RECORD, "w") as fh:
        json.dump(rec, fh, indent=1)
    return rec


def recorded_traffic():
    try:
        with open(TRAFFIC_RECORD) as fh:
            rec = json.load(fh)
    except (OSError, ValueError):
        return None
    if rec.get("kernel_source_sha256") != kernel_source_hash():
        return None
    return float(rec["traffic_bytes_per_launch"])


def hip():
    h = C.CDLL("libamdhip64.so")
    h.hipEventCreate.argtypes = [C.POINTER(C.c_void_p)]
    h.hipEventRecord.argtypes = [C.c_void_p, C.c_void_p]
    h.hipEventSynchronize.argtypes = [C.c_void_p]
    h.hipEventElapsedTime.argtypes = [C.POINTER(C.c_float), C.c_void_p, C.c_void_p]
    h.hipEventDestroy.argtypes = [C.c_void_p]
    h.hipDeviceSynchronize.argtypes = []
    return h


def host_cores() -> int:
    """Cores this process may run on (its CPU affinity set)."""
    try:
        return max(1, len(os.sched_getaffinity(0)))
    except (AttributeError, OSError):
        return max(1, os.cpu_count() or 1)


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as fh:
            return [l.split(":", 1)[1].strip() for l in fh if l.startswith("model name")][0]
    except (OSError, IndexError):
        return "unknown CPU"


def cpu_baseline(m: int, v: int, buff: int, budget_s: float = 60.0):
    """The reference's own CPU OpenMP build (oracle/_ref/weno_standalone,
    compiled from the reference source) in weno5/kernel scope on every host
    core this process may use, with the reference's published protocol of 5
    warm-up and 20 timed iterations (results/amd_mp_roof_ex/log.txt:31-32);
    falls back to the oracle's C port."""
    cores = host_cores()
    faces = 3 * 2 * (m + 1 + 2 * buff - 6) * (m + 1 + 2 * buff) ** 2 * v
    ref = os.path.join(ROOT, "oracle", "_ref", "weno_standalone")
    if os.path.exists(ref):
        with tempfile.TemporaryDirectory() as td:
            nml = os.path.join(td, "in.nml")
            with open(nml, "w") as fh:
                fh.write(f"&grid_params\n m_cells={m}, n_cells={m}, p_cells={m}, v_size={v},\n"
                         f" buff_size={buff}, weno_eps=1e-16, wenojs=T, teno_CT=1e-5\n/\n")
            env = dict(os.environ, OMP_TARGET_OFFLOAD="DISABLED", OMP_NUM_THREADS=str(cores),
                       OMP_PROC_BIND="close", WENO_KERNEL_MODE="weno5", WENO5_BENCH_SCOPE="kernel",
                       WENO5_SPECIALIZED_COMBINED="1", WENO_WARMUP_ITERS="5", WENO_BENCH_ITERS="20")
            out = subprocess.run([ref, nml], env=env, capture_output=True, text=True,
                                 timeout=budget_s * 4).stdout
        t = float(re.search(r"Kernel time per iter\s*:\s*([0-9.Ee+-]+)", out).group(1))
        return {"value": faces / t, "unit": "face values/s", "cores": cores, "kind": "reference",
                "sample": f"reference OpenMP build (amdflang -fopenmp -O3, host OpenMP), weno5 kernel scope "
                          f"(x+y+z), {m + 1}^3x{v}, OMP_NUM_THREADS={cores} OMP_PROC_BIND=close, "
                          f"5 warm-up + 20 timed iterations, {t:.4f} s/iter; host {cpu_model()}, "
                          f"{cores} cores in the affinity set"}
    import oracle
    from wenostandalone_amd import fields
    mm = 99   # bounded sample
    g = oracle.grid(mm, v=v, buff=buff)
    f = fields.seeded_field(mm, mm, mm, v, buff, seed=0)
    t0 = time.perf_counter()
    n = 0
    for d in range(3):
        ns = oracle.dir_extents(g, d)[0]
        cs = oracle.fill_coefficients(ns - 6, 0.03)
        rs = oracle.reshape(g, d, f)
        vL, vR = oracle.order1_copy(g, d, f)
        t1 = time.perf_counter()
        oracle.weno(g, 5, d, rs, cs, vL, vR)
        n += time.perf_counter() - t1
    sfaces = 3 * 2 * (mm + 1 + 2 * buff - 6) * (mm + 1 + 2 * buff) ** 2 * v
    return {"value": sfaces / n, "unit": "face values/s", "cores": cores, "kind": "port",
            "sample": f"oracle C port, WENO5 x+y+z at {mm + 1}^3x{v} ({n:.3f} s); host {cpu_model()}"}


def bf16_bits(f32: np.ndarray) -> np.ndarray:
    """FP32 -> BF16 bit patterns, round to nearest even."""
    u = f32.view(np.uint32)
    return np.asfortranarray(((u + 0x7FFF + ((u >> 16) & 1)) >> 16).astype(np.uint16))


def precision_leg(h, stream, m: int, v: int, buff: int, steps: int, warmup: int) -> dict:
    """The metric's second half: WENO5-JS x+y+z at (m+1)^3 x v with the state
    stored in FP64, FP32 and BF16 (FP32 faces for both), on the recipe-2 field
    (1e3 jumps and exactly constant boxes), core-range outputs as in the
    headline step.  Each form is timed over `steps` steps with HIP events on
    the library stream after `warmup` steps; the FP32 / BF16 faces are
    compared element-wise with the FP64 faces (the FP64 path is held to 1e-13
    of the oracle by the test suite): relative error |a-b|/|b| over every core
    face of the six outputs (values of this field are >= 0.1 in magnitude)."""
    import wenostandalone_amd as wm
    from wenostandalone_amd import fields
    from wenostandalone_amd.device import DeviceArray

    g = wm.make_grid(m, v=v, buff=buff)
    nx = m + 1 + 2 * buff
    f = fields.seeded_field(m, m, m, v, buff, seed=3, recipe=2)
    f32 = f.astype(np.float32)
    srcs = {wm.F64: DeviceArray.from_host(f), wm.F32: DeviceArray.from_host(f32),
            wm.BF16: DeviceArray.from_host(bf16_bits(f32))}
    del f, f32
    coefs = []
    for d in range(3):
        devs = [DeviceArray.from_host(c) for c in fields.harness_coefficients(nx - 6, fields.DIR_OFFSETS[d])]
        coefs.append((devs, wm.Coeffs(*[x.ptr for x in devs])))
    faces = {}
    for dt in (wm.F64, wm.F32, wm.BF16):
        odt = np.float64 if dt == wm.F64 else np.float32
        faces[dt] = [(DeviceArray((nx, nx, nx, v), odt), DeviceArray((nx, nx, nx, v), odt)) for _ in range(3)]
        for d in range(3):
            wm.check(wm.lib.weno_order1_copy(C.byref(g), d, dt, srcs[dt].vp, faces[dt][d][0].vp,
                                             faces[dt][d][1].vp))

    def step(dt):
        for d in range(3):
            wm.check(wm.lib.weno_reconstruct(C.byref(g), 5, d, wm.SRC_FIELD, dt, srcs[dt].vp, C.byref(coefs[d][1]),
                                             faces[dt][d][0].vp, faces[dt][d][1].vp, 0))

    wm.check(wm.lib.weno_set_async(1))
    ms = {}
    a, b = C.c_void_p(), C.c_void_p()
    h.hipEventCreate(C.byref(a))
    h.hipEventCreate(C.byref(b))
    for dt in (wm.F64, wm.F32, wm.BF16):
        for _ in range(warmup):
            step(dt)
        wm.check(wm.lib.weno_sync())
        h.hipEventRecord(a, stream)
        for _ in range(steps):
            step(dt)
        h.hipEventRecord(b, stream)
        wm.check(wm.lib.weno_sync())
        t = C.c_float()
        h.hipEventElapsedTime(C.byref(t), a, b)
        ms[dt] = t.value / steps
    wm.check(wm.lib.weno_set_async(0))
    core = slice(3, nx - 3)
    err = {}
    for dt, name in ((wm.F32, "fp32"), (wm.BF16, "bf16")):
        mx, ss, n = 0.0, 0.0, 0
        for d in range(3):
            for k in range(2):
                ref = faces[wm.F64][d][k].download()[core]
                got = faces[dt][d][k].download()[core].astype(np.float64)
                rel = np.abs(got - ref) / np.maximum(np.abs(ref), 1e-300)
                del got, ref
                mx = max(mx, float(rel.max()))
                ss += float(np.sum(rel * rel))
                n += rel.size
                del rel
        err[name] = (mx, float(np.sqrt(ss / n)))
    for dt in faces:
        for L, R in faces[dt]:
            L.free()
            R.free()
    for s in srcs.values():
        s.free()
    return {"fp32_speedup": ms[wm.F64] / ms[wm.F32], "fp32_max_rel_err": err["fp32"][0],
            "fp32_rms_rel_err": err["fp32"][1], "bf16_speedup": ms[wm.F64] / ms[wm.BF16],
            "bf16_max_rel_err": err["bf16"][0], "bf16_rms_rel_err": err["bf16"][1],
            "ms_per_step": {"fp64": ms[wm.F64], "fp32": ms[wm.F32], "bf16": ms[wm.BF16]},
            "field": "recipe 2 (sine modes + step + noise, 1e3 jump, exactly constant boxes), seed 3",
            "error_reference": "FP64 path on the FP64 field, core faces of vL/vR x/y/z"}


def _events(h, n):
    out = []
    for _ in range(n):
        e = C.c_void_p()
        h.hipEventCreate(C.byref(e))
        out.append(e)
    return out


def _time(h, stream, fn, reps=10, warm=2, warm_s=0.3):
    """Mean milliseconds of fn() over reps launches, HIP events on stream,
    after at least warm_s seconds of warm-up launches (the GPU clocks up
    after host-side pauses such as the error checks between forms)."""
    import wenostandalone_amd as wm
    for _ in range(warm):
        fn()
    wm.check(wm.lib.weno_sync())
    t0 = time.perf_counter()
    while time.perf_counter() - t0 < warm_s:
        fn()
        wm.check(wm.lib.weno_sync())
    a, b = _events(h, 2)
    h.hipEventRecord(a, stream)
    for _ in range(reps):
        fn()
    h.hipEventRecord(b, stream)
    wm.check(wm.lib.weno_sync())
    ms = C.c_float()
    h.hipEventElapsedTime(C.byref(ms), a, b)
    return ms.value / reps


def suite(path, parts=("precision", "batched", "adjoint")):
    """Secondary measurements (written to `path` as JSON):
      precision  - WENO5/7 at 200^3x8 on the recipe-2 field (1e3 jumps and
                   exactly constant boxes): FP64 vs FP32 state vs BF16 state,
                   WENO5 also through the general kernel (FP32 stencil
                   arithmetic), time per x+y+z and per-output max / RMS
                   relative error
                   against the FP64 GPU path (itself held to 1e-13 of the
                   oracle by the test suite);
      batched    - 64 blocks, sides U{16..128} per axis, v in {5,8,10},
                   WENO5/WENO7 mixed, x/y/z as one hipGraph after autotuning;
      adjoint    - WENO5/7-JS at 128^3x8: forward x+y+z vs forward + VJP.
    `parts` selects which of the three run."""
    import wenostandalone_amd as wm
    from wenostandalone_amd import fields
    from wenostandalone_amd.device import DeviceArray

    wm.check(wm.lib.weno_init(0))
    wm.check(wm.lib.weno_set_async(1))
    stream = wm.lib.weno_get_stream()
    h = hip()
    report = {}
    b = 6   # buff_size of every section

    if "precision" in parts:
        # ---- precision forms ------------------------------------------------
        m, v, b = 199, 8, 6
        g = wm.make_grid(m, v=v, buff=b)
        nx = m + 1 + 2 * b
        f = fields.seeded_field(m, m, m, v, b, seed=3, recipe=2)
        f32 = f.astype(np.float32)
        u = f32.view(np.uint32)
        bf = np.asfortranarray(((u + 0x7FFF + ((u >> 16) & 1)) >> 16).astype(np.uint16))
        del u
        srcs = {wm.F64: DeviceArray.from_host(f), wm.F32: DeviceArray.from_host(f32),
                wm.BF16: DeviceArray.from_host(bf)}
        del f, f32, bf
        coefs = []
        for d in range(3):
            devs = [DeviceArray.from_host(c) for c in fields.harness_coefficients(nx - 6, fields.DIR_OFFSETS[d])]
            coefs.append((devs, wm.Coeffs(*[x.ptr for x in devs])))
        faces = {dt: [(DeviceArray((nx, nx, nx, v), np.float64 if dt == wm.F64 else np.float32),
                       DeviceArray((nx, nx, nx, v), np.float64 if dt == wm.F64 else np.float32))
                      for _ in range(3)] for dt in (wm.F64, wm.F32)}
        prec = {}
        for order in (5, 7):
            ref = None
            row = {}
            forms = [(wm.F64, "fp64", -1), (wm.F32, "fp32", -1), (wm.BF16, "bf16", -1)]
            if order == 5:
                # the general kernel's form: stencil arithmetic in FP32
                forms += [(wm.F32, "fp32_general_kernel", 8 | 2), (wm.BF16, "bf16_general_kernel", 8 | 2)]
            for dt, name, variant in forms:
                fc = faces[wm.F64 if dt == wm.F64 else wm.F32]

                def sweep():
                    for d in range(3):
                        wm.check(wm.lib.weno_reconstruct(C.byref(g), order, d, wm.SRC_FIELD, dt, srcs[dt].vp,
                                                         C.byref(coefs[d][1]), fc[d][0].vp, fc[d][1].vp,
                                                         wm.FUSE_COPY))
                for d in range(3):
                    wm.check(wm.lib.weno_set_tile(order, dt, d, 0, 0, 0, variant))
                try:
                    ms = _time(h, stream, sweep)
                finally:
                    for d in range(3):
                        wm.check(wm.lib.weno_set_tile(order, dt, d, 0, 0, 0, -1))
                ent = {"ms_xyz": ms}
                if dt == wm.F64:
                    ref = [(L.download(), R.download()) for L, R in fc]
                else:
                    errs = {}
                    for d, ax in enumerate("xyz"):
                        for side, k in (("L", 0), ("R", 1)):
                            got = fc[d][k].download().astype(np.float64)
                            r = ref[d][k]
                            rel = np.abs(got - r) / np.maximum(np.abs(r), 1e-3 * np.abs(r).max())
                            errs[f"{ax}{side}"] = {"max": float(rel.max()), "rms": float(np.sqrt(np.mean(rel ** 2)))}
                            del got, rel
                    ent["rel_err_vs_fp64"] = errs
                row[name] = ent
            for name in list(row):
                if name != "fp64":
                    row[f"speedup_{name}"] = row["fp64"]["ms_xyz"] / row[name]["ms_xyz"]
            prec[f"weno{order}"] = row
        report["precision"] = prec
        for dt in faces:
            for L, R in faces[dt]:
                L.free()
                R.free()
        for s in srcs.values():
            s.free()

    if "batched" in parts:
        # ---- batched blocks --------------------------------------------------
        rng = np.random.default_rng(2024)
        specs = [(int(rng.integers(16, 129)) - 1, int(rng.integers(16, 129)) - 1, int(rng.integers(16, 129)) - 1,
                  int(rng.choice([5, 8, 10])), int(rng.choice([5, 7]))) for _ in range(64)]
        blocks = (wm.Block * 64)()
        keep = []
        nfaces = 0
        for i, (bm, bn, bp, bv, order) in enumerate(specs):
            fld = fields.seeded_field(bm, bn, bp, bv, b, seed=100 + i)
            vf = DeviceArray.from_host(fld)
            keep.append(vf)
            blk = blocks[i]
            blk.grid = wm.make_grid(bm, bn, bp, v=bv, buff=b)
            blk.order = order
            blk.v_vf = vf.ptr
            for d in range(3):
                ns, nk, nl = fields.dir_extents(bm, bn, bp, b, d)
                devs = [DeviceArray.from_host(c) for c in fields.harness_coefficients(ns - 6, fields.DIR_OFFSETS[d])]
                keep += devs
                blk.coeffs[d] = wm.Coeffs(*[x.ptr for x in devs])
                L, R = DeviceArray((ns, nk, nl, bv)), DeviceArray((ns, nk, nl, bv))
                keep += [L, R]
                blk.vL[d], blk.vR[d] = L.ptr, R.ptr
                nfaces += 2 * ns * nk * nl * bv
        best = (C.c_int * 3)()
        wm.check(wm.lib.weno_batched_autotune(blocks, 64, wm.FUSE_COPY, best))
        gh = C.c_void_p()
        wm.check(wm.lib.weno_batched_graph_create(blocks, 64, wm.FUSE_COPY, C.byref(gh)))
        ms = _time(h, stream, lambda: wm.check(wm.lib.weno_graph_launch(gh)))
        wm.check(wm.lib.weno_graph_destroy(gh))
        report["batched"] = {"blocks": 64, "sides": "U{16..128} per axis", "v_size": "{5,8,10}", "orders": "{5,7}",
                             "face_values_per_xyz": nfaces, "ms_xyz_graph": ms,
                             "face_values_per_s": nfaces / (ms * 1e-3),
                             "autotuned": {"threads": best[0], "pencils": best[1]}}
        for x in keep:
            x.free()
        del keep

    if "adjoint" in parts:
        # ---- adjoint -------------------------------------------------------------
        m, v = 127, 8
        g = wm.make_grid(m, v=v, buff=b)
        nx = m + 1 + 2 * b
        vf = DeviceArray.from_host(fields.seeded_field(m, m, m, v, b, seed=4))
        gv = DeviceArray(vf.shape)
        coefs = []
        for d in range(3):
            devs = [DeviceArray.from_host(c) for c in fields.harness_coefficients(nx - 6, fields.DIR_OFFSETS[d])]
            coefs.append((devs, wm.Coeffs(*[x.ptr for x in devs])))
        L3 = [DeviceArray((nx, nx, nx, v)) for _ in range(3)]
        R3 = [DeviceArray((nx, nx, nx, v)) for _ in range(3)]
        cot = np.asfortranarray(np.random.default_rng(1).standard_normal((nx, nx, nx, v)))
        gL, gR = DeviceArray.from_host(cot), DeviceArray.from_host(cot[::-1].copy(order="F"))
        del cot
        adj = {}
        for order in (5, 7):
            def fwd():
                for d in range(3):
                    wm.check(wm.lib.weno_reconstruct(C.byref(g), order, d, wm.SRC_FIELD, wm.F64, vf.vp,
                                                     C.byref(coefs[d][1]), L3[d].vp, R3[d].vp, 0))

            def bwd():
                for d in range(3):
                    wm.check(wm.lib.weno_reconstruct_vjp(C.byref(g), order, d, wm.SRC_FIELD, vf.vp,
                                                         C.byref(coefs[d][1]), gL.vp, gR.vp, gv.vp))
            t_f = _time(h, stream, fwd)
            t_b = _time(h, stream, bwd)
            adj[f"weno{order}"] = {"fwd_ms_xyz": t_f, "vjp_ms_xyz": t_b, "fwd_plus_bwd_over_fwd": (t_f + t_b) / t_f}
        report["adjoint"] = adj
    wm.check(wm.lib.weno_set_async(0))
    with open(path, "w") as fh:
        json.dump(report, fh, indent=1)
    print(json.dumps(report), file=sys.stderr, flush=True)


def families(path, m=199, v=8, buff=6, reps=10):
    """Every kernel family at m+1 cubed x v FP64 (default 200^3 x 8), x, y and
    z sweeps: HIP-event time per launch on the library stream and its
    algorithmic HBM bytes (each distinct input element read once, each output
    written once), as a fraction of the 8 TB/s HBM peak.  Per direction with
    ns = nk = nl = m+1+2*buff (212), nc = ns-6 (206):
      reshape .................. read (ns+6) nk nl v, write the same
      order-1 copy ............. read ns nk nl v, write 2 ns nk nl v
      WENO3, fused copy ........ read (ns+2) nk nl v, write 2 ns nk nl v
      WENO5/7, fused copy (+MP). read ns nk nl v, write 2 ns nk nl v
      WENO5/7, core only ....... read (nc+2R) nk nl v, write 2 nc nk nl v
      MP standalone ............ read (nc+4) nk nl v of v_rs and vL, vR on
                                 the core, write vL, vR on the core
    WENO forms run from v_vf (direction-aware loads); MP reads the reshaped
    array as the reference's s_preserve_monotonicity does (:1544)."""
    import wenostandalone_amd as wm
    from wenostandalone_amd import fields
    from wenostandalone_amd.device import DeviceArray

    wm.check(wm.lib.weno_init(0))
    h = hip()
    nx = m + 1 + 2 * buff
    nc = nx - 6
    plane = nx * nx * v * 8           # bytes of one j-slab of a face array
    f = fields.seeded_field(m, m, m, v, buff, seed=0)
    vf = DeviceArray.from_host(f)
    del f
    rs = DeviceArray((nx + 6, nx, nx, v))
    L, R = DeviceArray((nx, nx, nx, v)), DeviceArray((nx, nx, nx, v))
    coefs = []
    for d in range(3):
        devs = [DeviceArray.from_host(c) for c in fields.harness_coefficients(nc, fields.DIR_OFFSETS[d])]
        coefs.append((devs, wm.Coeffs(*[x.ptr for x in devs])))
    branches = {"js": {}, "teno": {"wenojs": False, "teno": True},
                "mapped": {"wenojs": False, "mapped_weno": True}, "wenoz": {"wenojs": False, "wenoz": True}}
    grids = {k: wm.make_grid(m, v=v, buff=buff, **kw) for k, kw in branches.items()}
    g = grids["js"]

    forms = [("reshape", lambda d: wm.lib.weno_reshape(C.byref(g), d, wm.F64, vf.vp, rs.vp),
              2 * (nx + 6) * plane),
             ("order1_copy", lambda d: wm.lib.weno_order1_copy(C.byref(g), d, wm.F64, vf.vp, L.vp, R.vp),
              3 * nx * plane)]

    def recon(order, br, flags):
        return lambda d: wm.lib.weno_reconstruct(C.byref(grids[br]), order, d, wm.SRC_FIELD, wm.F64, vf.vp,
                                                 C.byref(coefs[d][1]), L.vp, R.vp, flags)
    for order in (3, 5, 7):
        for br in branches:
            if order == 3 and br == "teno":
                continue      # the reference has no TENO3 branch
            rd = nx + 2 if order == 3 else nx
            forms.append((f"weno{order}_{br}_fused_copy", recon(order, br, wm.FUSE_COPY), (rd + 2 * nx) * plane))
    for order in (5, 7):
        forms.append((f"weno{order}_js_core_only", recon(order, "js", 0), (nc + order - 1 + 2 * nc) * plane))
        forms.append((f"weno{order}_js_fused_copy_mp", recon(order, "js", wm.FUSE_COPY | wm.FUSE_MP),
                      3 * nx * plane))
    forms.append(("mp_standalone", lambda d: wm.lib.weno_mp_limit(C.byref(g), d, wm.SRC_RESHAPED, wm.F64, rs.vp,
                                                                 L.vp, R.vp), (nc + 4 + 4 * nc) * plane))

    wm.check(wm.lib.weno_set_async(1))
    stream = wm.lib.weno_get_stream()
    out = {"workload": f"{m + 1}^3 x {v} FP64, harness coefficients, seeded field recipe 0",
           "peak_GBs": HBM_PEAK_GBS, "forms": {}}
    for name, fn, nbytes in forms:
        row = {"bytes_per_launch": nbytes}
        for d, ax in enumerate("xyz"):
            if name == "mp_standalone":
                # the reshaped array of this direction and faces from WENO5 on it
                wm.check(wm.lib.weno_reshape(C.byref(g), d, wm.F64, vf.vp, rs.vp))
                wm.check(recon(5, "js", wm.FUSE_COPY)(d))
            ms = _time(h, stream, lambda: wm.check(fn(d)), reps=reps)
            row[f"{ax}_us"] = ms * 1e3
            row[f"{ax}_GBs"] = nbytes / (ms * 1e-3) / 1e9
        mean_us = float(np.mean([row[f"{ax}_us"] for ax in "xyz"]))
        row["mean_us"] = mean_us
        row["GBs"] = nbytes / (mean_us * 1e-6) / 1e9
        row["frac_of_peak"] = row["GBs"] / HBM_PEAK_GBS
        out["forms"][name] = row
        print(f"families: {name:28s} {mean_us:9.1f} us  {row['GBs']:7.0f} GB/s  {row['frac_of_peak']:.3f}",
              file=sys.stderr, flush=True)
    wm.check(wm.lib.weno_set_async(0))
    for x in (vf, rs, L, R):
        x.free()
    with open(path, "w") as fh:
        json.dump(out, fh, indent=1)
    return out


def consumer(path, m=199, v=8, buff=6, reps=20):
    """The consumer step (flux divergence of one sweep's faces) at m+1 cubed
    x v FP64, WENO5-JS, Burgers flux, x/y/z: the fused pass
    (weno_reconstruct_flux: state read once, rhs written once) against
    reconstruction followed by the separate flux pass (weno_reconstruct +
    weno_flux_divergence).  Algorithmic bytes per sweep with ns = m+1+2*buff
    points along it and ns^2 pencils of v variables:
      fused ......... read (ns-6+4) inputs, write (ns-8) rhs per pencil
      two passes .... reconstruction: read (ns-2), write 2 (ns-6) faces;
                      flux: read 2 (ns-6) faces, write (ns-8) rhs"""
    import wenostandalone_amd as wm
    from wenostandalone_amd import fields
    from wenostandalone_amd.device import DeviceArray

    wm.check(wm.lib.weno_init(0))
    h = hip()
    ns = m + 1 + 2 * buff
    g = wm.make_grid(m, v=v, buff=buff)
    f = fields.seeded_field(m, m, m, v, buff, seed=0)
    vf = DeviceArray.from_host(f)
    del f
    coefs = []
    for d in range(3):
        devs = [DeviceArray.from_host(c) for c in fields.harness_coefficients(ns - 6, fields.DIR_OFFSETS[d])]
        coefs.append((devs, wm.Coeffs(*[x.ptr for x in devs])))
    L, R, rhs = DeviceArray((ns, ns, ns, v)), DeviceArray((ns, ns, ns, v)), DeviceArray((ns, ns, ns, v))
    fx = wm.Flux(0.0, 1.0, float(m + 1))
    pen = ns * ns * v * 8
    fused_bytes = ((ns - 2) + (ns - 8)) * pen
    two_bytes = ((ns - 2) + 2 * (ns - 6)) * pen + (2 * (ns - 6) + (ns - 8)) * pen
    wm.check(wm.lib.weno_set_async(1))
    stream = wm.lib.weno_get_stream()
    out = {"workload": f"{m + 1}^3 x {v} FP64, WENO5-JS, Burgers (a=0, b=1) Rusanov flux divergence",
           "fused_bytes_per_sweep": fused_bytes, "two_pass_bytes_per_sweep": two_bytes, "sweeps": {}}
    for d, ax in enumerate("xyz"):
        def fused():
            wm.check(wm.lib.weno_reconstruct_flux(C.byref(g), 5, d, wm.SRC_FIELD, vf.vp, C.byref(coefs[d][1]),
                                                  C.byref(fx), rhs.vp, 0))

        def two():
            wm.check(wm.lib.weno_reconstruct(C.byref(g), 5, d, wm.SRC_FIELD, wm.F64, vf.vp, C.byref(coefs[d][1]),
                                             L.vp, R.vp, 0))
            wm.check(wm.lib.weno_flux_divergence(C.byref(g), d, L.vp, R.vp, C.byref(fx), rhs.vp, 0))
        tf = _time(h, stream, fused, reps=reps)
        tt = _time(h, stream, two, reps=reps)
        out["sweeps"][ax] = {"fused_us": tf * 1e3, "two_pass_us": tt * 1e3, "speedup": tt / tf,
                             "fused_frac_of_hbm": fused_bytes / (tf * 1e-3) / (HBM_PEAK_GBS * 1e9)}
        print(f"consumer {ax}: fused {tf * 1e3:.1f} us, two passes {tt * 1e3:.1f} us", file=sys.stderr, flush=True)
    wm.check(wm.lib.weno_set_async(0))
    for x in (vf, L, R, rhs):
        x.free()
    with open(path, "w") as fh:
        json.dump(out, fh, indent=1)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--m", type=int, default=199)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-precision", action="store_true",
                    help="skip the FP32/BF16-state leg (the metric's 'FP32-form speedup @ err')")
    ap.add_argument("--flags", type=int, default=0, help="extra weno_reconstruct flags")
    ap.add_argument("--suite", default=None, help="also run the secondary measurements into this JSON file")
    ap.add_argument("--adjoint", default=None,
                    help="run only the suite's adjoint measurement (WENO5/7 VJP at 128^3x8) into this JSON file")
    ap.add_argument("--batched", default=None,
                    help="run only the suite's batched measurement (64 blocks, hipGraph) into this JSON file")
    ap.add_argument("--precision", default=None,
                    help="run only the suite's precision measurement (FP32/BF16 state) into this JSON file")
    ap.add_argument("--consumer", default=None,
                    help="time the fused reconstruction + flux divergence against two passes into this JSON file")
    ap.add_argument("--families", default=None,
                    help="time every kernel family against its roofline into this JSON file, then exit")
    ap.add_argument("--copy-yardstick", action="store_true",
                    help="run one plain streaming copy of a face array first (counter calibration)")
    ap.add_argument("--record-traffic", nargs=2, metavar=("FETCH_CSV", "WRITE_CSV"),
                    help="CPU only: turn two rocprofv3 --pmc counter files into the traffic record")
    args = ap.parse_args()
    if args.record_traffic:
        print(json.dumps(record_traffic(*args.record_traffic, m=args.m)))
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: there is no multi-GPU path (the sweep would shard by "
              "sub-domain, one GPU per sub-box, with a buff_size halo exchange between neighbours).",
              file=sys.stderr)
        sys.exit(2)

    if args.families:
        families(args.families, m=args.m)
        return
    if args.consumer:
        consumer(args.consumer, m=args.m)
        return
    for part in ("adjoint", "batched", "precision"):
        if getattr(args, part):
            suite(getattr(args, part), parts=(part,))
            return

    import wenostandalone_amd as wm
    from wenostandalone_amd import fields
    from wenostandalone_amd.device import DeviceArray

    m, v, buff = args.m, 8, 6
    wm.check(wm.lib.weno_init(0))
    g = wm.make_grid(m, v=v, buff=buff)
    f = fields.seeded_field(m, m, m, v, buff, seed=0)
    vf = DeviceArray.from_host(f)
    del f
    nx = m + 1 + 2 * buff
    ncore = nx - 6
    coefs, faces = [], []
    for d in range(3):
        devs = [DeviceArray.from_host(c) for c in fields.harness_coefficients(ncore, fields.DIR_OFFSETS[d])]
        coefs.append((devs, wm.Coeffs(*[x.ptr for x in devs])))
        L = DeviceArray((nx, nx, nx, v))
        R = DeviceArray((nx, nx, nx, v))
        wm.check(wm.lib.weno_order1_copy(C.byref(g), d, wm.F64, vf.vp, L.vp, R.vp))
        faces.append((L, R))

    wm.check(wm.lib.weno_set_async(1))
    stream = wm.lib.weno_get_stream()
    h = hip()
    if args.copy_yardstick:
        # one plain streaming copy of a face array (read + write of 8*212^3*8
        # bytes): the reference kernel for scaling HBM traffic counters
        src_c, dst_c = faces[0]
        wm.check(wm.lib.weno_stream_copy(dst_c.vp, src_c.vp, src_c.nbytes))
        wm.check(wm.lib.weno_sync())

    def step(evs=None):
        for d in range(3):
            if evs is not None:
                h.hipEventRecord(evs[d][0], stream)
            wm.check(wm.lib.weno_reconstruct(C.byref(g), 5, d, wm.SRC_FIELD, wm.F64, vf.vp,
                                             C.byref(coefs[d][1]), faces[d][0].vp, faces[d][1].vp,
                                             args.flags))
            if evs is not None:
                h.hipEventRecord(evs[d][1], stream)

    for _ in range(args.warmup):
        step()
    wm.check(wm.lib.weno_sync())

    evs = []
    for _ in range(args.steps):
        e = []
        for d in range(3):
            a, b = C.c_void_p(), C.c_void_p()
            h.hipEventCreate(C.byref(a))
            h.hipEventCreate(C.byref(b))
            e.append((a, b))
        evs.append(e)
    t_begin, t_end = C.c_void_p(), C.c_void_p()
    h.hipEventCreate(C.byref(t_begin))
    h.hipEventCreate(C.byref(t_end))

    h.hipDeviceSynchronize()
    w0 = time.perf_counter()
    h.hipEventRecord(t_begin, stream)
    for s in range(args.steps):
        step(evs[s])
    h.hipEventRecord(t_end, stream)
    wm.check(wm.lib.weno_sync())
    h.hipDeviceSynchronize()
    wall = time.perf_counter() - w0
    ms = C.c_float()
    h.hipEventElapsedTime(C.byref(ms), t_begin, t_end)
    total_ms = ms.value
    per_dir = [[], [], []]
    for e in evs:
        for d in range(3):
            h.hipEventElapsedTime(C.byref(ms), e[d][0], e[d][1])
            per_dir[d].append(ms.value)
    wm.check(wm.lib.weno_set_async(0))

    faces_per_sweep = 2 * ncore * nx * nx * v
    ms_per_step = total_ms / args.steps
    value = 3 * faces_per_sweep / (ms_per_step * 1e-3)
    # algorithmic bytes per launch: distinct inputs (core +- 2 along the sweep)
    # and the core-range outputs, FP64
    bytes_launch = ((ncore + 4) * nx * nx * v + faces_per_sweep) * 8
    launch_ms = float(np.mean([x for p in per_dir for x in p]))
    achieved = bytes_launch / (launch_ms * 1e-3) / 1e9
    out = {
        "metric": "WENO5 cell-faces/s, 200^3x8 FP64 (% HBM roofline); FP32-form speedup @ err",
        "value": value,
        "unit": "face values/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded field recipe 0: sine modes + step + noise; harness coefficients)",
        "config": {"workload": f"WENO5-JS x+y+z sweeps from v_vf, {m + 1}^3 interior x {v} vars, "
                               f"buff_size {buff}, core-range outputs",
                   "m_cells": m, "v_size": v, "buff_size": buff, "parallelism": "single GPU"},
        "per_sweep_ms": {k: float(np.median(p)) for k, p in zip("xyz", per_dir)},
        "wall_s_timed_region": wall,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "kernel": "sweep5_kernel (WENO5-JS FP64, producer/consumer waves)",
                     "bytes_per_launch": bytes_launch, "avg_launch_ms": launch_ms},
    }
    out["roofline"]["traffic"] = recorded_traffic() if m == 199 and args.flags == 0 else None
    for L, R in faces:
        L.free()
        R.free()
    vf.free()
    if not args.no_precision:
        prec = precision_leg(h, stream, m, v, buff, args.steps, args.warmup)
        out["fp32_speedup"] = prec["fp32_speedup"]
        out["fp32_max_rel_err"] = prec["fp32_max_rel_err"]
        out["fp32_rms_rel_err"] = prec["fp32_rms_rel_err"]
        out["fp32_form"] = prec
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(m, v, buff)
    if args.suite:
        suite(args.suite)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
