"""Benchmark: frequency-point solves/s of the harmonic sweep on one MI355X.

Workload (default): 3D acoustic cavity 1.0 x 0.8 x 0.6 m, 100 x 80 x 60 Hex20
fluid elements (1,976,881 DOF), air, Robin impedance face z = 0.6 with
Z = rho c (1 + 0.5i), unit point source at (0.2, 0.3, 0.1); 256 frequencies
10-2000 Hz.  One step = one sweep of all 256 frequencies, each solved to a
FP64 relative residual of 1e-8 (complex-FP32 COCG with the shifted-Laplacian
multigrid preconditioner inside FP64 iterative refinement), continuously
batched on 32 slots: a slot takes the next frequency as soon as its current
one passes the FP64 residual check.  Synthetic mesh, matrices resident in HBM
before the timed region.

Also reported: the plain per-frequency FP64 GPU solve on a sample of the same
frequencies (the denominator of the mixed-vs-FP64 speedup), sweep latency
p50/p99, the roofline of the fused multi-frequency SpMV (HIP events around
its launches), element assembly of a 200 x 200 x 4 Brick8 steel block (K, M,
C = 1e-5 K + 2 M) against the HBM roofline, and the CPU oracle timed on this
host.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "elpaso-core_amd"))

HBM_PEAK_GBS = 8000.0
# HBM bytes per dispatch measured with rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE
# on the default workloads, scaled by the factors a 1 GiB copy gives (x2, x1);
# the commands and counter values are in the profile below.
PMC_PROFILE = "profiles/pmc_traffic_operator_variants_and_assembly_kernels.md"
PMC_TRAFFIC = {("thgrp_op_kernel<0>", "100,80,60"): 2.934e9, ("tgrp_op_kernel<0, 4>", "100,80,60"): 2.933e9,
               ("row_op_kernel<0>", "100,80,60"): 2.790e9, ("rgrp_op_kernel<0, 4>", "100,80,60"): 2.833e9,
               ("tile_op_kernel<0>", "100,80,60"): 2.884e9, ("tgrp_op_kernel<0, 2>", "100,80,60"): 2.900e9,
               ("hex8_node_kernel<3, false, true>", "200,200,4"): 1.116e9}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def operator_kernel(B, variant):
    """Name of the level-0 FP32 operator kernel a sweep ran (the q = A p launch),
    from the slot count and the operator variant in the sweep statistics."""
    if B != 32:
        return f"op_kernel<float, {B}, 0>"
    if variant == -1:
        return "row_op_kernel<0>"
    if variant in (-2, -4):
        return f"rgrp_op_kernel<0, {-variant}>"
    if 0 < variant < 1000:
        return "tile_op_kernel<0>"
    if variant // 1000 == 6:
        return "thgrp_op_kernel<0>"
    return f"tgrp_op_kernel<0, {variant // 1000}>"


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--nel", type=str, default="100,80,60", help="Hex20 elements per direction")
    ap.add_argument("--nfreq", type=int, default=256)
    ap.add_argument("--block", type=int, default=32)
    ap.add_argument("--op-tile-rows", type=int, default=0,
                    help="32-slot operator variant: 0 timed at set-up, -1 row kernel, -2/-4 rows grouped to a "
                         "wavefront, t LDS tiles of t rows, 1000 R + t tiles with rows grouped R (e.g. 4032), "
                         "6000 + t tiles grouped 4 with their frequencies split over two workgroups")
    ap.add_argument("--fp64-sample", type=int, default=32,
                    help="frequencies timed with the FP64 solver (the error against FP64 covers all of them)")
    ap.add_argument("--no-config3", action="store_true", help="skip the coupled plate/cavity sweep")
    ap.add_argument("--no-assembly", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-peaks", action="store_true", help="skip re-measuring copy bandwidth and FP64 GEMM")
    ap.add_argument("--no-config0", action="store_true",
                    help="skip the 20x16x12 Fluid8 cavity leg (GPU sweep vs the CPU oracle's direct solves)")
    ap.add_argument("--cpu-procs", type=int, default=0,
                    help="processes of the CPU oracle's frequency-parallel run (capped by the host's cores)")
    ap.add_argument("--assembly-nel", type=str, default="200,200,4")
    ap.add_argument("--serve-requests", type=int, default=16, help="requests of the serving stream (0: skip)")
    ap.add_argument("--node-results", type=int, default=13,
                    help="frequencies of the sweep whose element/node results are timed (0: skip)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="also run 3 device-to-device copies of 1 GiB (PMC byte-counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on one GPU: this build has no multi-GPU path (frequency blocks would shard "
              "across GPUs with no per-iteration communication; see the design notes).", file=sys.stderr)
        sys.exit(2)

    from elpaso_amd import Context, GpuModel
    from elpaso_amd import _lib
    from elpaso_amd import meshgen as MG

    nx, ny, nz = (int(v) for v in args.nel.split(","))
    ctx = Context(0)
    t0 = time.perf_counter()
    model = MG.acoustic_cavity(nx, ny, nz, quadratic=True)
    g = GpuModel(ctx, model)
    g.assemble()
    log(f"[bench] cavity Hex20 {nx}x{ny}x{nz}: n={g.n} nnz={g.nnz} setup {time.perf_counter() - t0:.1f}s")
    freqs = MG.linear_frequencies(10.0, 2000.0, args.nfreq)
    B = args.block
    opts = _lib.default_opts(block=B, op_tile_rows=args.op_tile_rows)

    def run_sweep():
        t = time.perf_counter()
        X, rel, it, st = g.sweep(freqs, opts=opts, want_x=False)
        return time.perf_counter() - t, rel, it, st

    for w in range(args.warmup):
        dt, rel, it, st = run_sweep()
        log(f"[bench] warmup sweep {w}: {dt:.2f}s iters max {it.max()} relres max {rel.max():.2e} "
            f"(mg setup {st.setup_ms:.0f} ms, operator variant {st.op_tile_rows})")
    lat, rels, iters = [], [], []
    spmv_ms = spmv_calls = 0.0
    spmv_bytes = 0
    t_start = time.perf_counter()
    for k in range(args.steps):
        dt, rel, it, st = run_sweep()
        lat.append(dt)
        rels.append(rel)
        iters.append(it)
        spmv_ms += st.spmv_ms
        spmv_calls += st.spmv_calls
        spmv_bytes = st.spmv_bytes
        log(f"[bench] step {k}: sweep of {len(freqs)} freqs {dt:.2f}s iters {int(it.min())}-{int(it.max())} "
            f"mean {it.mean():.1f} relres max {rel.max():.2e} refinements {st.refinements} "
            f"spmv launches {st.spmv_calls}")
    elapsed = time.perf_counter() - t_start
    n_solved = len(freqs) * args.steps
    rels = np.concatenate(rels)
    iters = np.concatenate(iters)
    converged = int((rels <= opts.rtol).sum())
    value = n_solved / elapsed

    # plain per-frequency FP64 GPU solve on a sample of the same frequencies,
    # and the error of the mixed solutions against it (the mixed run of the
    # sample on the same 32 slots gives the sweep's own bits: a frequency's
    # result does not depend on the others in the batch)
    fp64 = None
    if args.fp64_sample > 0:
        # timed: the plain per-frequency FP64 solve on an evenly spaced sample
        idx = np.linspace(0, len(freqs) - 1, min(args.fp64_sample, len(freqs))).round().astype(int)
        o64 = _lib.default_opts(precision=_lib.PREC_FP64)
        t = time.perf_counter()
        _, rel64s, it64s, _ = g.sweep(freqs[idx], opts=o64, want_x=False)
        dt64 = time.perf_counter() - t
        # untimed: every frequency of the sweep solved both ways, the error of
        # the mixed solution against the FP64 one per frequency
        X64, rel64, it64, _ = g.sweep(freqs, opts=o64, want_x=True)
        Xm, relm, itm, _ = g.sweep(freqs, opts=opts, want_x=True)
        err = np.linalg.norm(Xm - X64, axis=1) / np.linalg.norm(X64, axis=1)
        del Xm, X64
        fp64 = {"timed_sample_freqs": len(idx), "fp64_solves_per_s": len(idx) / dt64,
                "fp64_relres_max": float(max(rel64.max(), rel64s.max())),
                "fp64_iters_mean": float(it64.mean()), "fp64_iters_max": int(it64.max()),
                "speedup_mixed_batched_vs_fp64": value / (len(idx) / dt64),
                "err_mixed_vs_fp64": {"freqs": len(freqs), "max": float(err.max()),
                                      "median": float(np.median(err)),
                                      "per_freq": [float("%.3g" % e) for e in err]}}
        log(f"[bench] fp64: {len(idx) / dt64:.2f} solves/s, error vs fp64 max {err.max():.2e} "
            f"median {np.median(err):.2e} over {len(freqs)} frequencies")

    # element and node results (pressure gradients at the nodes) of a sample
    # of the sweep's own solutions
    node_res = None
    if args.node_results > 0:
        idx = np.linspace(0, len(freqs) - 1, args.node_results).round().astype(int)
        Xs, _, _, _ = g.sweep(freqs[idx], opts=opts, want_x=True)
        g.node_results(Xs[:1], freqs[idx[:1]])  # warm-up
        tm = {}
        t = time.perf_counter()
        _, nd = g.node_results(Xs, freqs[idx], cells=False, timing=tm)
        wall = time.perf_counter() - t
        assert np.isfinite(nd).all()
        node_res = {"workload": f"grad p at the nodes of the Hex20 cavity, {len(idx)} of the sweep's solutions",
                    "kernel_ms_per_freq": round(tm["kernel_ms"] / len(idx), 3),
                    "wall_ms_per_freq": round(1e3 * wall / len(idx), 2),
                    "kernel": "fluid_item_gradp_kernel<20> + node_results_kernel",
                    "bound": "latency of the gathers of the solutions (20 per element and node)",
                    "note": "wall time includes the host copies of the solutions in and the 9 results per node out"}
        del Xs, nd
        log(f"[bench] node results: {node_res}")

    per_launch_ms = spmv_ms / max(1, spmv_calls)
    achieved = spmv_bytes / (per_launch_ms * 1e-3) / 1e9 if per_launch_ms > 0 else 0.0
    op_kernel = operator_kernel(B, int(st.op_tile_rows))
    traffic = PMC_TRAFFIC.get((op_kernel, args.nel))
    roof = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
            "traffic_source": PMC_PROFILE if traffic else None,
            "kernel": f"{op_kernel} (fused {B}-frequency FP32 operator, level 0, variant {int(st.op_tile_rows)})",
            "bytes_per_launch": int(spmv_bytes), "avg_launch_ms": round(per_launch_ms, 4),
            "launches": int(spmv_calls)}

    assembly = None
    if not args.no_assembly:
        ax, ay, az = (int(v) for v in args.assembly_nel.split(","))
        sm = MG.steel_block(ax, ay, az)
        sg = GpuModel(ctx, sm)
        sg.assemble()
        ms = sg.assemble_timed(repeats=10)
        n_nodes = len(sm.node_ids)
        alg = 3 * sg.nnz * 8 + n_nodes * 3 * 8 + sm.n_elements * 8 * 4
        assembly = {"workload": f"steel Brick8 {ax}x{ay}x{az} jittered, K+M+C FP64 CSR",
                    "n": sg.n, "nnz": sg.nnz, "ms": round(ms, 4), "algorithmic_bytes": alg,
                    "achieved_GBs": round(alg / (ms * 1e-3) / 1e9, 1),
                    "frac_hbm": round(alg / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                    "kernel": "hex8_node_kernel<3, false, true>",
                    "traffic": PMC_TRAFFIC.get(("hex8_node_kernel<3, false, true>", args.assembly_nel))}
        log(f"[bench] assembly: {assembly}")
        if args.node_results > 0:
            # brick centroid stresses and their node averages for synthetic
            # displacements of the block (timing only)
            rng = np.random.default_rng(0)
            nfr = args.node_results
            Xb = rng.standard_normal((nfr, sg.n)) + 1j * rng.standard_normal((nfr, sg.n))
            fb = np.linspace(10.0, 2000.0, nfr)
            sg.node_results(Xb[:1], fb[:1])  # warm-up
            tm = {}
            t = time.perf_counter()
            sg.node_results(Xb, fb, timing=tm)
            wall = time.perf_counter() - t
            assembly["node_results"] = {
                "workload": f"brick centroid stresses + node averages, {nfr} synthetic displacement fields",
                "kernel_ms_per_freq": round(tm["kernel_ms"] / nfr, 4), "wall_ms_per_freq": round(1e3 * wall / nfr, 2),
                "kernel": "brick_centroid_kernel + node_results_kernel"}
            del Xb
        sg.close()
        # the coupled vibroacoustic model at full size: steel plate 100x100x2
        # Brick8 (loss factor) on an air cavity 100x100x120 Fluid8, 10,000
        # interface Quad4: K, M, S assembled together
        cm = MG.coupled_plate_cavity()
        cg = GpuModel(ctx, cm)
        cg.assemble()
        cms = cg.assemble_timed(repeats=5)
        calg = 2 * cg.nnz * 8 + len(cm.node_ids) * 3 * 8 + sum(len(gr.elem_ids) * 8 * 4 for gr in cm.groups)
        assembly["coupled"] = {
            "workload": "coupled plate 100x100x2 Brick8 + cavity 100x100x120 Fluid8 + 10,000 interface Quad4, "
                        "K + M FP64 CSR over the coupled pattern (volume kernels; the interface matrix S is "
                        "a separate small pass)", "n": cg.n, "nnz": cg.nnz, "ms": round(cms, 4),
            "algorithmic_bytes": calg, "achieved_GBs": round(calg / (cms * 1e-3) / 1e9, 1),
            "frac_hbm": round(calg / (cms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}
        log(f"[bench] coupled assembly: {assembly['coupled']}")
        cg.close()

    if args.copy_calibration:
        copy_calibration()

    peaks = None
    if not args.no_peaks:
        peaks = measure_peaks()
        peaks.update({"published_hbm_GBs": HBM_PEAK_GBS, "published_fp64_TFLOPs": 78.6})
        log(f"[bench] measured peaks: {peaks}")

    serving = None
    if args.serve_requests > 0:
        serving = serve_stream(ctx, args.serve_requests)
        log(f"[bench] serving: {serving}")

    config0 = None
    if not args.no_config0:
        config0 = config0_leg(ctx, MG)
        log(f"[bench] config 0: {config0}")

    config3 = None
    if not args.no_config3:
        config3 = config3_leg(ctx, MG)
        log(f"[bench] config 3: {config3}")

    cpu = None
    if not args.no_cpu:
        steel = MG.steel_block(*(int(v) for v in args.assembly_nel.split(","))) if assembly else None
        cpu = cpu_baseline(model, freqs, steel=steel, config0=config0 is not None, procs=args.cpu_procs,
                           post_model=MG.acoustic_cavity() if node_res else None)
        if node_res:
            # the GPU leg's element-node gradients: 20 per Fluid20 element of the cavity
            gpu_us = node_res["kernel_ms_per_freq"] * 1e3 / (model.n_elements * 20)
            node_res["gpu_us_per_element_node"] = round(gpu_us, 6)
            node_res["gpu_vs_cpu_1core_per_element_node"] = round(cpu["node_results_1core"]["us_per_element_node"]
                                                                  / gpu_us)
        if assembly:
            assembly["gpu_vs_cpu_1core"] = round(cpu["brick8_assembly_1core_s"]["assembly"] / (assembly["ms"] * 1e-3))
        if config0:
            config0["gpu_vs_cpu_parallel"] = round(config0["gpu_solves_per_s"] /
                                                   cpu["config0"]["cpu_parallel_solves_per_s"], 1)

    line = {
        "metric": "freq-point solves/s at 1e-8 rel residual, mixed vs FP64; p99 sweep latency",
        "value": round(value, 3), "unit": "freq-point solves/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(1e3 * elapsed / args.steps, 2), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "complex-f32 Krylov + f64 refinement",
        "data": "synthetic",
        "config": {"workload": f"acoustic cavity 1.0x0.8x0.6 m, Hex20 {nx}x{ny}x{nz} ({g.n} DOF), "
                               f"{len(freqs)} freqs 10-2000 Hz on {B} slots, Robin face Z=rho c(1+0.5i)",
                   "n_dof": g.n, "nnz": g.nnz, "slots": B,
                   "precond": "two-level cycle: damped l1-Jacobi around the exact Hex8 box solve "
                              "(fast diagonalization)",
                   "method": "COCG (FP32) + FP64 iterative refinement",
                   "operator_variant": int(st.op_tile_rows), "post_smoothing_tile_rows": int(st.box_tile_rows),
                   "parallelism": "single GPU"},
        "converged": f"{converged}/{len(rels)}", "relres_max": float(rels.max()),
        "iters_mean": float(iters.mean()), "iters_max": int(iters.max()),
        "sweep_latency_s": {"p50": float(np.percentile(lat, 50)), "p99": float(np.percentile(lat, 99))},
        "fp64_baseline": fp64, "roofline": roof, "assembly": assembly, "serving": serving,
        "node_results": node_res,
        "cpu_baseline": cpu, "config0": config0, "config3": config3, "measured_peaks": peaks,
    }
    print(json.dumps(line), flush=True)


def copy_calibration(nbytes=1 << 30, repeats=3):
    """Plain device-to-device copies of known size: hardware byte counters
    (FETCH_SIZE, WRITE_SIZE) read on them give the factor between counter
    and true bytes on this GPU."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    for ptr in (a, b):
        if hip.hipMalloc(ctypes.byref(ptr), ctypes.c_size_t(nbytes)) != 0:
            raise RuntimeError("hipMalloc failed")
    hip.hipMemset(a, 1, ctypes.c_size_t(nbytes))
    for _ in range(repeats):
        hip.hipMemcpy(b, a, ctypes.c_size_t(nbytes), 3)  # hipMemcpyDeviceToDevice
    hip.hipDeviceSynchronize()
    hip.hipFree(a)
    hip.hipFree(b)
    log(f"[bench] copy calibration: {repeats} x {nbytes} B device-to-device")


def measure_peaks(copy_bytes=1 << 30, n=8192, repeats=5):
    """The GPU's attainable peaks re-measured on this box: device-to-device
    copy of 1 GiB (HBM read + write bytes per second) and rocBLAS FP64 GEMM
    (n x n x n).  Reported beside the published peaks the rooflines use."""
    import ctypes
    C = ctypes
    hip = C.CDLL("libamdhip64.so")
    out = {}
    a, b = C.c_void_p(), C.c_void_p()
    try:
        for ptr in (a, b):
            if hip.hipMalloc(C.byref(ptr), C.c_size_t(copy_bytes)) != 0:
                raise RuntimeError("hipMalloc failed")
        hip.hipMemset(a, 1, C.c_size_t(copy_bytes))
        hip.hipMemcpy(b, a, C.c_size_t(copy_bytes), 3)
        hip.hipDeviceSynchronize()
        t = time.perf_counter()
        for _ in range(repeats):
            hip.hipMemcpy(b, a, C.c_size_t(copy_bytes), 3)  # hipMemcpyDeviceToDevice
        hip.hipDeviceSynchronize()
        dt = (time.perf_counter() - t) / repeats
        out["hbm_copy_GBs"] = round(2 * copy_bytes / dt / 1e9, 1)
    finally:
        for ptr in (a, b):
            if ptr.value:
                hip.hipFree(ptr)
    try:
        rb = C.CDLL("librocblas.so")
    except OSError:
        return out
    h = C.c_void_p()
    if rb.rocblas_create_handle(C.byref(h)) != 0:
        return out
    bufs = [C.c_void_p() for _ in range(3)]
    try:
        nb = n * n * 8
        for ptr in bufs:
            if hip.hipMalloc(C.byref(ptr), C.c_size_t(nb)) != 0:
                raise RuntimeError("hipMalloc failed")
            hip.hipMemset(ptr, 0x3F, C.c_size_t(nb))  # every double 4.8e-4
        one, zero = C.c_double(1.0), C.c_double(0.0)
        rb.rocblas_dgemm.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                     C.c_void_p, C.c_void_p, C.c_int, C.c_void_p, C.c_int, C.c_void_p,
                                     C.c_void_p, C.c_int]
        args = (h, 111, 111, n, n, n, C.byref(one), bufs[0], n, bufs[1], n, C.byref(zero), bufs[2], n)
        if rb.rocblas_dgemm(*args) != 0:
            return out
        hip.hipDeviceSynchronize()
        t = time.perf_counter()
        for _ in range(repeats):
            rb.rocblas_dgemm(*args)
        hip.hipDeviceSynchronize()
        dt = (time.perf_counter() - t) / repeats
        out["fp64_dgemm_TFLOPs"] = round(2.0 * n ** 3 / dt / 1e12, 1)
        out["dgemm_n"] = n
    finally:
        for ptr in bufs:
            if ptr.value:
                hip.hipFree(ptr)
        rb.rocblas_destroy_handle(h)
    return out


def serve_stream(ctx, n_req, rate=20.0, seed=0, budget_gb=200, resident=16):
    """The serving workload: a seeded Poisson stream (mean 20 requests/s) of
    sweep requests, each an acoustic cavity of the Fluid8 20x16x12 family
    (1.0x0.8x0.6 m, Robin face, point source) or a coupled plate/cavity of
    the 100x100x(2+120) family (1.0x1.0x1.2 m cavity, 5 mm clamped steel
    plate, eta 0.01), with element counts drawn log-uniform for 5k-1.5M DOF,
    16-512 frequencies uniform in 10-2000 Hz and Z = rho c (a + bi), a, b
    uniform in [0.2, 2].  Models are built (not assembled) before the stream
    starts; each request is submitted at its arrival time to a server with a
    200 GB HBM budget and up to 16 resident requests, which assembles and
    sets it up and batches its frequency slots with the other resident
    requests of its precision.  Latency = submit to result."""
    from elpaso_amd import GpuModel, Server
    from elpaso_amd import _lib
    from elpaso_amd import meshgen as MG
    rng = np.random.default_rng(seed)
    reqs = []
    for _ in range(n_req):
        coupled = bool(rng.integers(0, 2))
        ndof = float(np.exp(rng.uniform(np.log(5e3), np.log(1.5e6))))
        a, b = rng.uniform(0.2, 2.0, size=2)
        nf = int(rng.integers(16, 513))
        freqs = np.sort(rng.uniform(10.0, 2000.0, size=nf))
        if coupled:
            # n ~ 1.2 n_xy^3 (cavity 1.2 n_xy deep) + 9 n_xy^2 (plate)
            nxy = max(4, int(round((ndof / 1.2) ** (1.0 / 3.0))))
            reqs.append(("coupled", (nxy, 2, int(round(1.2 * nxy))), 0j, freqs))
        else:
            # n ~ 480 s^3 for 10s x 8s x 6s elements
            sc = (ndof / 480.0) ** (1.0 / 3.0)
            nel = tuple(max(2, int(round(sc * r))) for r in (10, 8, 6))
            reqs.append(("acoustic", nel, 1.21 * 343.0 * (a + 1j * b), freqs))
    arrivals = np.cumsum(rng.exponential(1.0 / rate, size=n_req))
    models = []
    for kind, nel, Z, _ in reqs:
        if kind == "acoustic":
            models.append(GpuModel(ctx, MG.acoustic_cavity(*nel, Z=Z)))
        else:
            models.append(GpuModel(ctx, MG.coupled_plate_cavity(n_xy=nel[0], n_plate_z=nel[1], n_cav_z=nel[2])))
    ndofs = [gm.n for gm in models]
    sv = Server(ctx, budget_gb << 30, max_slots=resident)
    opts = _lib.default_opts(block=32)
    t0 = time.perf_counter()
    jobs = []
    for k, (gm, (_, _, Z, freqs)) in enumerate(zip(models, reqs)):
        delay = arrivals[k] - (time.perf_counter() - t0)
        if delay > 0:
            time.sleep(delay)
        jobs.append(sv.submit(gm, freqs, Z=Z, opts=opts))
    lat, ok, solves, per = [], 0, 0, []
    for j, (kind, nel, _, freqs), n in zip(jobs, reqs, ndofs):
        _, rel, it, l = sv.wait(j, want_x=False)
        lat.append(l)
        good = bool((rel <= 1e-8).all())
        ok += int(good)
        solves += len(freqs)
        per.append({"kind": kind, "dof": int(n), "freqs": len(freqs), "latency_s": round(l, 3),
                    "iters_mean": round(float(it.mean()), 1), "converged": good})
        log(f"[bench] served {kind} n={n} {len(freqs)} freqs in {l:.2f}s")
    makespan = time.perf_counter() - t0
    sv.close()
    return {"workload": "Poisson stream of acoustic (Fluid8 cavity) and coupled (plate + cavity) sweep requests, "
                        "5k-1.5M DOF log-uniform, 16-512 freqs in 10-2000 Hz, Z = rho c (a + bi)",
            "requests": n_req, "arrival_rate_per_s": rate, "dof_range": [int(min(ndofs)), int(max(ndofs))],
            "frequencies": solves, "all_converged": f"{ok}/{n_req}",
            "latency_s": {"p50": float(np.percentile(lat, 50)), "p99": float(np.percentile(lat, 99))},
            "solves_per_s": solves / makespan, "makespan_s": makespan, "resident_max": resident,
            "hbm_budget_GB": budget_gb, "per_request": per}


def _oracle_direct(args):
    """Worker of the CPU oracle's frequency-parallel run: elPaSo's
    per-frequency flow (re-integration, scatter, BCs, direct LU) for a share
    of the frequencies.  Runs in a spawned process that never touches the GPU."""
    freqs, = args
    sys.path.insert(0, os.path.join(ROOT, "elpaso-core_amd"))
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    from elpaso_amd import meshgen as MG
    m = MG.acoustic_cavity()
    num = O.Numbering(m)
    pat = O.pattern(m, num)
    t = time.perf_counter()
    O.solve_direct(m, np.asarray(freqs), num, pat)
    return time.perf_counter() - t


def config0_leg(ctx, MG):
    """The reference's own CPU-runnable case on the GPU: Fluid8 cavity
    20x16x12 (4,641 DOF), Robin face z = 0.6 with Z = rho c (1 + 0.5i), 100
    frequencies 10-500 Hz, one sweep on 32 slots (solutions stay on the
    device), each Krylov iteration replayed from a captured HIP graph (the
    default at this size), and beside it the same sweep launching every
    kernel (graph = -1).  Its CPU counterpart is timed in the cpu_baseline
    leg."""
    from elpaso_amd import GpuModel
    from elpaso_amd import _lib
    m = MG.acoustic_cavity()
    freqs = MG.linear_frequencies(10.0, 500.0, 100)
    g = GpuModel(ctx, m)
    g.assemble()
    best = {}
    for mode in (0, -1):
        o = _lib.default_opts(block=32, graph=mode)
        g.sweep(freqs, opts=o, want_x=False)  # warm-up
        dts = []
        for _ in range(5):
            t = time.perf_counter()
            _, rel, it, st = g.sweep(freqs, opts=o, want_x=False)
            dts.append(time.perf_counter() - t)
        best[mode] = (float(np.median(dts)), rel, it, st)
    g.close()
    dt, rel, it, st = best[0]
    return {"workload": "acoustic cavity 1.0x0.8x0.6 m, Fluid8 20x16x12 (4,641 DOF), 100 freqs 10-500 Hz",
            "gpu_solves_per_s": round(len(freqs) / dt, 1), "gpu_sweep_ms": round(1e3 * dt, 2),
            "gpu_relres_max": float(rel.max()), "gpu_iters_max": int(it.max()),
            "graph_replays": int(st.graph_replays), "timing": "median of 5 sweeps",
            "gpu_sweep_ms_without_graph": round(1e3 * best[-1][0], 2)}


def config3_leg(ctx, MG, sample=8):
    """The coupled vibroacoustic case: steel plate 100x100x2 Brick8 (clamped
    edges, loss factor 0.01) on an air cavity 100x100x120 Fluid8 through
    10,000 interface Quad4, 1,326,130 DOF, unit normal force at a seeded plate
    node, 512 frequencies 20-1000 Hz on 32 slots.  The inner Krylov solve is
    FP64 here (the plate's stiffness is too ill-conditioned for FP32 inner
    solves); the preconditioner is the exact box solve of the cavity and the
    block cyclic reduction of the plate, one factorization per window of 32
    frequencies prepared in the background.  The first sweep keeps its
    solutions (and warms up); the second is timed with the solutions left on
    the device.  Beside it: the plain per-frequency FP64 solve (the plate
    factored at each frequency) on an evenly spaced sample, and the error of
    the sweep's solutions against it."""
    from elpaso_amd import GpuModel
    from elpaso_amd import _lib
    m = MG.coupled_plate_cavity()
    freqs = MG.linear_frequencies(20.0, 1000.0, 512)
    g = GpuModel(ctx, m)
    g.assemble()
    opts = _lib.default_opts(block=32)
    X, rel0, _, _ = g.sweep(freqs, opts=opts, want_x=True)
    t = time.perf_counter()
    _, rel, it, st = g.sweep(freqs, opts=opts, want_x=False)
    dt = time.perf_counter() - t
    idx = np.linspace(0, len(freqs) - 1, sample).round().astype(int)
    o64 = _lib.default_opts(precision=_lib.PREC_FP64)
    t = time.perf_counter()
    X64, rel64, it64, st64 = g.sweep(freqs[idx], opts=o64, want_x=True)
    dt64 = time.perf_counter() - t
    err = np.linalg.norm(X[idx] - X64, axis=1) / np.linalg.norm(X64, axis=1)
    del X, X64
    g.close()
    return {"workload": "coupled plate 1.0x1.0x0.005 m Brick8 100x100x2 (clamped, eta 0.01) + air cavity "
                        "1.0x1.0x1.2 m Fluid8 100x100x120, 10,000 interface Quad4 (1,326,130 DOF), "
                        "512 freqs 20-1000 Hz on 32 slots",
            "solves_per_s": round(len(freqs) / dt, 2), "sweep_s": round(dt, 3),
            "converged": f"{int((rel <= 1e-8).sum())}/{len(rel)}", "relres_max": float(max(rel.max(), rel0.max())),
            "iters_mean": float(it.mean()), "iters_max": int(it.max()),
            "precision": "complex-f64 Krylov (batched on 32 slots) + f64 refinement",
            "precond": "box solve (cavity) + block cyclic reduction (plate), 16 windows",
            "structure_factorizations": st.refactors, "factorization_wait_ms": round(st.refactor_ms, 1),
            "fp64_per_frequency": {"sample_freqs": len(idx), "solves_per_s": round(len(idx) / dt64, 3),
                                   "iters": it64.tolist(), "relres_max": float(rel64.max())},
            "speedup_batched_vs_per_frequency_fp64": round((len(freqs) / dt) / (len(idx) / dt64), 2),
            "err_vs_fp64_per_frequency": [float("%.3g" % e) for e in err]}


def _cpu_config0(procs):
    """CPU side of the reference case: the oracle's restatement of elPaSo's
    per-frequency loop (re-integration + scatter + BCs + SuperLU, standing in
    for MUMPS) on one core for 10 of the frequencies, and over all 100
    frequencies split into contiguous shares across P processes."""
    sys.path.insert(0, os.path.join(ROOT, "elpaso-core_amd"))
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    from elpaso_amd import meshgen as MG
    from threadpoolctl import threadpool_limits
    m = MG.acoustic_cavity()
    freqs = MG.linear_frequencies(10.0, 500.0, 100)
    num = O.Numbering(m)
    pat = O.pattern(m, num)
    tm = []
    with threadpool_limits(limits=1):  # one core: SuperLU's BLAS calls single-threaded
        t = time.perf_counter()
        O.solve_direct(m, freqs[::10], num, pat, timings=tm)
        dt1 = time.perf_counter() - t
    out = {"cpu_1core_solves_per_s": round(10 / dt1, 3),
           "cpu_1core_ms_per_freq": {"assembly": round(1e3 * float(np.mean([a for a, _ in tm])), 1),
                                     "lu_solve": round(1e3 * float(np.mean([b for _, b in tm])), 1)}}
    p = max(1, procs)
    import multiprocessing as mp
    shares = [freqs[i * len(freqs) // p:(i + 1) * len(freqs) // p] for i in range(p)]
    # one thread per process (inherited at spawn): BLAS threads of P processes
    # on the same cores would contend and slow every LU by orders of magnitude
    keys = ("OMP_NUM_THREADS", "OPENBLAS_NUM_THREADS", "MKL_NUM_THREADS")
    saved = {k: os.environ.get(k) for k in keys}
    os.environ.update({k: "1" for k in keys})
    try:
        t = time.perf_counter()
        with mp.get_context("spawn").Pool(p) as pool:
            pool.map(_oracle_direct, [(s.tolist(),) for s in shares])
        dtp = time.perf_counter() - t
    finally:
        for k, v in saved.items():
            if v is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = v
    out["cpu_procs"] = p
    out["cpu_parallel_solves_per_s"] = round(len(freqs) / dtp, 2)
    out["sample"] = ("1 core: 10 of the 100 frequencies; parallel: all 100, contiguous shares per process "
                     "(process start-up and per-process symbolic set-up included)")
    return out


def host_info(requested):
    """The host the CPU baseline runs on: logical CPUs, the CPUs this process
    may use, the CPU model, and the processes the baseline uses (the box's
    CPU share for one GPU, 16, unless asked otherwise)."""
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    n = os.cpu_count() or 1
    aff = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else n
    procs = requested if requested > 0 else min(16, aff)
    return {"nproc": n, "usable_cpus": aff, "cpu_model": model, "processes": procs}


def _oracle_assembly_worker(lo, hi, barrier, out):
    """One process of the parallel oracle assembly: the Brick8 block's
    symbolic phase (untimed), a barrier, then K + M + C of elements [lo, hi)
    into the full pattern."""
    sys.path.insert(0, os.path.join(ROOT, "elpaso-core_amd"))
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    from elpaso_amd import meshgen as MG
    m = MG.steel_block()
    num = O.Numbering(m)
    pat = O.pattern(m, num)
    g = m.groups[0]
    g.elem_ids, g.conn = g.elem_ids[lo:hi], g.conn[lo:hi]
    barrier.wait()
    t = time.perf_counter()
    O.assemble_km(m, num, pat)
    out.put(time.perf_counter() - t)


def _cpu_parallel_assembly(procs):
    """The oracle's K + M + C assembly of the 200x200x4 Brick8 block on P
    processes, each on a contiguous range of elements (into its own copy of
    the pattern; the final sum of the P partial matrices is not timed)."""
    import multiprocessing as mp
    ctx = mp.get_context("spawn")
    n_el = 200 * 200 * 4
    barrier = ctx.Barrier(procs)
    q = ctx.Queue()
    ps = [ctx.Process(target=_oracle_assembly_worker,
                      args=(i * n_el // procs, (i + 1) * n_el // procs, barrier, q)) for i in range(procs)]
    for p in ps:
        p.start()
    times = [q.get() for _ in ps]
    for p in ps:
        p.join()
    return {"processes": procs, "wall_s": round(max(times), 3),
            "sample": "element integration + scatter of the whole block, P contiguous element ranges after a "
                      "barrier (symbolic phase and the sum of the partial matrices untimed)"}


def _cpu_lu_probe():
    """The direct LU (SuperLU, COLAMD ordering, complex128, one thread) of the
    Hex20 cavity's A(w) at 1 kHz on three small meshes: its time and fill
    grow so fast with n that the 1.98M-DOF system is out of reach; the fitted
    growth and the extrapolation to n = 1,976,881 are reported as such."""
    sys.path.insert(0, os.path.join(ROOT, "elpaso-core_amd"))
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    import scipy.sparse.linalg as sla
    from elpaso_amd import meshgen as MG
    from threadpoolctl import threadpool_limits
    pts = []
    for nel in ((10, 8, 6), (15, 12, 9), (20, 16, 12)):
        m = MG.acoustic_cavity(*nel, quadratic=True)
        num = O.Numbering(m)
        pat = O.pattern(m, num)
        rp, ci, A, F = O.dynamic_matrix(m, 2 * np.pi * 1000.0, num, pat)
        Am = O.csr(rp, ci, A, num.n).tocsc()
        with threadpool_limits(limits=1):
            t = time.perf_counter()
            lu = sla.splu(Am)
            dt = time.perf_counter() - t
        pts.append({"mesh": "x".join(map(str, nel)), "n": num.n, "lu_s": round(dt, 3),
                    "lu_entries": int(lu.L.nnz + lu.U.nnz)})
    ln = np.log([p["n"] for p in pts])
    et = float(np.polyfit(ln, np.log([p["lu_s"] for p in pts]), 1)[0])
    ef = float(np.polyfit(ln, np.log([p["lu_entries"] for p in pts]), 1)[0])
    ratio = 1976881 / pts[-1]["n"]
    return {"points": pts, "time_exponent": round(et, 2), "fill_exponent": round(ef, 2),
            "extrapolated_to_config2": {"n": 1976881, "lu_s": float("%.3g" % (pts[-1]["lu_s"] * ratio ** et)),
                                        "lu_GB": float("%.3g" % (pts[-1]["lu_entries"] * ratio ** ef * 16 / 1e9)),
                                        "note": "extrapolation of the three measured points (SuperLU with "
                                                "COLAMD ordering); a nested-dissection ordering, as MUMPS "
                                                "uses, fills less"}}


def cpu_baseline(model, freqs, steel=None, config0=False, procs=16, post_model=None):
    """elPaSo's per-frequency flow restated on the CPU (oracle), one core:
    re-integrate every element at w, scatter into CSR (SetupKM).  The direct
    LU solve (MUMPS in elPaSo, SuperLU here) of the 1.98M-DOF Hex20 system
    does not fit a bounded sample, so the value is an upper bound on CPU
    solves/s (assembly time only).  Also timed here: the oracle's K + M + C
    assembly of the Brick8 block (one core; element integrals in C, scatter
    in numpy) and the reference case's full per-frequency loop."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    t = time.perf_counter()
    num = O.Numbering(model)
    pat = O.pattern(model, num)
    t_sym = time.perf_counter() - t
    t = time.perf_counter()
    O.dynamic_matrix(model, 2 * np.pi * float(freqs[len(freqs) // 2]), num, pat)
    t_asm = time.perf_counter() - t
    log(f"[bench] cpu oracle: symbolic {t_sym:.1f}s, one-frequency assembly {t_asm:.1f}s")
    host = host_info(procs)
    out = {"value": round(1.0 / t_asm, 5), "unit": "freq-point solves/s (upper bound: assembly only)",
           "cores": 1, "kind": "port", "host": host,
           "sample": f"1 frequency ({float(freqs[len(freqs) // 2]):.1f} Hz) of the same model: elPaSo's "
                     f"per-frequency element re-integration + CSR scatter + BCs ({t_asm:.1f} s); the "
                     f"direct LU of this system is beyond a bounded sample and is not included"}
    if steel is not None:
        t = time.perf_counter()
        snum = O.Numbering(steel)
        spat = O.pattern(steel, snum)
        t_ssym = time.perf_counter() - t
        t = time.perf_counter()
        O.assemble_km(steel, snum, spat)
        out["brick8_assembly_1core_s"] = {"assembly": round(time.perf_counter() - t, 2),
                                          "symbolic": round(t_ssym, 2)}
        log(f"[bench] cpu oracle Brick8 assembly: {out['brick8_assembly_1core_s']}")
        out["brick8_assembly_parallel"] = _cpu_parallel_assembly(host["processes"])
        log(f"[bench] cpu oracle Brick8 assembly on {host['processes']} processes: "
            f"{out['brick8_assembly_parallel']}")
    out["config2_lu_probe"] = _cpu_lu_probe()
    log(f"[bench] cpu LU probe: {out['config2_lu_probe']}")
    if config0:
        out["config0"] = _cpu_config0(host["processes"])
        log(f"[bench] cpu oracle config 0: {out['config0']}")
    if post_model is not None:
        # element and node results of one solution (pressure gradients at the
        # nodes), the oracle's pure-Python restatement of elPaSo's loops
        pnum = O.Numbering(post_model)
        Xp = np.random.default_rng(0).standard_normal((1, pnum.n)) + 0j
        t = time.perf_counter()
        O.node_results(post_model, Xp, [100.0], pnum)
        tp = time.perf_counter() - t
        n_en = sum(len(gr.elem_ids) * gr.conn.shape[1] for gr in post_model.groups)
        out["node_results_1core"] = {"sample": f"1 frequency of the Fluid8 20x16x12 cavity ({n_en} element-node "
                                               f"gradients), pure-Python restatement",
                                     "s_per_freq": round(tp, 2), "us_per_element_node": round(1e6 * tp / n_en, 1)}
        log(f"[bench] cpu oracle node results: {out['node_results_1core']}")
    return out

if __name__ == "__main__":
    main()
