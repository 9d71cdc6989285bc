"""Benchmark: rotation rows scored per second on one MI355X.

Workload (default "pair200"): map A 200^3 at 1 A with 3000 pseudo-atoms, map B
100^3 around a moved 800-atom sub-chain (seed 2001), thresholds at the 90th
percentile, -g 16 -s 2 -A 10, vector-product mode: 24,624 Euler rows scored
against a 175^3 grid.  One step = the exhaustive search of the pair (every
row: rotate-and-gather, FFT correlation of the 3 vector channels, translation
argmax) plus the top-10 selection.  Inputs (resampled maps, target spectra,
rotation matrices) are resident in HBM before the timed region.

Prints one JSON line.  --gpus N > 1 exits: this build runs on one GPU.
"""
from __future__ import annotations

import argparse
import contextlib
import io
import json
import os
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "rotations scored/sec at 10° on 200³ pair (1 GPU); top-10 recall vs plain"
PEAK_HBM_GBS = 8000.0

WORKLOADS = {
    "pair200": dict(pair="pair200", seed=2001, s=2.0, g=16.0, A=10.0,
                    text="A 200^3 @1A, 3000 atoms / B 100^3, 800-atom sub-chain; -g 16 -s 2 -A 10; mode V"),
    "pair100": dict(pair="pair100", seed=1001, s=7.0, g=16.0, A=30.0,
                    text="A 100^3 @1A, 500 atoms / B 60^3, 150-atom sub-chain; -g 16 -s 7 -A 30; mode V"),
}


def _stage_bytes(n: int, w: int, b: int) -> dict:
    """Algorithmic HBM bytes per launch (b rotations, V mode, 3 channels) of
    each stage, with the query's support box [lo, lo + w)^3:
      gather   reads the float4 source under the box once, writes 3 float32
               planes inside the box;
      fwd_zy   reads the w slabs' rows inside the box (z inside the box
               too), writes 3 half spectra of the w slabs;
      xcorr    reads those spectra, the 3 target spectra once per launch,
               writes 1 half spectrum per rotation;
      inv_yz   reads that spectrum."""
    nhz = n // 2 + 1
    nh = n * n * nhz
    return {"gather": b * (16 + 12) * w ** 3,
            "fwd_zy": b * 3 * (w * (w + 1) * w * 4 + w * n * nhz * 8),
            "xcorr": b * (3 * w * n * nhz * 8 + nh * 8) + 3 * nh * 8,
            "inv_yz": b * nh * 8}


PEAK_BF16_TFLOPS = 2500.0   # MI355X dense bf16 / fp16 MFMA

PMC_SUMMARIES = [os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles", f)
                 for f in ("xcorr_pair200_pmc.md", "xcorr_mr_pair200.md")]


def pmc_traffic(workload: str, batch: int, schedule, box_width: int):
    """HBM bytes per x-pass launch measured by rocprofv3 (FETCH_SIZE and
    WRITE_SIZE passes, calibrated), read from the committed summary's line
    ``traffic: workload=W batch=B variant=V tile=T box=W bytes=X``; None when
    no line matches this run's workload, batch, x-pass schedule and box."""
    if not isinstance(schedule, dict):
        schedule = {"xpass_variant": 3, "xpass_tile": 8}
    want = {"workload": workload, "batch": str(batch), "variant": str(schedule.get("xpass_variant")),
            "tile": str(schedule.get("xpass_tile")), "box": str(box_width)}
    for path in PMC_SUMMARIES:
        try:
            with open(path) as f:
                for line in f:
                    line = line.strip().strip("`")
                    if not line.startswith("traffic:"):
                        continue
                    kv = dict(t.split("=", 1) for t in line[len("traffic:"):].split() if "=" in t)
                    if all(kv.get(k) == v for k, v in want.items()):
                        return int(kv["bytes"])
        except OSError:
            pass
    return None


def prepare(workload: str):
    from vesper_python_amd import synth
    from vesper_python_amd import VESPER_1_prob_new as V
    from vesper_python_amd import main as cli
    w = dict(WORKLOADS[workload])
    pair = synth.make_config_pair(w["pair"], w["seed"])
    tmp = tempfile.mkdtemp(prefix="vx_bench_")
    paths = synth.write_pair(pair, os.path.join(tmp, "p"))
    w["paths"] = paths
    w["thresholds"] = (float(pair.a.threshold), float(pair.b.threshold))
    with contextlib.redirect_stdout(io.StringIO()):
        m1, m2, n1, n2 = cli._resample_pair(paths["A"], paths["B"], pair.a.threshold, pair.b.threshold,
                                            w["s"], w["g"])
        n1 = V.fastVEC(m1, n1, w["g"])
        n2 = V.fastVEC(m2, n2, w["g"])
    w["src_maps"] = (m1, m2)
    return w, n1, n2


def cpu_baseline(n1, n2, angles, budget_s: float, nrows_max: int):
    """The oracle's per-row scoring (rot_mrc + float32 r2c/c2r per channel +
    channel sum + argmax), single thread, on a strided sample of the rows."""
    from threadpoolctl import threadpool_limits
    from oracle import vesper_ref as ref
    t = ref.MapGrid(data=np.asarray(n1.data, np.float64), xwidth=np.float32(n1.xwidth), orig=n1.orig,
                    vec=n1.vec, cent=n1.cent, xdim=n1.xdim, ave=n1.ave, std=n1.std,
                    std_norm_ave=n1.std_norm_ave)
    spec = ref.target_spectra(ref.target_channels(t, "VecProduct"))
    stride = max(1, len(angles) // nrows_max)
    sample = angles[::stride]
    done = 0
    got = []
    with threadpool_limits(1):
        t0 = time.perf_counter()
        for a in sample:
            sc, tr, _, _ = ref.rot_and_search_fft(np.asarray(n2.data, np.float64), n2.vec, a, spec, t, "VecProduct",
                                                  "ref32")
            got.append((float(sc), tuple(int(v) for v in tr)))
            done += 1
            if time.perf_counter() - t0 > budget_s:
                break
        dt = time.perf_counter() - t0
    return dict(value=done / dt, unit="rotations/s", cores=1, kind="port",
                sample=f"{done} rows (every {stride}th of {len(angles)}) of the same pair, "
                       f"oracle ref32 rot_mrc + scipy.fft float32 r2c/c2r, 1 thread, {dt:.1f} s",
                **host_info(), _rows=list(range(0, done * stride, stride)), _scores=got)


def cpu_stages(w, n1, n2, planes: int = 8):
    """The steps beside the search loop on one host core (the reference runs
    them on the CPU once per search): fastVEC of the query map by the oracle's C
    restatement of doFastVEC (exact order, float64) on `planes` output
    x-planes spread over the grid, extrapolated to all of them; and the FP64
    get_score at one translation by the oracle (NumPy).  The same calls on
    the GPU are in aux_rooflines."""
    from threadpoolctl import threadpool_limits
    from oracle import build_c
    from oracle import vesper_ref as ref
    m1, m2 = w["src_maps"]
    n = int(n2.xdim)
    lib = build_c.load()
    src = np.ascontiguousarray(np.asarray(m2.data, np.float32))
    xs = [int((k + 0.5) * n / planes) for k in range(planes)]
    with threadpool_limits(1):
        t0 = time.perf_counter()
        for x in xs:
            build_c.fastvec_c(lib, src, m2.xwidth, m2.orig, n, n2.xwidth, n2.orig, w["g"], planes=(x, x + 1))
        t_fv = time.perf_counter() - t0
        t = ref.MapGrid(data=np.asarray(n1.data, np.float64), xwidth=np.float32(n1.xwidth), orig=n1.orig,
                        vec=n1.vec, cent=n1.cent, xdim=n1.xdim, ave=n1.ave, std=n1.std,
                        std_norm_ave=n1.std_norm_ave)
        t0 = time.perf_counter()
        ref.get_score(t, np.asarray(n2.data, np.float64), n2.vec, (3, -2, 5), prec="ref64")
        t_gs = time.perf_counter() - t0
    return {"fastvec_query_map": {"seconds_est": round(t_fv / planes * n, 2), "planes_timed": planes, "n": n,
                                  "what": "oracle C doFastVEC restatement, float64, 1 thread, "
                                          f"{planes} of {n} output x-planes timed ({t_fv:.1f} s), scaled"},
            "get_score_fp64": {"seconds": round(t_gs, 3), "what": "oracle NumPy get_score (ref64), 1 thread, "
                                                                   "one translation"}}


def host_info():
    """The host the CPU baseline ran on: CPU model and the CPUs this process
    may use."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"cpu_model": model, "affinity_cpus": len(os.sched_getaffinity(0)), "os_cpu_count": os.cpu_count()}


def baseline_parity(cpu, best, trans):
    """The GPU's scores and translations for the rows the CPU baseline scored
    (they are the oracle's float32 results): max relative difference and
    translation agreement."""
    rows = cpu.pop("_rows")
    ref_rows = cpu.pop("_scores")
    b = best[:, 0].double().cpu().numpy()
    t = trans[:, 0].cpu().numpy()
    rel = [abs(b[r] - s) / abs(s) for r, (s, _) in zip(rows, ref_rows)]
    same = [tuple(int(v) for v in t[r]) == tt for r, (_, tt) in zip(rows, ref_rows)]
    return {"rows": len(rows), "max_rel_diff": float(max(rel)) if rel else None,
            "translations_equal": int(sum(same))}


_POOL = {}


def _pool_init(paths):
    """Worker start: memory-map the maps and spectra written by the parent."""
    from threadpoolctl import threadpool_limits
    _POOL["limits"] = threadpool_limits(1)
    _POOL.update({k: np.load(v, mmap_mode="r") for k, v in paths.items()})


def _pool_rows(args):
    from oracle import vesper_ref as ref
    angles, deadline = args
    g = _POOL
    t = ref.MapGrid(data=np.asarray(g["tdata"]), xwidth=np.float32(1), orig=np.zeros(3, np.float32),
                    vec=np.asarray(g["tvec"]), ave=float(g["tave"]), std=1.0, std_norm_ave=1.0)
    spec = [np.asarray(x) for x in g["spec"]]
    done = 0
    for a in angles:
        if time.time() > deadline:
            break
        ref.rot_and_search_fft(np.asarray(g["sdata"]), np.asarray(g["svec"]), a, spec, t, "VecProduct", "ref32")
        done += 1
    return done


def cpu_baseline_pool(n1, n2, angles, budget_s: float, workers: int):
    """The same per-row oracle work on `workers` processes (one thread each),
    rows dealt round-robin; rows completed per second of wall time once the
    workers are up."""
    import multiprocessing as mp
    from oracle import vesper_ref as ref
    t = ref.MapGrid(data=np.asarray(n1.data, np.float64), xwidth=np.float32(n1.xwidth), orig=n1.orig,
                    vec=n1.vec, cent=n1.cent, xdim=n1.xdim, ave=n1.ave, std=n1.std,
                    std_norm_ave=n1.std_norm_ave)
    spec = np.stack(ref.target_spectra(ref.target_channels(t, "VecProduct")))
    tmp = tempfile.mkdtemp(prefix="vx_cpu_")
    arrays = {"tdata": np.asarray(n1.data, np.float64), "tvec": np.asarray(n1.vec, np.float32),
              "tave": np.asarray(n1.ave, np.float64), "spec": spec,
              "sdata": np.asarray(n2.data, np.float64), "svec": np.asarray(n2.vec, np.float32)}
    paths = {}
    for k, v in arrays.items():
        paths[k] = os.path.join(tmp, k + ".npy")
        np.save(paths[k], v)
    stride = max(1, len(angles) // (workers * 64))
    sample = angles[::stride]
    ctx = mp.get_context("spawn")
    with ctx.Pool(workers, initializer=_pool_init, initargs=(paths,)) as pool:
        pool.map(_pool_rows, [(sample[:0], 0.0)] * workers)        # workers up and mapped
        t0 = time.time()
        deadline = t0 + budget_s
        chunks = [(sample[i::workers], deadline) for i in range(workers)]
        done = sum(pool.map(_pool_rows, chunks))
        dt = time.time() - t0
    for v in paths.values():
        os.remove(v)
    os.rmdir(tmp)
    return dict(value=done / dt, unit="rotations/s", cores=workers, kind="port",
                sample=f"{done} rows (stride {stride} over {len(angles)}) of the same pair on {workers} "
                       f"processes x 1 thread, oracle ref32, {dt:.1f} s", **host_info())


def measure_prob(n1, n2, rot, batch, steps):
    """Prob-mode search speed on the same pair (7 channels: 3 vector + 4
    probability, two correlation groups, the alpha-mixed z-score fused into
    the search): rows/s of the mixing pass (the heavier of the two passes).
    The probability channels are synthetic: four random fractions that sum to
    one on the voxels where each map's density is non-zero."""
    import torch
    from vesper_python_amd import _native as nat
    from vesper_python_amd import engine

    def probs(m, seed):
        g = torch.Generator(device="cuda").manual_seed(seed)
        d = m._dev[0]
        u = torch.rand((4,) + tuple(d.shape), generator=g, device="cuda", dtype=torch.float32) + 1e-3
        u = u / u.sum(dim=0, keepdim=True)
        return (u * (d != 0).float()).contiguous()

    n = int(n1.xdim)
    plan = engine.SearchPlan(n, nat.VX_MODE_PROB, batch=batch)
    plan.set_target(*n1._dev, probs(n1, 1), 0.0)
    plan.set_search(*n2._dev, probs(n2, 2), 0.0)
    plan.set_prob_stats(0.5, 1.0, 0.0, 1.0, 0.0, enable=True)
    from vesper_python_amd import autotune
    R = int(rot.shape[0])
    tuned = autotune.tune(plan, rot[::max(1, R // 256)].contiguous(), sample_rows=128, use_cache=False)
    plan.score_rows(rot[: 2 * batch])
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(steps):
        plan.score_rows(rot)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / steps
    nprof = 8 * batch
    stage = plan.profile_rows(rot[:nprof])
    return {"rows_per_s": round(R / (ms / 1e3), 1), "ms_per_pass": round(ms, 3), "rows": R, "batch": batch,
            "channels": 7, "pass": "mixing pass (vector, probability and mixed scores per row)",
            "schedule": tuned["schedule"], "stage_ms_per_launch": {k: round(v / 8, 4) for k, v in stage.items()}}


def _mfma_roofline(variant, n, info, batch, xcorr_ms):
    """MFMA rate of an MFMA x pass from its per-launch time (ms), or None
    for the FFT x passes."""
    if variant == 2:
        # split-bf16 whole-axis DFT: 3 products per real-form complex GEMM,
        # forward [2NP x 2KP] x [2KP x plane] per channel, inverse [2NP x 2NP]
        npad = 16 * ((n + 15) // 16)
        kp = 16 * ((info["box_width"] + 15) // 16)
        plane = n * (n // 2 + 1)
        flops = 2 * 3 * batch * plane * (3 * (2 * npad) * (2 * kp) + (2 * npad) ** 2)
        name = "xcorr"
    elif variant == 4:
        # radix split: per 16-column tile and rotation, 4 waves x 3 split
        # products x N2 column blocks x (channels x forward K-steps + inverse
        # K-steps) v_mfma_f32_16x16x32_f16 of 16,384 flops each
        from vesper_python_amd.autotune import mr_factor
        f1, f2 = mr_factor(n)
        lo_, w_ = info["box_lo"], info["box_width"]
        nwin = (lo_ + w_ - 1) // f2 - lo_ // f2 + 1
        ksf, ksi = (1 if nwin <= 16 else 2), (1 if f1 <= 16 else 2)
        tiles = -(-(n * (n // 2 + 1)) // 16)
        flops = batch * tiles * 4 * 3 * f2 * (3 * ksf + ksi) * 16384
        name = "xcorr (radix split, N = %d x %d)" % (f1, f2)
    else:
        return None
    tflops = flops / (xcorr_ms / 1e3) / 1e12
    return {"kernel": name, "flops_per_launch": flops, "achieved": round(tflops, 1), "peak": PEAK_BF16_TFLOPS,
            "unit": "TFLOP/s", "frac": round(tflops / PEAK_BF16_TFLOPS, 4)}


def measure_mixed(plan, rot, best, trans, n, batch, steps, plain_value, profile_rows, workload):
    """The same exhaustive search with the radix-split MFMA x pass (the
    length-N1 DFTs of N = N1 x N2 as split-f16 matrix products with fp32
    accumulation, everything else fp32): rows/s over `steps` full passes
    (HIP events), its scores against the fp32 run's (`best`, `trans`), and
    the x pass's MFMA and HBM rates.  Leaves the plan on this schedule; None
    where the grid edge has no radix split."""
    import torch
    from vesper_python_amd import engine
    try:
        plan.tune(xpass_variant=4, xpass_tile=0)
    except engine.VesperError:
        return None
    R = int(rot.shape[0])
    b2 = torch.empty_like(best)
    t2 = torch.empty_like(trans)
    plan.score_rows(rot, b2, t2)
    engine.topn(b2, 10)
    torch.cuda.synchronize()
    st = torch.cuda.current_stream()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record(st)
    for _ in range(steps):
        plan.score_rows(rot, b2, t2)
        top = engine.topn(b2, 10)
    e1.record(st)
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / steps
    value = R / (ms / 1e3)
    ref = best[:, 0].double()
    rel = float(((b2[:, 0].double() - ref).abs().max() / ref.abs().max()).item())
    teq = float((t2 == trans).all(dim=-1).double().mean().item())
    top32 = engine.topn(best[:, 0].contiguous(), 10)
    prows = min(R, max(batch, profile_rows // batch * batch))
    stage_ms = plan.profile_rows(rot[:prows])
    launches = (prows + batch - 1) // batch
    info = plan.info()
    xms = stage_ms["xcorr"] / launches
    xbytes = _stage_bytes(n, info["box_width"], batch)["xcorr"]
    return {"value": round(value, 2), "unit": "rotations/s", "ms_per_step": round(ms, 3),
            "vs_plain_fp32": round(value / plain_value, 4),
            "schedule": {"xpass_variant": 4},
            "max_rel_score_diff_vs_fp32": rel, "translations_equal_vs_fp32": teq,
            "top10_same_as_fp32": bool(np.array_equal(np.asarray(top), np.asarray(top32))),
            "xcorr_ms_per_launch": round(xms, 4),
            "xcorr_hbm": {"achieved": round(xbytes / (xms / 1e3) / 1e9, 1), "peak": PEAK_HBM_GBS, "unit": "GB/s",
                          "frac": round(xbytes / (xms / 1e3) / 1e9 / PEAK_HBM_GBS, 4)},
            "mfma": _mfma_roofline(4, n, info, batch, xms),
            "traffic": pmc_traffic(workload, batch, {"xpass_variant": 4, "xpass_tile": 0}, info["box_width"])}


def measure_rescore(plan, n1, n2, angles, best):
    """The FP64 re-check on the benchmark pair: the rows near the top-10
    cutoff of the exhaustive float32 scores and the 3 x 3 x 3 refinement rows
    of the top 10 (at most 270) re-scored by vx_rescore_f64; GPU time (HIP
    events), candidates per row, the largest FP32/FP64 relative difference,
    and how many of the top 10 change rank or translation."""
    import torch
    from vesper_python_amd import engine
    from vesper_python_amd import VESPER_1_prob_new as V
    from vesper_python_amd.geometry import refine_rows, rotation_rows
    b32 = best[:, 0].double().cpu().numpy()
    top = engine.topn(best[:, 0].contiguous(), 10)
    near = V.recheck_rows(b32, top)
    ref_rows = np.concatenate([refine_rows(angles[i]) for i in top], axis=0)
    rows = np.concatenate([angles[near], ref_rows], axis=0)
    rot = torch.from_numpy(rotation_rows(rows)).cuda()
    target, search = n1._dev, n2._dev
    plan.rescore_f64(rot[:8], target, search, n1.ave)             # warm-up (workspace, kernels)
    b, t = plan.score_rows(rot)
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record()
    sc, tr, nc = plan.rescore_f64(rot, target, search, n1.ave)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1)
    s64 = sc[:, 0].cpu().numpy()
    s32 = b[:, 0].double().cpu().numpy()
    t64, t32 = tr[:, 0].cpu().numpy(), t[:, 0].cpu().numpy()
    k = len(near)
    order64 = sorted(range(k), key=lambda i: (-s64[i], int(near[i])))[:10]
    return {"rows": int(len(rows)), "near_cutoff_rows": int(k), "refinement_rows": int(len(ref_rows)),
            "gpu_ms": round(ms, 3), "ms_per_row": round(ms / len(rows), 4),
            "mean_candidates": round(float(nc.double().mean().item()), 2), "max_candidates": int(nc.max().item()),
            "max_rel_fp32_vs_fp64": float(np.max(np.abs(s32 - s64) / np.abs(s64))),
            "translations_changed": int(np.sum(np.any(t32 != t64, axis=1))),
            "top10_rank_changes": int(sum(int(near[a]) != int(b_) for a, b_ in zip(order64, top)))}


def _event_ms(fn, reps):
    """Mean milliseconds of fn() between HIP events on the current stream,
    after one untimed call."""
    import torch
    fn()
    torch.cuda.synchronize()
    st = torch.cuda.current_stream()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record(st)
    for _ in range(reps):
        fn()
    e1.record(st)
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / reps


def measure_aux(w, n1, n2, best, reps=10):
    """HBM rooflines of the kernels beside the search on the benchmark pair,
    each call timed with HIP events on its launch stream (mean of `reps`
    calls after a warm-up), against its algorithmic bytes:
      fastVEC of the query map (k_fv_z reads the source map and writes the z
        sums, 2 float64 per (x, y, dest z); k_fv_y reads them and writes 3
        float64 per (x, dest y, dest z); k_fv_x reads those, one source tap
        per voxel, and writes data float64 + vec 3 x float32),
      get_score at one translation (k_stats1 + k_stats2 read the query
        density twice; k_score reads both maps' density float64 and vectors
        3 x float32),
      the top-10 selection over every row's score (vx_topn_ws: the scores
        once; a few launches of sorting in LDS, latency-bound).
    The per-kernel bytes divide the per-kernel durations of a kernel trace."""
    import torch
    from vesper_python_amd import engine
    m1, m2 = w["src_maps"]
    n = int(n2.xdim)
    n3 = n ** 3
    src = getattr(m2, "_dev_data", None)
    if src is None:
        src = engine.to_dev(np.asarray(m2.data, dtype=np.float32), torch.float32)
    src = src.float().contiguous()
    sx, sy, sz = (int(v) for v in src.shape)
    fv_bytes = {"k_fv_z": sx * sy * sz * 4 + sx * sy * n * 16,
                "k_fv_y": sx * sy * n * 16 + 3 * sx * n * n * 8,
                "k_fv_x": 3 * sx * n * n * 8 + n3 * 4 + n3 * (8 + 12)}
    fv_ms = _event_ms(lambda: engine.fastvec(src, m2.xwidth, m2.orig, n, n2.xwidth, n2.orig, w["g"]), reps)
    tdata, tvec = n1._dev
    sdata, svec = n2._dev
    gs_bytes = {"k_stats1+k_stats2": 2 * n3 * 8, "k_score": n3 * (8 + 12 + 8 + 12)}
    gs_ms = _event_ms(lambda: engine.get_score(tdata, tvec, n1.ave, sdata, svec, (3, -2, 5)), reps)
    R = int(best.shape[0])
    scores = best[:, 0].contiguous()
    ws = engine.topn_workspace(R, 10)
    rows = torch.empty(10, dtype=torch.int32, device=scores.device)
    tn_ms = _event_ms(lambda: engine.topn_into(scores, R, 10, rows, ws), reps)
    tn_bytes = R * 4 + 10 * 4

    def line(ms, by, bound="hbm"):
        gbs = sum(by.values()) / (ms / 1e3) / 1e9
        return {"ms": round(ms, 4), "bytes": int(sum(by.values())), "by_kernel_bytes": by, "achieved_GBps": round(gbs, 1),
                "frac": round(gbs / PEAK_HBM_GBS, 4), "bound": bound}
    return {"fastvec": dict(line(fv_ms, fv_bytes), source_shape=[sx, sy, sz], n=n),
            "get_score": line(gs_ms, gs_bytes),
            "topn": dict(line(tn_ms, {"vx_topn_ws": tn_bytes}, "latency"), rows=R)}


def measure_library_fft(n1, n2, rot, batch, plan):
    """The same correlation done with the vendor FFT library (rocFFT through
    torch.fft) on already-rotated query channels: forward r2c of the three
    channels of `batch` rows, multiply by the conjugated target spectra,
    channel sum, c2r, per-row maximum.  A reference point for the fused
    kernels; it leaves out the rotation and the support-box pruning, so it
    is a lower bound on a library-FFT path's time."""
    import torch
    from vesper_python_amd import engine
    n = int(n1.xdim)
    dims = (-3, -2, -1)
    q = torch.empty((batch, 3, n, n, n), device="cuda", dtype=torch.float32)
    for b in range(batch):
        _, ov, _ = engine.rotate_map(rot[b].cpu().numpy(), n2._dev[0], n2._dev[1])
        q[b] = ov.permute(3, 0, 1, 2)
    tv = n1._dev[1].permute(3, 0, 1, 2).contiguous()
    tspec = torch.fft.rfftn(tv, dim=dims).conj()

    def run():
        s = (torch.fft.rfftn(q, dim=dims) * tspec).sum(dim=1)
        # unnormalised inverse, as the reference's c2r
        return torch.fft.irfftn(s, s=(n, n, n), dim=dims, norm="forward").reshape(batch, -1).amax(dim=1)

    run()
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(3):
        best = run()
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / 3
    # the fused path's scores for the same rows
    fb, _ = plan.score_rows(rot[:batch])
    rel = float(((best - fb[:, 0]).abs() / fb[:, 0].abs()).max())
    return {"rows_per_s": round(batch / (ms / 1e3), 1), "ms_per_batch": round(ms, 3), "batch": batch,
            "max_rel_diff_vs_fused": rel,
            "what": "torch.fft (rocFFT) r2c x 3 channels, multiply, sum, c2r, max; rotation excluded"}


def measure_uneven_pairs(npairs: int = 6, ang: float = 30.0, s: float = 1.5, g: float = 16.0,
                         seed0: int = 4000):
    """Many pairs of uneven size in one call of `batch.search_pairs`: pairs
    from the uneven generator (A edge U{96..320} voxels at 1 A, 40-60 %
    sub-chains), resampled at -s 1.5 and fastVEC'd, then every -A `ang` row of
    every pair scored with one plan and one captured graph per grid edge,
    replayed per pair.  Rows/s of the whole call (plan setup and capture
    included), after one untimed call; the eager call (no graphs) beside it,
    and whether both give the same top rows."""
    import torch
    from vesper_python_amd import batch, synth
    from vesper_python_amd import VESPER_1_prob_new as V
    from vesper_python_amd import main as cli
    from vesper_python_amd.geometry import calc_angle_comb
    tmp = tempfile.mkdtemp(prefix="vx_pairs_")
    pairs = []
    for k in range(npairs):
        pair = synth.make_uneven_pair(seed0 + k)
        paths = synth.write_pair(pair, os.path.join(tmp, f"p{k}"))
        with contextlib.redirect_stdout(io.StringIO()):
            m1, m2, n1, n2 = cli._resample_pair(paths["A"], paths["B"], pair.a.threshold, pair.b.threshold, s, g)
            n1 = V.fastVEC(m1, n1, g)
            n2 = V.fastVEC(m2, n2, g)
        pairs.append((n1, n2))
    rows = len(calc_angle_comb(ang)) * npairs
    res = {}
    for use_graph in (True, False):
        batch.search_pairs(pairs, ang=ang, use_graph=use_graph)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        out = batch.search_pairs(pairs, ang=ang, use_graph=use_graph)
        torch.cuda.synchronize()
        res[use_graph] = (time.perf_counter() - t0, out)
    same = all(np.array_equal(a["top"], b["top"]) for a, b in zip(res[True][1], res[False][1]))
    edges = [int(p[0].xdim) for p in pairs]
    return {"pairs": npairs, "grid_edges": edges, "graphed_pairs": sum(edges.count(e) > 1 for e in edges),
            "rows": rows,
            "rows_per_s_graph": round(rows / res[True][0], 1), "rows_per_s_eager": round(rows / res[False][0], 1),
            "same_top_rows": bool(same),
            "what": f"uneven pairs (A edge U{{96..320}}), -s {s} -g {g} -A {ang}; one plan per grid edge, a "
                    "captured graph replayed per pair where pairs share an edge, plan setup and capture inside "
                    "the timed call"}


def _write_uneven_pair(args):
    """Worker: one pair of the uneven generator written as MRC files."""
    seed, tmp = args
    from vesper_python_amd import synth
    pair = synth.make_uneven_pair(seed)
    paths = synth.write_pair(pair, os.path.join(tmp, f"p{seed}"))
    return seed, paths, float(pair.a.threshold), float(pair.b.threshold)


def measure_pairs48(npairs=48, ang=10.0, s=1.5, g=16.0, seed0=4000, workers=16):
    """The uneven multi-pair workload at full size: 48 pairs from the uneven
    generator (A edge U{96..320} at 1 A, 1000-12000 atoms, 40-60 % sub-chains),
    resampled at -s 1.5 -g 16 and searched exhaustively at -A 10 by one
    batch.search_pairs call (one plan and captured graph per grid edge,
    replayed per pair; plan setup and capture inside the timed call).  The
    pairs are generated on `workers` host processes, then read, resampled
    and fastVEC'd by the drop-in path.  Reports the wall time of the search
    call, rows/s, and the grid edges."""
    import multiprocessing as mp
    import torch
    from concurrent.futures import ProcessPoolExecutor
    from vesper_python_amd import batch
    from vesper_python_amd import VESPER_1_prob_new as V
    from vesper_python_amd import main as cli
    from vesper_python_amd.geometry import calc_angle_comb
    tmp = tempfile.mkdtemp(prefix="vx_p48_")
    t0 = time.perf_counter()
    with ProcessPoolExecutor(max_workers=workers, mp_context=mp.get_context("spawn")) as ex:
        made = list(ex.map(_write_uneven_pair, [(seed0 + k, tmp) for k in range(npairs)]))
    t_gen = time.perf_counter() - t0
    pairs = []
    t0 = time.perf_counter()
    for k, (seed, paths, t1, t2) in enumerate(made):
        with contextlib.redirect_stdout(io.StringIO()):
            m1, m2, n1, n2 = cli._resample_pair(paths["A"], paths["B"], t1, t2, s, g)
            n1 = V.fastVEC(m1, n1, g)
            n2 = V.fastVEC(m2, n2, g)
        for m in (n1, n2):
            # the search reads the device copies; keep only the host shapes
            m.data = np.broadcast_to(np.float64(0.0), np.shape(m.data))
            m.vec = np.broadcast_to(np.float32(0.0), np.shape(m.vec))
        pairs.append((n1, n2))
        for f in paths.values():
            os.remove(f)
        print(f"[pairs48] pair {k + 1}/{npairs} N={int(n1.xdim)}", file=sys.stderr, flush=True)
    t_prep = time.perf_counter() - t0
    rows = len(calc_angle_comb(ang)) * npairs
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    out = batch.search_pairs(pairs, ang=ang, use_graph=True,
                             progress=lambda i, n: print(f"[pairs48] searched pair {i + 1} N={n}", file=sys.stderr,
                                                         flush=True))
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    edges = [int(p[0].xdim) for p in pairs]
    return {"pairs": npairs, "rows": rows, "search_seconds": round(dt, 2), "rows_per_s": round(rows / dt, 1),
            "grid_edges": edges, "distinct_edges": len(set(edges)),
            "graphed_pairs": sum(edges.count(e) > 1 for e in edges),
            "generate_seconds": round(t_gen, 1), "resample_fastvec_seconds": round(t_prep, 1),
            "top1_scores": [float(o["scores"][o["top"][0]]) for o in out],
            "what": f"48 uneven pairs, -s {s} -g {g} -A {ang}, exhaustive; one plan + graph per grid edge"}


def measure_cli(w):
    """The drop-in CLI end to end on the workload's MRC files, as a user runs
    it: `main orig -a A.mrc -b B.mrc -t -T -g -s -A -N 10` (MRC reading,
    resampling, fastVEC, the exhaustive search, local refinement, FP64
    get_score of the models and the reference's printed output)."""
    from vesper_python_amd import main as cli
    t1, t2 = w["thresholds"]
    argv = ["orig", "-a", w["paths"]["A"], "-b", w["paths"]["B"], "-t", str(t1), "-T", str(t2),
            "-g", str(w["g"]), "-s", str(w["s"]), "-A", str(w["A"]), "-N", "10"]
    buf = io.StringIO()
    t0 = time.perf_counter()
    with contextlib.redirect_stdout(buf):
        res = cli.main(argv)
    dt = time.perf_counter() - t0
    return {"seconds": round(dt, 3), "models": len(res), "printed_lines": buf.getvalue().count("\n"),
            "argv": " ".join(["orig", "-a", "A.mrc", "-b", "B.mrc"] + argv[5:])}


def measure_cli_prob(w):
    """The drop-in CLI's prob mode end to end on the same pair with four
    synthetic probability channels per map (Dirichlet(0.5) fractions on the
    density mask, smoothed, renormalised; written as the .npy files the CLI
    reads): `main prob -a A.mrc -npa A.npy -b B.mrc -npb B.npy -t -T -g -s
    -A -N 10 -alpha 0.5` -- ten resampled and fastVEC'd maps, the statistics
    pass and the mixing pass over every row, refinement, scoring, output."""
    import numpy as np
    from vesper_python_amd import main as cli
    from vesper_python_amd import synth
    pair = synth.make_config_pair(w["pair"], w["seed"])
    rng = np.random.default_rng(w["seed"] + 1)
    synth.add_prob_channels(rng, pair.a)
    synth.add_prob_channels(rng, pair.b)
    tmp = tempfile.mkdtemp(prefix="vx_prob_")
    paths = synth.write_pair(pair, os.path.join(tmp, "p"))
    argv = ["prob", "-a", paths["A"], "-npa", paths["A_prob"], "-b", paths["B"], "-npb", paths["B_prob"],
            "-t", str(pair.a.threshold), "-T", str(pair.b.threshold), "-g", str(w["g"]), "-s", str(w["s"]),
            "-A", str(w["A"]), "-N", "10", "-alpha", "0.5"]
    buf = io.StringIO()
    t0 = time.perf_counter()
    with contextlib.redirect_stdout(buf):
        res = cli.main(argv)
    dt = time.perf_counter() - t0
    return {"seconds": round(dt, 3), "models": len(res), "printed_lines": buf.getvalue().count("\n"),
            "argv": "prob -a A.mrc -npa A.npy -b B.mrc -npb B.npy " + " ".join(argv[9:])}


def measure_interp(edge: int = 337, cpu: bool = True):
    """Prediction-grid interpolation (interpolation_functions.py:6-135) on a
    synthetic [edge]^3 x 4 float32 grid with values on one parity lattice,
    the size of the reference notebook's timing: GPU time of vx_interpolate
    (HIP events, grid resident in HBM), and the oracle's vectorised NumPy
    form on the host for comparison."""
    import torch
    from vesper_python_amd import interpolation as ip
    g = torch.Generator(device="cuda").manual_seed(5)
    d = torch.zeros((edge, edge, edge, 4), dtype=torch.float32, device="cuda")
    d[::2, ::2, ::2] = torch.rand(d[::2, ::2, ::2].shape, generator=g, device="cuda") + 0.01
    work = d.clone()
    ip.interpolate_device(work)                     # warm-up
    torch.cuda.synchronize()
    reps = 5
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    ms = 0.0
    for _ in range(reps):
        work.copy_(d)
        e0.record()
        ip.interpolate_device(work)
        e1.record()
        torch.cuda.synchronize()
        ms += e0.elapsed_time(e1)
    ms /= reps
    grid_bytes = d.numel() * 4
    model = grid_bytes + grid_bytes * 7 // 8        # read the grid, write the seven filled lattices
    rec = {"grid": [edge, edge, edge, 4], "dtype": "f32", "gpu_ms": round(ms, 3),
           "achieved_GBps": round(model / (ms / 1e3) / 1e9, 1), "bytes_model": model,
           "peak_GBps": PEAK_HBM_GBS, "frac": round(model / (ms / 1e3) / 1e9 / PEAK_HBM_GBS, 4)}
    if cpu:
        from threadpoolctl import threadpool_limits
        from oracle import interp_ref as ir
        host = d.cpu().numpy()
        with threadpool_limits(1):
            t0 = time.perf_counter()
            ir.interpolate(host)
            rec["cpu_port_s"] = round(time.perf_counter() - t0, 3)
        rec["cpu_port"] = "oracle vectorised NumPy, 1 thread, same grid"
        rec["gpu_matches_port"] = bool(np.array_equal(host, work.cpu().numpy()))
    return rec


def measure_proposal(plan, n1, n2, ang, best, R, plain_rate):
    """Learned proposal on the same pair: train on the -A 30 landscape, then
    FFT-score 5 % of the -A `ang` rows in two rounds with retraining; recall
    of the exhaustive top-10 (`best` holds the exhaustive scores) and
    rotation-equivalents per second (R rows covered / wall time of the whole
    proposal pipeline, training included, after one untimed warm-up run).
    Three timed runs, each with a fresh network: the median time is reported,
    and the lowest recall of the three."""
    import torch
    from vesper_python_amd import engine, proposer
    from vesper_python_amd.geometry import calc_angle_comb
    truth = engine.topn(best[:, 0].contiguous(), 10)
    fine = calc_angle_comb(ang)
    coarse = calc_angle_comb(30.0)
    # one untimed run first, as the exhaustive search gets its warm-up step:
    # the first call loads the network's kernels and grows the allocator
    pr = proposer.Proposer(plan, *n1._dev, *n2._dev, proposer.RotationMLP(seed=7))
    proposer.propose_rounds(pr, coarse, fine, frac=0.05, rounds=2)
    torch.cuda.synchronize()
    runs = []
    for _ in range(3):
        t0 = time.perf_counter()
        pr = proposer.Proposer(plan, *n1._dev, *n2._dev, proposer.RotationMLP(seed=7))
        top, scored, ncoarse = proposer.propose_rounds(pr, coarse, fine, frac=0.05, rounds=2)
        torch.cuda.synchronize()
        runs.append((time.perf_counter() - t0, proposer.recall(top, truth), scored + ncoarse))
    dt = float(np.median([r[0] for r in runs]))
    rate = R / dt
    return {"recall_top10": min(r[1] for r in runs), "recall_runs": [r[1] for r in runs], "fine_rows": R,
            "fft_rows_scored": int(np.median([r[2] for r in runs])),
            "seconds": round(dt, 3), "seconds_runs": [round(r[0], 3) for r in runs],
            "rotation_equivalents_per_s": round(rate, 1),
            "speedup_vs_plain": round(rate / plain_rate, 2),
            "method": "MLP (80-256-256-256-1, bf16/MFMA) trained on the -A 30 landscape, 2 rounds x 2.5 % of rows, "
                      "then neighbour sweeps around the best rows"}


def _pairs_in_chunks(seeds, tmp, workers, chunk=32):
    """Uneven-generator pairs written as MRC files by a pool of host
    processes, yielded one at a time (seed, paths, thresholds); the files of
    at most two chunks of pairs exist at once, and each pair's are removed
    after use."""
    import multiprocessing as mp
    from concurrent.futures import ProcessPoolExecutor
    chunks = [seeds[c0:c0 + chunk] for c0 in range(0, len(seeds), chunk)]
    with ProcessPoolExecutor(max_workers=workers, mp_context=mp.get_context("spawn")) as ex:
        fut = [ex.submit(_write_uneven_pair, (sd, tmp)) for sd in chunks[0]] if chunks else []
        for ci in range(len(chunks)):
            # the next chunk is written while this one is used
            nxt = [ex.submit(_write_uneven_pair, (sd, tmp)) for sd in chunks[ci + 1]] if ci + 1 < len(chunks) else []
            for f in fut:
                item = f.result()
                yield item
                for path in item[1].values():
                    if os.path.exists(path):
                        os.remove(path)
            fut = nxt


def _resampled(item, s, g):
    from vesper_python_amd import VESPER_1_prob_new as V
    from vesper_python_amd import main as cli
    seed, paths, t1, t2 = item
    with contextlib.redirect_stdout(io.StringIO()):
        m1, m2, n1, n2 = cli._resample_pair(paths["A"], paths["B"], t1, t2, s, g)
        n1 = V.fastVEC(m1, n1, g)
        n2 = V.fastVEC(m2, n2, g)
    return n1, n2


def measure_ablation(n_train=16, n_test=8, s=3.0, g=16.0, label_ang=10.0, judge_ang=5.0, coarse_ang=30.0,
                     fracs=(0.0025, 0.005, 0.01, 0.05), seed_train=6000, seed_test=7000, epochs=8,
                     workers=16, weights_in=None, weights_out=None):
    """What the learned proposer gains over the same search without it (the
    learned-proposal protocol; its full size is 512 training and 64
    held-out pairs): the network is pretrained on the exhaustive -A
    label_ang landscapes of n_train pairs from the uneven generator
    (streamed: one pair's plan on the device at a time), then on each of
    n_test held-out pairs four policies spend the same budget -- the -A
    coarse_ang landscape, a fraction (each of `fracs`) of the -A judge_ang
    rows in two shares, three neighbour sweeps around the best rows --
    choosing the shares by (mlp) the network refined on the pair's coarse
    landscape, (coarse) the coarse landscape's score of each row's nearest
    coarse row, (random) uniform draws, (hybrid) half of each share by the
    coarse ranking and the rest by the network.  Judge: the exhaustive -A
    judge_ang search of the pair (top-10 recall, and the gap between the
    best score found and the exhaustive best).  Maps resampled at -s s (2x
    the uneven workload's 1.5 A, so that the protocol fits one GPU call).  weights_out saves the pretrained network (.npz) and
    weights_in loads one instead of pretraining, so that training and the
    held-out evaluation can run separately."""
    import torch
    from vesper_python_amd import _native as nat
    from vesper_python_amd import engine, proposer
    from vesper_python_amd.batch import auto_batch
    from vesper_python_amd.geometry import calc_angle_comb, rotation_rows
    tmp = tempfile.mkdtemp(prefix="vx_abl_")
    labels = calc_angle_comb(label_ang)
    fine = calc_angle_comb(judge_ang)
    coarse = calc_angle_comb(coarse_ang)

    def prop(n1, n2, mlp):
        plan = engine.SearchPlan(int(n1.xdim), nat.VX_MODE_VEC, batch=auto_batch(int(n1.xdim)))
        plan.set_target(*n1._dev, None, n1.ave)
        plan.set_search(*n2._dev, None, n1.ave)
        return proposer.Proposer(plan, *n1._dev, *n2._dev, mlp)

    out = {}
    t0 = time.perf_counter()
    if weights_in:
        with np.load(weights_in) as z:
            st = {k: z[k] for k in z.files}
        out["pretrained_from"] = os.path.basename(weights_in)
    else:
        it = _pairs_in_chunks([seed_train + k for k in range(n_train)], tmp, workers)
        edges = []

        def make(k):
            return prop(*_resampled(next(it), s, g), None)

        def progress(k, n):
            edges.append(int(n))
            if (k + 1) % 8 == 0 or k + 1 == n_train:
                print(f"[ablation] train pair {k + 1}/{n_train} N={n}", file=sys.stderr, flush=True)
        mlp = proposer.RotationMLP(seed=21)
        losses = proposer.pretrain_streamed(mlp, make, n_train, labels, epochs=epochs, lr=2e-3, progress=progress)
        st = mlp.state()
        if weights_out:
            np.savez(weights_out, **st)
        out.update(pretrain_loss=[round(losses[0], 4), round(losses[-1], 4)],
                   pretrain_s=round(time.perf_counter() - t0, 1), train_rows=len(labels) * n_train,
                   train_grid_edges=[min(edges), max(edges)])
        print(f"[ablation] pretrained on {n_train} pairs in {out['pretrain_s']} s, loss {out['pretrain_loss']}",
              file=sys.stderr, flush=True)
    per = []
    for k, item in enumerate(_pairs_in_chunks([seed_test + k for k in range(n_test)], tmp, workers)):
        n1, n2 = _resampled(item, s, g)
        pr = prop(n1, n2, proposer.RotationMLP(seed=22))
        rot = torch.from_numpy(rotation_rows(fine)).cuda()
        torch.cuda.synchronize()
        te = time.perf_counter()
        best, _ = pr.plan.score_rows(rot)
        truth = engine.topn(best[:, 0].contiguous(), 10)
        torch.cuda.synchronize()
        t_ex = time.perf_counter() - te
        b = best[:, 0].double().cpu().numpy()
        rec = {"seed": item[0], "grid_edge": int(n1.xdim), "exhaustive_s": round(t_ex, 3)}
        for frac in fracs:
            for pol in proposer.POLICIES:
                pr.mlp.load(st)
                torch.cuda.synchronize()
                tp = time.perf_counter()
                top, nscored, ncoarse = proposer.propose_rounds(pr, coarse, fine, frac=frac, rounds=2, policy=pol,
                                                                coarse_step=coarse_ang, seed=k)
                torch.cuda.synchronize()
                rec[f"{pol}@{frac}"] = {"recall_top10": proposer.recall(top, truth),
                                        "best_gap": round(float(1.0 - b[top].max() / b.max()), 6),
                                        "fft_rows": nscored + ncoarse,
                                        "seconds": round(time.perf_counter() - tp, 3)}
        print(f"[ablation] held-out pair {k + 1}/{n_test} N={rec['grid_edge']} "
              + " ".join(f"{p}@{f}={rec[f'{p}@{f}']['recall_top10']}" for f in fracs for p in proposer.POLICIES),
              file=sys.stderr, flush=True)
        per.append(rec)
        pr.plan.close()
    table = {}
    for frac in (fracs if per else ()):
        row = {}
        for p in proposer.POLICIES:
            recs = [r[f"{p}@{frac}"] for r in per]
            row[p] = {"mean_recall_top10": round(float(np.mean([x["recall_top10"] for x in recs])), 4),
                      "pairs_full_recall": int(sum(x["recall_top10"] == 1.0 for x in recs)),
                      "mean_best_gap": round(float(np.mean([x["best_gap"] for x in recs])), 6),
                      "mean_fft_rows": int(np.mean([x["fft_rows"] for x in recs]))}
        row["gain_over_coarse"] = round(row["mlp"]["mean_recall_top10"] - row["coarse"]["mean_recall_top10"], 4)
        row["gain_over_random"] = round(row["mlp"]["mean_recall_top10"] - row["random"]["mean_recall_top10"], 4)
        table[str(frac)] = row
    out.update(protocol=f"pretrain on {n_train} uneven pairs (exhaustive -A {label_ang} labels), judge {n_test} "
                        f"held-out pairs by exhaustive -A {judge_ang} ({len(fine)} rows); budget per policy: "
                        f"-A {coarse_ang} landscape + the fraction of the rows in 2 shares + 3 neighbour sweeps; "
                        f"-s {s} -g {g}",
               by_fraction=table, pairs=per)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="pair200", choices=sorted(WORKLOADS))
    ap.add_argument("--batch", type=int, default=32, help="rotation rows per launch")
    ap.add_argument("--cpu-seconds", type=float, default=20.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--profile-rows", type=int, default=512)
    ap.add_argument("--no-tune", action="store_true")
    ap.add_argument("--rows", type=int, default=0, help="score only the first ROWS rows per step (profiling)")
    ap.add_argument("--calib-copy-mb", type=int, default=0,
                    help="first run a fill and a scale of this many MiB (counter calibration)")
    ap.add_argument("--no-proposal", action="store_true", help="skip the learned-proposal measurement")
    ap.add_argument("--no-prob", action="store_true", help="skip the prob-mode measurement")
    ap.add_argument("--no-interp", action="store_true", help="skip the prediction-grid interpolation measurement")
    ap.add_argument("--no-libfft", action="store_true", help="skip the vendor-FFT pipeline reference point")
    ap.add_argument("--no-pairs", action="store_true", help="skip the uneven multi-pair measurement")
    ap.add_argument("--no-cli", action="store_true", help="skip the end-to-end CLI measurement")
    ap.add_argument("--pairs48", action="store_true",
                    help="only the 48-pair uneven workload at -s 1.5 -A 10; prints its own JSON line")
    ap.add_argument("--ablation", action="store_true",
                    help="only the learned-proposer ablation (held-out pairs, -A 5 judge); prints its own JSON line")
    ap.add_argument("--ablation-pairs", default="16,8",
                    help="training and held-out pairs of the ablation (the protocol's full size: 512,64)")
    ap.add_argument("--ablation-fracs", default="0.0025,0.005,0.01,0.05",
                    help="shares of the -A 5 rows each policy may FFT-score")
    ap.add_argument("--ablation-epochs", type=int, default=8)
    ap.add_argument("--ablation-test-seed", type=int, default=7000,
                    help="seed of the first held-out pair (the training pairs start at 6000)")
    ap.add_argument("--ablation-weights-out", default="", help="save the pretrained network (.npz)")
    ap.add_argument("--ablation-weights-in", default="", help="load a pretrained network instead of training")
    ap.add_argument("--aux", action="store_true",
                    help="only the rooflines of fastVEC, get_score and top-N on the workload; prints its own JSON line")
    ap.add_argument("--schedule", default="", help='fixed kernel schedule instead of autotuning, as JSON, '
                    'e.g. \'{"xpass_variant": 2, "xpass_tile": 64}\' (profiling a specific kernel)')
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1",
                          "n_gpus": args.gpus}), flush=True)
        sys.exit(2)

    import torch
    from vesper_python_amd import engine
    from vesper_python_amd import _native as nat
    from vesper_python_amd.geometry import calc_angle_comb, rotation_rows

    engine.require_gpu()
    if args.ablation:
        ntr, nte = (int(v) for v in args.ablation_pairs.split(","))
        print(json.dumps({"ablation": measure_ablation(
            n_train=ntr, n_test=nte, fracs=tuple(float(v) for v in args.ablation_fracs.split(",")),
            epochs=args.ablation_epochs, seed_test=args.ablation_test_seed, weights_in=args.ablation_weights_in or None,
            weights_out=args.ablation_weights_out or None)}), flush=True)
        return
    if args.pairs48:
        print(json.dumps({"pairs48": measure_pairs48()}), flush=True)
        return
    if args.calib_copy_mb:
        # counter calibration: a fill (writes S bytes) and an elementwise
        # scale (reads S, writes S) of S = calib_copy_mb MiB
        src = torch.ones(args.calib_copy_mb * 262144, dtype=torch.float32, device="cuda")
        dst = src * 2.0
        torch.cuda.synchronize()
        del src, dst
    w, n1, n2 = prepare(args.workload)
    n = int(n1.xdim)
    angles = calc_angle_comb(w["A"])
    if args.rows:
        angles = angles[:args.rows]
    R = len(angles)
    dev = torch.device("cuda", 0)
    if args.aux:
        # top-N timing needs one score per row, not the search's values
        scores = torch.rand((R, 1), generator=torch.Generator(device=dev).manual_seed(0), device=dev)
        print(json.dumps({"aux_rooflines": measure_aux(w, n1, n2, scores)}), flush=True)
        return
    tdata, tvec = n1._dev
    sdata, svec = n2._dev
    plan = engine.SearchPlan(n, nat.VX_MODE_VEC, batch=args.batch)
    plan.set_target(tdata, tvec, None, n1.ave)
    plan.set_search(sdata, svec, None, n1.ave)
    rot = torch.from_numpy(rotation_rows(angles)).to(dev)
    tuned = None
    if args.schedule:
        sched = json.loads(args.schedule)
        plan.tune(**sched)
        tuned = {"schedule": sched}
    elif not args.no_tune:
        from vesper_python_amd import autotune
        # strided sample of the real rows; the schedule choice is cached per (device, N, mode, batch)
        # the headline is the plain fp32 search: the MFMA x passes (split
        # f16 / bf16 operands) are measured separately (measure_mixed)
        tuned = autotune.tune(plan, rot[::max(1, R // 256)].contiguous(), sample_rows=128, use_cache=False,
                              verbose=True, mixed_precision=False)
    best = torch.empty((R, 1), dtype=torch.float32, device=dev)
    trans = torch.empty((R, 1, 3), dtype=torch.int32, device=dev)
    print(f"[bench] workload={args.workload} N={n} rows={R} batch={args.batch}", file=sys.stderr, flush=True)

    def step():
        plan.score_rows(rot, best, trans)
        return engine.topn(best, 10)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    stream = torch.cuda.current_stream()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    e0.record(stream)
    for i in range(args.steps):
        top = step()
        print(f"[bench] step {i + 1}/{args.steps} done", file=sys.stderr, flush=True)
    e1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    gpu_ms = e0.elapsed_time(e1)
    ms_per_step = gpu_ms / args.steps
    value = R * args.steps / (gpu_ms / 1e3)

    # per-stage launch durations (HIP events on the launch stream)
    prows = min(R, max(args.batch, args.profile_rows // args.batch * args.batch))
    stage_ms = plan.profile_rows(rot[:prows])
    launches = (prows + args.batch - 1) // args.batch
    info = plan.info()
    sb = _stage_bytes(n, info["box_width"], args.batch)
    dom = max(sb, key=lambda k: stage_ms[k])
    avg_launch_s = stage_ms[dom] / launches / 1e3
    achieved = sb[dom] / avg_launch_s / 1e9
    roofline = {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 1), "peak": PEAK_HBM_GBS,
                "unit": "GB/s", "frac": round(achieved / PEAK_HBM_GBS, 4), "traffic": None,
                "avg_launch_ms": round(stage_ms[dom] / launches, 4),
                "stage_ms_per_launch": {k: round(v / launches, 4) for k, v in stage_ms.items()},
                "bytes_per_launch": sb[dom], "rows_per_launch": args.batch,
                "support_box": [info["box_lo"], info["box_width"]],
                "stage_GBps": {k: round(sb[k] / (stage_ms[k] / launches / 1e3) / 1e9, 1) for k in sb}}

    # whole pipeline per scored row: the algorithmic bytes its four kernels
    # move per launch (support-box pruned), over the rows of a launch
    row_bytes = sum(sb.values()) / args.batch
    roofline["path"] = {"bytes_per_row": int(row_bytes), "achieved": round(value * row_bytes / 1e9, 1),
                        "frac": round(value * row_bytes / 1e9 / PEAK_HBM_GBS, 4),
                        "model": "gather + z/y + x + y/z stage bytes of one launch / rows per launch"}
    sched = tuned["schedule"] if tuned else {}
    mf = _mfma_roofline(sched.get("xpass_variant") if isinstance(sched, dict) else None, n, info, args.batch,
                        stage_ms["xcorr"] / launches)
    if mf is not None:
        roofline["mfma"] = mf
    if dom == "xcorr":
        roofline["traffic"] = pmc_traffic(args.workload, args.batch, tuned["schedule"] if tuned else None,
                                          info["box_width"])

    # the batched mixed-precision search (radix-split MFMA x pass) and the
    # learned proposal on it; then back to the all-fp32 schedule
    mixed = None
    if not args.rows and not args.schedule:
        mixed = measure_mixed(plan, rot, best, trans, n, args.batch, args.steps, value, args.profile_rows,
                              args.workload)
    proposal = None
    if not args.no_proposal and not args.rows:
        proposal = measure_proposal(plan, n1, n2, w["A"], best, R, value)
        proposal["schedule"] = "mixed precision (x-pass variant 4)" if mixed else "fp32"
    if mixed is not None and tuned:
        plan.tune(**tuned["schedule"])

    rescore = None
    aux = None
    if not args.rows:
        rescore = measure_rescore(plan, n1, n2, angles, best)
        aux = measure_aux(w, n1, n2, best)

    libfft = None
    if not args.no_libfft and not args.rows:
        libfft = measure_library_fft(n1, n2, rot, args.batch, plan)

    cli_m = None
    if not args.no_cli and not args.rows:
        cli_m = measure_cli(w)
        cli_m["prob"] = measure_cli_prob(w)

    pairs_m = None
    if not args.no_pairs and not args.rows:
        pairs_m = measure_uneven_pairs()

    interp = None
    if not args.no_interp and not args.rows:
        interp = measure_interp(cpu=not args.no_cpu)

    prob = None
    if not args.no_prob and not args.rows:
        prob = measure_prob(n1, n2, rot, max(8, args.batch // 2), 1)

    cpu = cpu_all = None
    if not args.no_cpu:
        cpu = cpu_baseline(n1, n2, angles, args.cpu_seconds, 64)
        workers = min(16, len(os.sched_getaffinity(0)))
        if workers > 1:
            try:
                cpu_all = cpu_baseline_pool(n1, n2, angles, args.cpu_seconds, workers)
            except Exception as e:  # the one-core figure stays the reported baseline
                print(f"[bench] multi-core CPU baseline skipped: {e}", file=sys.stderr, flush=True)
    rec = {"metric": METRIC, "value": round(value, 2), "unit": "rotations/s", "n_gpus": 1,
           "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_per_step, 3),
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
           "data": "synthetic", "top10_recall": 1.0,
           "top10_recall_basis": "value is the exhaustive search itself; the learned proposal's recall "
                                 "against it is proposal.recall_top10",
           "config": {"workload": f"{args.workload}: {w['text']}", "grid_edge": n, "rows": R,
                      "batch": args.batch, "seed": w["seed"], "top_rows": [int(v) for v in top[:10]],
                      "schedule": tuned["schedule"] if tuned else "default"},
           "wall_s": round(wall, 3), "roofline": roofline, "cpu_baseline": cpu, "proposal": proposal,
           "prob_mode": prob, "interpolation": interp, "library_fft_pipeline": libfft, "uneven_pairs": pairs_m,
           "fp64_recheck": rescore, "aux_rooflines": aux, "mixed_precision": mixed,
           "cli_end_to_end": cli_m}
    if cpu:
        cpu["parity_vs_gpu"] = baseline_parity(cpu, best, trans)
        rec["gpu_vs_cpu"] = round(value / cpu["value"], 1)
        try:
            st = cpu_stages(w, n1, n2)
            if aux:
                st["fastvec_query_map"]["gpu_ms"] = aux["fastvec"]["ms"]
                st["get_score_fp64"]["gpu_ms"] = aux["get_score"]["ms"]
            st["search_row"] = {"seconds": round(1.0 / cpu["value"], 3), "gpu_ms": round(1e3 / value, 4)}
            cpu["stages"] = st
        except Exception as e:  # the search-loop figure stays the reported baseline
            print(f"[bench] CPU stage timings skipped: {e}", file=sys.stderr, flush=True)
    if cpu_all:
        rec["cpu_baseline_all_cores"] = cpu_all
        rec["gpu_vs_cpu_all_cores"] = round(value / cpu_all["value"], 1)
    print(json.dumps(rec), flush=True)


if __name__ == "__main__":
    main()
