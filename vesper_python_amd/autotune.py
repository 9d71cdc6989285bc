"""Autotuner for the search kernels' schedules.

For a plan (grid edge N, mode, batch), times every candidate schedule of the
x pass (register- or LDS-accumulating kernel, tile width), the slab shape or
pencil passes, the order of the FFT's radix stages and the number of batches
in flight on a sample of the
actual rotation rows with the library's per-stage HIP-event profiler
(vx_score_rows_profile), keeps the fastest, and caches it per
(device, N, mode, batch) in a JSON file.  The FFT schedules (variants 0, 1,
3) compute identical results; the MFMA x passes (variant 2, split-bf16 DFT;
variant 4, split-f16 radix split) are accepted only if their best scores on
the sample agree with the FFT schedule's within 1e-4 of the largest score.
"""
from __future__ import annotations

import json
import os
import sys
import threading

import numpy as np
import torch

CACHE = os.environ.get("VX_TUNE_CACHE",
                       os.path.join(os.path.expanduser("~"), ".cache", "vesper_python_amd", "tune.json"))
_lock = threading.Lock()


APPROX_RTOL = 1e-4


def candidates(n: int, groups: int = 1):
    out = [dict(xpass_variant=0, xpass_tile=0)]
    for t in (2, 4, 8, 12, 16):
        out.append(dict(xpass_variant=1, xpass_tile=t))
    for t in (1, 2, 4, 8, 12):
        if n * t <= 2048:
            out.append(dict(xpass_variant=3, xpass_tile=t))
    if groups == 1 and n <= 192:
        for t in (32, 64):
            out.append(dict(xpass_variant=2, xpass_tile=t))
    if groups == 1 and mr_factor(n) is not None:
        out.append(dict(xpass_variant=4, xpass_tile=0))
    return out


MFMA_VARIANTS = (2, 4)


def mr_factor(n: int):
    """(N1, N2) of the radix-split MFMA x pass for grid edge n: N2 the
    smallest of 2, 3, 4, 5, 7, 8 dividing n with N1 = n / N2 <= 32, else None
    (the library's choice)."""
    for r in (2, 3, 4, 5, 7, 8):
        if n % r == 0 and 2 <= n // r <= 32:
            return n // r, r
    return None


def radix_orders(n: int) -> dict:
    """The distinct FFT stage orders of grid edge n that the library's
    VX_TUNE_RADIX_ORDER values 0..3 give: {value: radices}."""
    two, odd, m = [], [], n
    for r in (8, 4, 2):
        while m % r == 0:
            two.append(r)
            m //= r
    for q in (3, 5, 7, 11, 13):
        while m % q == 0:
            odd.append(q)
            m //= q
    orders = {0: two + odd, 1: (two + odd)[::-1], 2: odd + two, 3: odd[::-1] + two}
    out, seen = {}, set()
    for k in range(4):
        t = tuple(orders[k])
        if t not in seen:
            seen.add(t)
            out[k] = t
    return out


def _key(plan) -> str:
    name = torch.cuda.get_device_name(torch.cuda.current_device()) if torch.cuda.is_available() else "?"
    return f"{name}|n={plan.n}|mode={plan.mode}|batch={plan.batch}|groups={plan.groups}"


def _load() -> dict:
    try:
        with open(CACHE) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def _store(d: dict) -> None:
    try:
        os.makedirs(os.path.dirname(CACHE), exist_ok=True)
        tmp = CACHE + ".tmp"
        with open(tmp, "w") as f:
            json.dump(d, f, indent=1, sort_keys=True)
        os.replace(tmp, CACHE)
    except OSError:
        pass


# launches over which one and two batches in flight are compared
STREAM_LAUNCHES = 32


def tune(plan, rot: torch.Tensor, sample_rows: int = 128, use_cache: bool = True, verbose: bool = False,
         mixed_precision: bool = True) -> dict:
    """Pick the fastest schedule for `plan` on rows `rot` (device [R, 9]);
    applies it and returns {"schedule": ..., "ms": {...}}.  With
    mixed_precision=False the MFMA x passes (split bf16 / f16 operands) are
    not candidates: every stage computes in fp32."""
    key = _key(plan) + ("" if mixed_precision else "|fp32")
    with _lock:
        cache = _load() if use_cache else {}
    if use_cache and key in cache:
        sched = cache[key]["schedule"]
        plan.tune(**sched)
        return dict(cache[key], cached=True)
    nrows = min(int(rot.shape[0]), max(plan.batch, sample_rows // plan.batch * plan.batch))
    sample = rot[:nrows]
    results = []
    exact = None
    plan.tune(radix_order=0)
    for cand in candidates(plan.n, plan.groups):
        if not mixed_precision and cand["xpass_variant"] in MFMA_VARIANTS:
            continue
        try:
            plan.tune(**cand)
            got, _ = plan.score_rows(sample[:plan.batch])  # warm-up of this schedule
            ms = plan.profile_rows(sample)
        except Exception as e:  # a schedule the device rejects is skipped
            if verbose:
                print("[tune] skip", cand, e, file=sys.stderr)
            continue
        got = got.double().cpu().numpy()
        if cand["xpass_variant"] not in MFMA_VARIANTS:
            exact = got if exact is None else exact
        elif exact is None or np.max(np.abs(got - exact)) > APPROX_RTOL * max(np.max(np.abs(exact)), 1e-30):
            if verbose:
                print("[tune] reject", cand, "(accuracy)", file=sys.stderr)
            continue
        total = sum(ms.values())
        results.append((total, cand, ms))
        if verbose:
            print(f"[tune] {cand} total={total:.3f} ms xcorr={ms['xcorr']:.3f}", file=sys.stderr, flush=True)
    results.sort(key=lambda r: r[0])
    total, best, ms = results[0]
    # slab workgroup shape or separate z and y passes, on top of the best x
    # pass (the stages are independent); the pencil passes round differently,
    # so they are held to the same 1e-4 gate as the MFMA x pass
    try:
        plan.tune(pencil=False)
        slab_ok = True
    except Exception:
        slab_ok = False   # the slab exceeds LDS: the pencil passes are the only form
    for cand in ((dict(best, slab_threads=512), dict(best, pencil=True)) if slab_ok else ()):
        try:
            plan.tune(**cand)
            plan.score_rows(sample[:plan.batch])
            ms2 = plan.profile_rows(sample)
        except Exception as e:
            if verbose:
                print("[tune] skip", cand, e, file=sys.stderr)
            continue
        t2 = sum(ms2.values())
        if verbose:
            print(f"[tune] {cand} total={t2:.3f} ms fwd_zy={ms2['fwd_zy']:.3f} inv_yz={ms2['inv_yz']:.3f}",
                  file=sys.stderr, flush=True)
        if cand.get("pencil"):
            got, _ = plan.score_rows(sample[:plan.batch])
            got = got.double().cpu().numpy()
            if exact is None or np.max(np.abs(got - exact)) > APPROX_RTOL * max(np.max(np.abs(exact)), 1e-30):
                plan.tune(pencil=False)
                continue
        results.append((t2, cand, ms2))
        if t2 < total:
            total, best, ms = t2, cand, ms2
        if cand.get("pencil"):
            plan.tune(pencil=False)
    if "slab_threads" not in best:
        best = dict(best, slab_threads=0)
    # the order of the FFT stages (equal to fp32 rounding: held to the 1e-4 gate)
    best = dict(best, radix_order=0)
    for order, radices in radix_orders(plan.n).items():
        if order == 0:
            continue
        cand = dict(best, radix_order=order)
        try:
            plan.tune(**cand)
            got, _ = plan.score_rows(sample[:plan.batch])
            ms2 = plan.profile_rows(sample)
        except Exception as e:
            if verbose:
                print("[tune] skip", cand, e, file=sys.stderr)
            continue
        got = got.double().cpu().numpy()
        t2 = sum(ms2.values())
        if verbose:
            print(f"[tune] radices {radices} total={t2:.3f} ms", file=sys.stderr, flush=True)
        if exact is None or np.max(np.abs(got - exact)) > APPROX_RTOL * max(np.max(np.abs(exact)), 1e-30):
            continue
        results.append((t2, cand, ms2))
        if t2 < total:
            total, best, ms = t2, cand, ms2
    plan.tune(radix_order=best["radix_order"])
    # two batches in flight: only a wall-clock measurement shows the overlap,
    # and only over enough launches (a 4-launch sample is mostly the fill
    # and drain of the pipeline): the sample repeated to 32 launches; two
    # streams unless one is more than 1 % faster
    if nrows >= 2 * plan.batch:
        reps = max(1, (STREAM_LAUNCHES * plan.batch + nrows - 1) // nrows)
        long_sample = sample.repeat(reps, 1)[:STREAM_LAUNCHES * plan.batch]
        walls = {}
        for ns in (1, 2):
            try:
                plan.tune(**best, streams=ns)
                plan.score_rows(sample)
                torch.cuda.synchronize()
                e0 = torch.cuda.Event(enable_timing=True)
                e1 = torch.cuda.Event(enable_timing=True)
                e0.record()
                for _ in range(2):
                    plan.score_rows(long_sample)
                e1.record()
                torch.cuda.synchronize()
                walls[ns] = e0.elapsed_time(e1) / 2
            except Exception as e:
                if verbose:
                    print("[tune] skip streams", ns, e, file=sys.stderr)
        if verbose and walls:
            print(f"[tune] streams wall ms over {STREAM_LAUNCHES} launches {walls}", file=sys.stderr, flush=True)
        if 2 in walls and (1 not in walls or walls[1] >= 0.99 * walls[2]):
            best = dict(best, streams=2)
        else:
            best = dict(best, streams=1)
    plan.tune(**best)
    rec = {"schedule": best, "ms_per_row": total / nrows, "rows": nrows,
           "tried": [{"schedule": c, "ms_per_row": t / nrows} for t, c, _ in results]}
    if use_cache:
        with _lock:
            cache = _load()
            cache[key] = rec
            _store(cache)
    return dict(rec, cached=False)
