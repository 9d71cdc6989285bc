"""Device-side API over libvesper_hip: device maps, search plans, fastVEC,
map rotation, FP64 scoring.  PyTorch-ROCm tensors are used only as device
allocations and for the current HIP stream."""
from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import _native as nat
from ._native import VesperError


def require_gpu() -> torch.device:
    if not torch.cuda.is_available():
        raise VesperError("no ROCm GPU visible: vesper_python_amd runs on MI355X only "
                          "(there is no CPU fallback)")
    nat.load()
    return torch.device("cuda", torch.cuda.current_device())


def _p(t) -> ctypes.c_void_p:
    return ctypes.c_void_p(0 if t is None else t.data_ptr())


def _stream() -> ctypes.c_void_p:
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


_NUMPY_OF = {torch.float32: np.float32, torch.float64: np.float64, torch.int32: np.int32}


def to_dev(a, dtype) -> torch.Tensor:
    """Copy a host array (or tensor) to the GPU as a contiguous `dtype` tensor."""
    dev = require_gpu()
    if isinstance(a, torch.Tensor):
        return a.to(device=dev, dtype=dtype).contiguous()
    return torch.from_numpy(np.ascontiguousarray(a, dtype=_NUMPY_OF[dtype])).to(dev)


def _host3(v) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(v, dtype=np.float64).reshape(3))


def fastvec(src_data, src_width, src_orig, n, dest_width, dest_orig, dreso=16.0):
    """Gaussian-weighted vector field of a thresholded source map
    (doFastVEC, VESPER_1_prob_new.py:180-257).  Returns device tensors
    data [n,n,n] float64 and vec [n,n,n,3] float32."""
    src = to_dev(src_data if isinstance(src_data, torch.Tensor) else np.asarray(src_data, dtype=np.float32),
                 torch.float32)
    sx, sy, sz = (int(s) for s in src.shape)
    data = torch.empty((n, n, n), dtype=torch.float64, device=src.device)
    vec = torch.empty((n, n, n, 3), dtype=torch.float32, device=src.device)
    so = _host3(np.asarray(src_orig, dtype=np.float32))
    do = _host3(dest_orig)
    nat.call("vx_fastvec", _p(src), sx, sy, sz, float(np.float32(src_width)),
             so.ctypes.data_as(ctypes.c_void_p), int(n), float(dest_width),
             do.ctypes.data_as(ctypes.c_void_p), float(dreso), _p(data), _p(vec), _stream())
    return data, vec


def map_stats(data: torch.Tensor) -> dict:
    """Statistics of VESPER_1_prob_new.py:157-175 computed on the device."""
    out = np.zeros(5, dtype=np.float64)
    d = data.contiguous()
    nat.call("vx_map_stats", _p(d), int(d.numel()), out.ctypes.data_as(ctypes.c_void_p), _stream())
    return dict(dsum=out[0], Nact=int(out[1]), ave=out[2], std=out[3], std_norm_ave=out[4])


def mode_planes(mode: int, data=None, vec=None, prob=None, ave: float = 0.0) -> torch.Tensor:
    """The float32 FFT input planes [C, n, n, n] of a map for a mode
    (vx_mode_planes): the 3 vector components (VX_MODE_VEC), the mode's
    transform of the density (O/C/P/L), or 3 vector + 4 probability planes
    (VX_MODE_PROB)."""
    ref = vec if vec is not None else data
    n = int(ref.shape[0])
    c = {nat.VX_MODE_VEC: 3, nat.VX_MODE_PROB: 7}.get(int(mode), 1)
    out = torch.empty((c, n, n, n), dtype=torch.float32, device=ref.device)
    nat.call("vx_mode_planes", n, int(mode), _p(data), _p(vec), _p(prob), float(ave), _p(out), _stream())
    return out


def laplacian(data: torch.Tensor) -> torch.Tensor:
    out = torch.empty_like(data)
    nat.call("vx_laplacian", int(data.shape[0]), _p(data), _p(out), _stream())
    return out


def rotate_map(rot9, data: torch.Tensor, vec: torch.Tensor, prob: torch.Tensor | None = None,
               flags: int = 0):
    """rot_mrc / rot_mrc_prob for one rotation (VESPER_1_prob_new.py:279-354,
    :1211-1299).  rot9 = float32(flip(M)) row-major."""
    n = int(data.shape[0])
    r = np.ascontiguousarray(np.asarray(rot9, dtype=np.float32).reshape(9))
    od = torch.empty_like(data)
    ov = torch.empty_like(vec)
    op = torch.empty_like(prob) if prob is not None else None
    nat.call("vx_rotate_map", n, r.ctypes.data_as(ctypes.c_void_p), _p(data), _p(vec), _p(prob),
             _p(od), _p(ov), _p(op), int(flags), _stream())
    return od, ov, op


def get_score(tdata, tvec, t_ave, sdata, svec, trans) -> dict:
    """get_score sums in FP64 (VESPER_1_prob_new.py:1302-1393)."""
    n = int(tdata.shape[0])
    t = np.ascontiguousarray(np.asarray(trans, dtype=np.int32).reshape(3))
    out = np.zeros(10, dtype=np.float64)
    nat.call("vx_get_score", n, _p(tdata), _p(tvec), float(t_ave), _p(sdata), _p(svec),
             t.ctypes.data_as(ctypes.c_void_p), out.ctypes.data_as(ctypes.c_void_p), _stream())
    return dict(Nm=int(out[0]), total=int(out[1]), cc_raw=out[2], pcc_raw=out[3], score=out[4],
                ave2=out[5], std2=out[6], pstd2=out[7])


def _ptrs(tensors):
    arr = (ctypes.c_void_p * len(tensors))(*[t.data_ptr() for t in tensors])
    return arr, ctypes.cast(arr, ctypes.c_void_p)


def sum_argmax(maps) -> tuple:
    """find_best_trans_list on the device: (float32 best, flat index) of the
    float32 channel sum of `maps` (device float32 tensors of one shape)."""
    maps = [m.contiguous() for m in maps]
    keep, ptr = _ptrs(maps)
    best = torch.empty(1, dtype=torch.float32, device=maps[0].device)
    idx = torch.empty(1, dtype=torch.int64, device=maps[0].device)
    nat.call("vx_sum_argmax", ptr, len(maps), int(maps[0].numel()), _p(best), _p(idx), _stream())
    return np.float32(best.item()), int(idx.item())


def mixed_argmax(vmaps, pmaps, alpha, vstd, vave, pstd, pave) -> tuple:
    """find_best_trans_mixed on the device: (float32 best, flat index)."""
    vmaps = [m.contiguous() for m in vmaps]
    pmaps = [m.contiguous() for m in pmaps]
    kv, pv = _ptrs(vmaps)
    kp, pp = _ptrs(pmaps)
    best = torch.empty(1, dtype=torch.float32, device=vmaps[0].device)
    idx = torch.empty(1, dtype=torch.int64, device=vmaps[0].device)
    nat.call("vx_mixed_argmax", pv, len(vmaps), pp, len(pmaps), int(vmaps[0].numel()), float(alpha),
             float(vstd), float(vave), float(pstd), float(pave), _p(best), _p(idx), _stream())
    return np.float32(best.item()), int(idx.item())


def sco_arr(tvec: torch.Tensor, svec: torch.Tensor, trans, out_dtype=torch.float64) -> torch.Tensor:
    """get_score's sco_arr (per-voxel float32 products at the search positions)."""
    n = int(tvec.shape[0])
    t = np.ascontiguousarray(np.asarray(trans, dtype=np.int32).reshape(3))
    out = torch.empty((n, n, n), dtype=out_dtype, device=tvec.device)
    nat.call("vx_sco_arr", n, _p(tvec), _p(svec), t.ctypes.data_as(ctypes.c_void_p),
             int(out_dtype == torch.float64), _p(out), _stream())
    return out


def topn(scores: torch.Tensor, k: int, stride: int = 1, col: int = 0) -> np.ndarray:
    """Rows of the k best scores, (score desc, row asc)."""
    nrows = scores.numel() // stride
    k_eff = min(int(k), nrows)
    rows = torch.empty(max(k_eff, 1), dtype=torch.int32, device=scores.device)
    nat.call("vx_topn", _p(scores), int(nrows), int(stride), int(col), int(max(k, 1)), _p(rows), _stream())
    return rows[:k_eff].cpu().numpy()


def topn_workspace(nrows: int, k: int, device=None) -> torch.Tensor:
    """A device workspace for vx_topn_ws over `nrows` rows and k <= 1024."""
    nb = ctypes.c_int64(0)
    nat.call("vx_topn_ws_bytes", int(nrows), int(max(k, 1)), ctypes.byref(nb))
    dev = device if device is not None else require_gpu()
    return torch.empty((nb.value + 7) // 8, dtype=torch.int64, device=dev)


def topn_into(scores: torch.Tensor, nrows: int, k: int, rows: torch.Tensor, ws: torch.Tensor,
              stride: int = 1, col: int = 0) -> None:
    """Rows of the k best scores into the device tensor `rows` (score desc,
    row asc), on the caller's workspace `ws` (topn_workspace): allocates
    nothing, so it can be recorded into a HIP graph."""
    nat.call("vx_topn_ws", _p(scores), int(nrows), int(stride), int(col), int(max(k, 1)), _p(rows), _p(ws),
             int(ws.numel() * ws.element_size()), _stream())


class SearchPlan:
    """A vx_plan: target spectra + query source for one (N, mode) search."""

    def __init__(self, n: int, mode: int, batch: int = 8, flags: int = 0):
        require_gpu()
        h = ctypes.c_void_p()
        nat.call("vx_plan_create", int(n), int(mode), int(batch), int(flags), ctypes.byref(h))
        self._h = h
        self.n, self.mode, self.batch, self.flags = int(n), int(mode), int(batch), int(flags)
        self._keep = []
        self.mix = False

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            nat.load().vx_plan_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def info(self) -> dict:
        out = np.zeros(8, dtype=np.int64)
        nat.call("vx_plan_info", self._h, out.ctypes.data_as(ctypes.c_void_p))
        return dict(n=int(out[0]), mode=int(out[1]), batch=int(out[2]), groups=int(out[3]),
                    channels=int(out[4]), workspace_bytes=int(out[5]), box_lo=int(out[6]),
                    box_width=int(out[7]))

    def tune(self, xpass_variant: int | None = None, xpass_tile: int | None = None,
             pencil: bool | None = None, slab_threads: int | None = None, streams: int | None = None,
             static_plans: bool | None = None, radix_order: int | None = None,
             slab_fused: bool | None = None, pencil_chunks: int | None = None):
        """Select kernel schedules.  x-pass variants 0, 1 and 3 (FFT), the
        slab shape and the inverse z pass's chunks per workgroup
        (pencil_chunks, 0 = automatic) give identical results; the
        compile-time stage plans (grid edges 16..200) and the fused slab
        kernels on them agree to fp32 rounding (bit-identical at 175); the
        pencil passes, the radix orders and variant 2 (split-bf16 MFMA DFT)
        agree within 1e-4 relative."""
        if radix_order is not None:
            nat.call("vx_plan_tune", self._h, nat.VX_TUNE_RADIX_ORDER, int(radix_order))
        if static_plans is not None:
            nat.call("vx_plan_tune", self._h, nat.VX_TUNE_STATIC_PLANS, int(bool(static_plans)))
        if slab_fused is not None:
            nat.call("vx_plan_tune", self._h, nat.VX_TUNE_SLAB_FUSED, int(bool(slab_fused)))
        if pencil_chunks is not None:
            nat.call("vx_plan_tune", self._h, nat.VX_TUNE_PENCIL_CHUNKS, int(pencil_chunks))
        if pencil is not None:
            nat.call("vx_plan_tune", self._h, nat.VX_TUNE_FORCE_PENCIL, int(bool(pencil)))
        if xpass_variant is not None:
            nat.call("vx_plan_tune", self._h, nat.VX_TUNE_XPASS_VARIANT, int(xpass_variant))
        if xpass_tile is not None:
            nat.call("vx_plan_tune", self._h, nat.VX_TUNE_XPASS_TILE, int(xpass_tile))
        if slab_threads is not None:
            nat.call("vx_plan_tune", self._h, nat.VX_TUNE_SLAB_THREADS, int(slab_threads))
        if streams is not None:
            nat.call("vx_plan_tune", self._h, nat.VX_TUNE_STREAMS, int(streams))

    def pin_box(self, lo: int, width: int):
        """Fix the support box to [lo, lo + width) on every axis (width 0 =
        recompute per query)."""
        nat.call("vx_plan_tune", self._h, nat.VX_TUNE_PIN_BOX_LO, int(lo))
        nat.call("vx_plan_tune", self._h, nat.VX_TUNE_PIN_BOX_WIDTH, int(width))

    def set_target(self, data, vec, prob=None, ave=0.0):
        nat.call("vx_set_target", self._h, _p(data), _p(vec), _p(prob), float(ave), _stream())

    def set_search(self, data, vec, prob=None, ave=0.0):
        self._keep = [data, vec, prob]
        nat.call("vx_set_search", self._h, _p(data), _p(vec), _p(prob), float(ave), _stream())

    def set_prob_stats(self, alpha, vstd, vave, pstd, pave, enable=True):
        self.mix = bool(enable)
        nat.call("vx_set_prob_stats", self._h, int(bool(enable)), float(alpha), float(vstd),
                 float(vave), float(pstd), float(pave))

    @property
    def groups(self) -> int:
        if self.mode == nat.VX_MODE_PROB:
            return 3 if self.mix else 2
        return 1

    def correlate(self, tspec: torch.Tensor, q: torch.Tensor) -> torch.Tensor:
        """Per-channel unnormalised correlation maps irfftn(tspec * rfftn(q))
        (fft_search_best_dot, VESPER_1_prob_new.py:464-491).  tspec: [C, n, n,
        n/2+1] complex64 (conjugated target spectra), q: [C, n, n, n] float32."""
        tspec = tspec.contiguous()
        q = q.contiguous()
        c = int(q.shape[0])
        if tuple(q.shape[1:]) != (self.n,) * 3 or tuple(tspec.shape) != (c, self.n, self.n, self.n // 2 + 1):
            raise VesperError("correlate: shapes do not match the plan's grid edge")
        out = torch.empty_like(q)
        nat.call("vx_correlate", self._h, _p(tspec), c, _p(q), _p(out), _stream())
        return out

    def score_rows(self, rot: torch.Tensor, best: torch.Tensor | None = None,
                   trans: torch.Tensor | None = None):
        """Score every row of rot [R, 9] (device float32).  Returns (best [R, G]
        float32, trans [R, G, 3] int32) device tensors."""
        R = int(rot.shape[0])
        G = self.groups
        if best is None:
            best = torch.empty((R, G), dtype=torch.float32, device=rot.device)
        if trans is None:
            trans = torch.empty((R, G, 3), dtype=torch.int32, device=rot.device)
        if R:
            nat.call("vx_score_rows", self._h, _p(rot), R, _p(best), _p(trans), _stream())
        return best, trans

    def query_planes(self, rot: torch.Tensor) -> torch.Tensor:
        """The float32 query planes [C, n, n, n] that score_rows correlates for
        the rotation rot [9] (device float32), formed by the search's own
        gather (vx_query_planes); zero outside the support box."""
        C = int(self.info()["channels"])
        out = torch.empty((C, self.n, self.n, self.n), dtype=torch.float32, device=rot.device)
        nat.call("vx_query_planes", self._h, _p(rot.contiguous()), _p(out), _stream())
        return out

    def rescore_f64(self, rot: torch.Tensor, target, search, ave: float = 0.0, window: float = 1e-5,
                    max_cand: int = 64):
        """FP64 scores and translations of the rows of rot [R, 9] (device
        float32), decided among the translations within `window` (relative)
        of each row's float32 maximum (vx_rescore_f64).  target / search =
        (data float64, vec float32, prob float32 [4, n, n, n] or None): the
        maps of set_target / set_search.  Returns (score [R, G] float64, trans
        [R, G, 3] int32, ncand [R, G] int32) device tensors.  A row with more
        candidates than max_cand (a plateau of equal values) is re-run with
        room for all of them, up to 4096; beyond that it raises."""
        R = int(rot.shape[0])
        G = self.groups
        score = torch.empty((R, G), dtype=torch.float64, device=rot.device)
        trans = torch.empty((R, G, 3), dtype=torch.int32, device=rot.device)
        ncand = torch.zeros((R, G), dtype=torch.int32, device=rot.device)
        if R:
            td, tv, tp = (tuple(target) + (None, None, None))[:3]
            sd, sv, sp = (tuple(search) + (None, None, None))[:3]
            mc = int(max_cand)
            while True:
                nat.call("vx_rescore_f64", self._h, _p(td), _p(tv), _p(tp), float(ave), _p(sd), _p(sv), _p(sp),
                         _p(rot), R, float(window), mc, _p(score), _p(trans), _p(ncand), _stream())
                worst = int(ncand.max().item())
                if worst <= mc:
                    break
                if mc >= 4096:
                    raise VesperError(f"FP64 re-scoring: a row has {worst} translations within the window "
                                      f"{window} of its maximum (at most 4096)")
                mc = min(4096, 1 << (worst - 1).bit_length())
        return score, trans, ncand

    STAGES = ("gather", "fwd_zy", "xcorr", "inv_yz", "mixed", "finalize")

    def profile_rows(self, rot: torch.Tensor) -> dict:
        """Per-stage milliseconds (HIP events on the launch stream) of scoring
        every row of rot, summed over batches."""
        R = int(rot.shape[0])
        G = self.groups
        best = torch.empty((R, G), dtype=torch.float32, device=rot.device)
        trans = torch.empty((R, G, 3), dtype=torch.int32, device=rot.device)
        ms = np.zeros(6, dtype=np.float64)
        nat.call("vx_score_rows_profile", self._h, _p(rot), R, _p(best), _p(trans),
                 ms.ctypes.data_as(ctypes.c_void_p), _stream())
        return dict(zip(self.STAGES, ms.tolist()))
