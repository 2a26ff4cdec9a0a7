"""Host-side mirror of the reference estimator interface over the C ABI (ctypes).

`EmbreeBre` and `PhotonBeamsEvaluator` keep the method shapes of the reference
classes (SmallUPBP/src/Bre/Bre.hxx:40-184, SmallUPBP/src/Beams/PhBeams.hxx:46-160):
build(...), evalBre / evalBeamBeamEstimate(...), destroy(). Evaluation is batched:
one call evaluates every query ray of a `Queries` batch on the GPU.

The library is libupbp_hip.so, built in-tree for gfx950. There is no CPU
fallback: importing works anywhere, but creating a Context raises when the
library is missing or no gfx950 device is present.
"""
from __future__ import annotations

import ctypes as C
import os
from typing import Optional

import numpy as np

from . import _abi
from .scenes import Beams, Photons, Queries

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libupbp_hip.so")
_lib = None

EXPORTED = [
    "upbp_ctx_create", "upbp_ctx_destroy", "upbp_ctx_last_error", "upbp_ctx_set_stream",
    "upbp_ctx_get_stream", "upbp_ctx_set_sync", "upbp_set_media", "upbp_bre_build", "upbp_bre_destroy",
    "upbp_bre_set_power", "upbp_bre_count", "upbp_bre_eval_brute_force", "upbp_beams_build", "upbp_beams_destroy", "upbp_beams_count",
    "upbp_beams_piece_count", "upbp_beams_selection_pdf", "upbp_beams_selection_pdfs", "upbp_beams_grid_stats",
    "upbp_eval", "upbp_eval_noalloc", "upbp_hit_signatures",
    "upbp_adjoint", "upbp_batch_plan_create", "upbp_batch_plan_destroy", "upbp_batch_run",
    "upbp_set_tuning", "upbp_get_tuning", "upbp_autotune", "upbp_eval_upbp", "upbp_bre_radii",
    "upbp_beams_radii", "upbp_hash_grid_build", "upbp_hash_grid_destroy", "upbp_hash_grid_count",
    "upbp_pp3d_eval", "upbp_surf_eval", "upbp_trace_light_paths",
]


class UpbpError(RuntimeError):
    def __init__(self, status: int, message: str):
        super().__init__("%s: %s" % (_abi.STATUS_NAMES.get(status, status), message))
        self.status = status
        self.message = message


def load():
    """Load libupbp_hip.so (raises ImportError when it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError("libupbp_hip.so is not built (run __graft_entry__.build() or "
                          "python smallupbp_amd/build_ext.py)")
    try:
        # torch ships its own HIP runtime (SONAME libamdhip64.so.7). Loading it first makes the
        # dynamic loader resolve this library's libamdhip64.so.7 to that same copy, so one HIP
        # runtime serves the process; two runtimes in one process cannot share the GPU.
        import torch  # noqa: F401
    except ImportError:
        pass
    L = C.CDLL(LIB_PATH)
    vp, i64, i32 = C.c_void_p, C.c_int64, C.c_int32
    st = C.c_int
    L.upbp_ctx_create.restype = st
    L.upbp_ctx_create.argtypes = [C.c_int, C.POINTER(vp)]
    L.upbp_ctx_destroy.restype = st
    L.upbp_ctx_destroy.argtypes = [vp]
    L.upbp_ctx_last_error.restype = C.c_char_p
    L.upbp_ctx_last_error.argtypes = [vp]
    L.upbp_ctx_set_stream.restype = st
    L.upbp_ctx_set_stream.argtypes = [vp, vp]
    L.upbp_ctx_get_stream.restype = vp
    L.upbp_ctx_get_stream.argtypes = [vp]
    L.upbp_ctx_set_sync.restype = st
    L.upbp_ctx_set_sync.argtypes = [vp, C.c_int]
    L.upbp_set_media.restype = st
    L.upbp_set_media.argtypes = [vp, vp, i32]
    L.upbp_bre_build.restype = st
    L.upbp_bre_build.argtypes = [vp, C.POINTER(_abi.PhotonInput), C.POINTER(_abi.BuildParams), C.POINTER(vp),
                                 C.POINTER(i64)]
    L.upbp_bre_destroy.restype = st
    L.upbp_bre_destroy.argtypes = [vp]
    L.upbp_bre_set_power.restype = st
    L.upbp_bre_set_power.argtypes = [vp, vp, C.c_int]
    L.upbp_bre_count.restype = i64
    L.upbp_bre_count.argtypes = [vp]
    L.upbp_bre_radii.restype = st
    L.upbp_bre_radii.argtypes = [vp, vp, C.c_int]
    L.upbp_beams_radii.restype = st
    L.upbp_beams_radii.argtypes = [vp, vp, vp, C.c_int]
    L.upbp_beams_build.restype = st
    L.upbp_beams_build.argtypes = [vp, C.POINTER(_abi.BeamInput), C.POINTER(_abi.BuildParams), C.POINTER(vp)]
    L.upbp_beams_destroy.restype = st
    L.upbp_beams_destroy.argtypes = [vp]
    L.upbp_beams_count.restype = i64
    L.upbp_beams_count.argtypes = [vp]
    L.upbp_beams_piece_count.restype = i64
    L.upbp_beams_piece_count.argtypes = [vp]
    L.upbp_beams_selection_pdf.restype = C.c_float
    L.upbp_beams_selection_pdf.argtypes = [vp, vp]
    L.upbp_beams_selection_pdfs.restype = C.c_int
    L.upbp_beams_selection_pdfs.argtypes = [vp, vp, C.c_int64, vp, C.c_int]
    L.upbp_beams_grid_stats.restype = C.c_int
    L.upbp_bre_eval_brute_force.restype = C.c_int
    L.upbp_bre_eval_brute_force.argtypes = [vp, vp, vp, vp, vp, vp, vp, C.c_int64, vp, C.c_int]
    L.upbp_beams_grid_stats.argtypes = [vp, vp, C.POINTER(_abi.QueryBatch), vp, C.c_int]
    L.upbp_eval.restype = st
    L.upbp_eval.argtypes = [vp, vp, vp, C.POINTER(_abi.QueryBatch), vp, vp, C.c_int, C.POINTER(_abi.EvalStats)]
    L.upbp_eval_noalloc.restype = st
    L.upbp_eval_noalloc.argtypes = [vp, vp, vp, C.POINTER(_abi.QueryBatch), vp, vp, C.POINTER(_abi.EvalStats)]
    L.upbp_eval_upbp.restype = st
    L.upbp_eval_upbp.argtypes = [vp, vp, vp, C.POINTER(_abi.QueryBatch), C.POINTER(_abi.MisInputs), vp, vp, C.c_int,
                                 C.POINTER(_abi.EvalStats)]
    L.upbp_hit_signatures.restype = st
    L.upbp_hit_signatures.argtypes = [vp, vp, vp, C.POINTER(_abi.QueryBatch), vp, vp, C.c_int]
    L.upbp_adjoint.restype = st
    L.upbp_adjoint.argtypes = [vp, vp, vp, C.POINTER(_abi.QueryBatch), vp, vp, vp, vp, vp, vp, C.c_int,
                               C.POINTER(_abi.EvalStats)]
    L.upbp_batch_plan_create.restype = st
    L.upbp_batch_plan_create.argtypes = [vp, C.POINTER(_abi.Instance), i32, C.POINTER(vp)]
    L.upbp_batch_plan_destroy.restype = st
    L.upbp_batch_plan_destroy.argtypes = [vp]
    L.upbp_batch_run.restype = st
    L.upbp_batch_run.argtypes = [vp, vp, C.c_int, C.POINTER(_abi.EvalStats)]
    L.upbp_set_tuning.restype = st
    L.upbp_set_tuning.argtypes = [vp, C.POINTER(_abi.Tuning)]
    L.upbp_get_tuning.restype = st
    L.upbp_get_tuning.argtypes = [vp, C.POINTER(_abi.Tuning)]
    L.upbp_autotune.restype = st
    L.upbp_autotune.argtypes = [vp, vp, vp, C.POINTER(_abi.QueryBatch), C.POINTER(_abi.Tuning),
                                C.POINTER(C.c_float)]
    L.upbp_hash_grid_build.restype = st
    L.upbp_hash_grid_build.argtypes = [vp, vp, vp, i64, C.c_float, i64, C.c_uint32, C.c_int, C.POINTER(vp)]
    L.upbp_hash_grid_destroy.restype = st
    L.upbp_hash_grid_destroy.argtypes = [vp]
    L.upbp_hash_grid_count.restype = i64
    L.upbp_hash_grid_count.argtypes = [vp]
    L.upbp_trace_light_paths.restype = st
    L.upbp_trace_light_paths.argtypes = [vp, vp, i32, vp, i32, C.POINTER(_abi.LightTraceParams), i64,
                                         C.POINTER(_abi.LightPaths), C.c_int]
    for name in ("upbp_pp3d_eval", "upbp_surf_eval"):
        getattr(L, name).restype = st
        getattr(L, name).argtypes = [vp, vp, vp, i64, C.POINTER(_abi.MisInputs), C.c_uint32, vp, vp, C.c_int,
                                     C.POINTER(_abi.EvalStats)]
    _lib = L
    return L


def _is_torch(x) -> bool:
    return hasattr(x, "data_ptr") and hasattr(x, "is_cuda")


class Context:
    """One upbp_ctx on one GPU (the library's own stream unless set_stream is used)."""

    def __init__(self, device: int = 0):
        self.lib = load()
        h = C.c_void_p()
        r = self.lib.upbp_ctx_create(device, C.byref(h))
        if r:
            raise UpbpError(r, self.lib.upbp_ctx_last_error(None).decode())
        self.h = h
        self.device = device
        self.n_media = 0

    def check(self, r: int):
        if r:
            raise UpbpError(r, self.lib.upbp_ctx_last_error(self.h).decode())

    def set_media(self, media: np.ndarray):
        m = np.ascontiguousarray(media)
        self.check(self.lib.upbp_set_media(self.h, m.ctypes.data, len(m)))
        self.n_media = len(m)

    def set_stream(self, stream_handle: Optional[int]):
        self.check(self.lib.upbp_ctx_set_stream(self.h, stream_handle))

    def stream(self) -> int:
        return self.lib.upbp_ctx_get_stream(self.h) or 0

    def set_sync(self, sync: bool):
        self.check(self.lib.upbp_ctx_set_sync(self.h, 1 if sync else 0))

    def set_tuning(self, waves_per_block: int, blocks_per_cu: int = 0, photon_waves_per_block: int = 0,
                   frustum_leaf: int = 0):
        """Waves per workgroup of the traversal kernels (the photon kernel's own count when
        photon_waves_per_block > 0), resident blocks per CU of the persistent kernels, and the
        largest subtree a wave whose rays share their origin tests as one leaf (0 = default)."""
        t = _abi.Tuning(waves_per_block, blocks_per_cu, photon_waves_per_block, frustum_leaf)
        self.check(self.lib.upbp_set_tuning(self.h, C.byref(t)))

    def get_tuning(self):
        t = _abi.Tuning()
        self.check(self.lib.upbp_get_tuning(self.h, C.byref(t)))
        return t.waves_per_block, t.blocks_per_cu, t.photon_waves_per_block, t.frustum_leaf

    def close(self):
        if getattr(self, "h", None):
            self.lib.upbp_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---------------------------------------------------------------------------
class QueryBatchView:
    """A upbp_query_batch over host numpy arrays or device torch tensors."""

    def __init__(self, origin, direction, seg_offsets, segments, beam_type=_abi.LONG_BEAM, lite=False,
                 order=_abi.ORDER_SORTED):
        self.device = _is_torch(origin) and origin.is_cuda
        mem = _abi.MEM_DEVICE if self.device else _abi.MEM_HOST
        self.count = int(origin.shape[0])
        self._keep = (origin, direction, seg_offsets, segments)
        self.qb = _abi.QueryBatch(self.count, _abi.ptr(origin), _abi.ptr(direction), _abi.ptr(seg_offsets),
                                  _abi.ptr(segments), beam_type, 1 if lite else 0, order, mem)

    @staticmethod
    def host(q: Queries, order=_abi.ORDER_SORTED):
        return QueryBatchView(np.ascontiguousarray(q.origin, np.float32), np.ascontiguousarray(q.direction, np.float32),
                              np.ascontiguousarray(q.seg_offsets, np.int64), np.ascontiguousarray(q.segments),
                              q.beam_type, q.lite, order)

    @staticmethod
    def to_device(q: Queries, device="cuda", order=_abi.ORDER_SORTED):
        import torch

        seg = np.ascontiguousarray(q.segments).view(np.float32).reshape(-1, 8)
        return QueryBatchView(
            torch.from_numpy(np.ascontiguousarray(q.origin, np.float32)).to(device),
            torch.from_numpy(np.ascontiguousarray(q.direction, np.float32)).to(device),
            torch.from_numpy(np.ascontiguousarray(q.seg_offsets, np.int64)).to(device),
            torch.from_numpy(seg.copy()).to(device),
            q.beam_type, q.lite, order)


def _arr(x, dtype):
    if x is None:
        return None
    if _is_torch(x):
        return x.contiguous()
    return np.ascontiguousarray(x, dtype=dtype)


class EmbreeBre:
    """PB2D photon-point estimator (reference: EmbreeBre, SmallUPBP/src/Bre/Bre.hxx:40-184)."""

    def __init__(self, ctx: Context):
        self.ctx = ctx
        self.h = None
        self.num_photons = 0

    def build(self, photons: Photons, radius: Optional[float] = None, in_medium=None, leaf_size: int = 0,
              radius_calculation: int = _abi.CONSTANT_RADIUS, knn: int = 0) -> int:
        """EmbreeBre::build (Bre.cxx:526-632). Returns the number of photons kept."""
        assert self.h is None, "destroy() before building again (Bre.cxx:529-532)"
        arrs = [_arr(photons.position, np.float32), _arr(photons.inc_dir, np.float32),
                _arr(photons.power, np.float32), _arr(photons.radius, np.float32), _arr(in_medium, np.uint8)]
        dev = any(_is_torch(a) and a.is_cuda for a in arrs if a is not None)
        pin = _abi.PhotonInput(photons.count, *[_abi.ptr(a) for a in arrs],
                               _abi.MEM_DEVICE if dev else _abi.MEM_HOST)
        r = photons.const_radius if radius is None else radius
        params = _abi.BuildParams(radius_calculation, float(r or 0.0), knn, leaf_size, 0.0, 0)
        h = C.c_void_p()
        n = C.c_int64()
        self.ctx.check(self.ctx.lib.upbp_bre_build(self.ctx.h, C.byref(pin), C.byref(params), C.byref(h), C.byref(n)))
        self.h = h
        self.num_photons = n.value
        self.num_input = photons.count
        return n.value

    def radii(self) -> np.ndarray:
        """Radius the build gave each input photon (0 where the in-medium filter dropped it)."""
        r = np.zeros(self.num_input, np.float32)
        self.ctx.check(self.ctx.lib.upbp_bre_radii(self.h, r.ctypes.data, _abi.MEM_HOST))
        return r

    def set_power(self, power):
        p = _arr(power, np.float32)
        mem = _abi.MEM_DEVICE if (_is_torch(p) and p.is_cuda) else _abi.MEM_HOST
        self.ctx.check(self.ctx.lib.upbp_bre_set_power(self.h, _abi.ptr(p), mem))

    def evalBreBruteForce(self, origin, direction, min_t, max_t, medium_id):
        """EmbreeBre::evalBreBruteForce (Bre.cxx:980-1029) for n rays (host arrays: origin and
        direction [n,3], min_t, max_t [n], medium_id [n]) -> [n,3]."""
        o = np.ascontiguousarray(origin, np.float32).reshape(-1, 3)
        d = np.ascontiguousarray(direction, np.float32).reshape(-1, 3)
        n = o.shape[0]
        t0 = np.ascontiguousarray(min_t, np.float32).reshape(n)
        t1 = np.ascontiguousarray(max_t, np.float32).reshape(n)
        m = np.ascontiguousarray(medium_id, np.int32).reshape(n)
        if d.shape[0] != n:
            raise ValueError("origin and direction must hold the same number of rays")
        out = np.zeros((n, 3), np.float32)
        self.ctx.check(self.ctx.lib.upbp_bre_eval_brute_force(self.ctx.h, self.h, _abi.ptr(o), _abi.ptr(d),
                                                              _abi.ptr(t0), _abi.ptr(t1), _abi.ptr(m), n,
                                                              _abi.ptr(out), _abi.MEM_HOST))
        return out

    def evalBre(self, queries: QueryBatchView, out=None, stats=None):
        """EmbreeBre::evalBre (Bre.cxx:812-964) for every query of the batch -> [Q,3]."""
        return eval_batch(self.ctx, self, None, queries, out_photons=out, stats=stats)[0]

    def destroy(self):
        if self.h is not None:
            self.ctx.lib.upbp_bre_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


class PhotonBeamsEvaluator:
    """BB1D photon-beam estimator (reference: PhotonBeamsEvaluator, SmallUPBP/src/Beams/PhBeams.hxx:46-160)."""

    def __init__(self, ctx: Context):
        self.ctx = ctx
        self.h = None

    def build(self, beams: Beams, radius: Optional[float] = None, start_radius=None, end_radius=None,
              leaf_size: int = 0, piece_factor: float = 0.0, max_pieces: int = 0,
              radius_calculation: int = _abi.CONSTANT_RADIUS, knn: int = 0, max_beams_in_cell: int = 0,
              reduction_type: int = 0, grid_resolution: int = 0, reduction_seed: int = 1234):
        """PhotonBeamsEvaluator::build (PhBeams.cxx:66-148); max_beams_in_cell > 0 turns on the
        per-cell beam reduction of the reference grid (-gridmax / -gridred)."""
        assert self.h is None, "destroy() before building again"
        arrs = [_arr(beams.origin, np.float32), _arr(beams.direction, np.float32), _arr(beams.length, np.float32),
                _arr(beams.throughput, np.float32), _arr(beams.flags, np.uint32), _arr(beams.medium_id, np.int32),
                _arr(start_radius, np.float32), _arr(end_radius, np.float32)]
        dev = any(_is_torch(a) and a.is_cuda for a in arrs if a is not None)
        bin_ = _abi.BeamInput(beams.count, *[_abi.ptr(a) for a in arrs], _abi.MEM_DEVICE if dev else _abi.MEM_HOST)
        r = beams.radius if radius is None else radius
        params = _abi.BuildParams(radius_calculation, float(r), knn, leaf_size, float(piece_factor), int(max_pieces),
                                  int(max_beams_in_cell), int(reduction_type), int(grid_resolution), int(reduction_seed))
        h = C.c_void_p()
        self.ctx.check(self.ctx.lib.upbp_beams_build(self.ctx.h, C.byref(bin_), C.byref(params), C.byref(h)))
        self.h = h
        self.num_beams = beams.count
        return self

    @property
    def pieces(self) -> int:
        return self.ctx.lib.upbp_beams_piece_count(self.h)

    def radii(self):
        """(mStartRadius, mEndRadius) the build gave each beam."""
        r0 = np.zeros(self.num_beams, np.float32)
        r1 = np.zeros(self.num_beams, np.float32)
        self.ctx.check(self.ctx.lib.upbp_beams_radii(self.h, r0.ctypes.data, r1.ctypes.data, _abi.MEM_HOST))
        return r0, r1

    def getBeamSelectionPdf(self, pos) -> float:
        p = np.ascontiguousarray(pos, np.float32)
        if p.size != 3:
            raise ValueError("pos must be one xyz position")
        pdf = self.ctx.lib.upbp_beams_selection_pdf(self.h, p.ctypes.data)
        if pdf != pdf:  # NaN: the device read failed
            raise UpbpError(_abi.UPBP_EDEVICE, self.ctx.lib.upbp_ctx_last_error(self.ctx.h).decode())
        return pdf

    def getBeamSelectionPdfs(self, positions):
        """Selection pdf of the cell holding each position ([n,3] numpy or device tensor -> [n])."""
        if _is_torch(positions) and positions.is_cuda:
            import torch

            if positions.ndim != 2 or positions.shape[1] != 3:
                raise ValueError("positions must have shape [n, 3], got %s" % (tuple(positions.shape),))
            p = positions.contiguous().float()
            out = torch.empty(p.shape[0], dtype=torch.float32, device=p.device)
            mem = _abi.MEM_DEVICE
        else:
            p = np.ascontiguousarray(positions, np.float32)
            if p.size % 3 != 0:
                raise ValueError("positions must hold xyz triples, got %d floats" % p.size)
            p = p.reshape(-1, 3)
            out = np.empty(p.shape[0], np.float32)
            mem = _abi.MEM_HOST
        self.ctx.check(self.ctx.lib.upbp_beams_selection_pdfs(self.h, _abi.ptr(p), p.shape[0], _abi.ptr(out), mem))
        return out

    def grid_stats(self, queries: QueryBatchView):
        """GridStats of the reference grid's walk per query (Grid.hxx:318-399, PhBeams.cxx:177-230)
        -> host [Q,3] u32: beams tested, cells crossed, overfull cells crossed."""
        out = np.zeros((queries.count, 3), np.uint32)
        self.ctx.check(self.ctx.lib.upbp_beams_grid_stats(self.ctx.h, self.h, C.byref(queries.qb), _abi.ptr(out),
                                                          _abi.MEM_HOST))
        return out

    def evalBeamBeamEstimate(self, queries: QueryBatchView, out=None, stats=None):
        """PhotonBeamsEvaluator::evalBeamBeamEstimate (PhBeams.cxx:177-324) per query -> [Q,3]."""
        return eval_batch(self.ctx, None, self, queries, out_beams=out, stats=stats)[1]

    def destroy(self):
        if self.h is not None:
            self.ctx.lib.upbp_beams_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


def trace_light_paths(ctx: Context, scene, path_count: int, device: bool = True, capacity=None):
    """upbp_trace_light_paths over a scenes.LightScene: light sub-paths traced in the scene's
    media from its point lights. Returns a dict of arrays (torch tensors on the GPU when
    `device`, numpy otherwise) trimmed to the vertices and beams produced, with path_ends."""
    from . import scenes

    cap_v, cap_b = capacity or scenes.light_path_capacity(scene, path_count)
    if device:
        import torch

        def zeros(shape, dtype):
            dt = np.dtype(dtype)
            n = int(np.prod(shape)) if isinstance(shape, tuple) else int(shape)
            return torch.zeros(max(n, 1) * dt.itemsize, dtype=torch.uint8, device="cuda")

        arr = {k: zeros(v.shape, v.dtype) for k, v in scenes.light_path_arrays(cap_v, cap_b, path_count).items()}
    else:
        arr = scenes.light_path_arrays(cap_v, cap_b, path_count)
    lp = _abi.light_paths(arr, cap_v, cap_b)
    regions = np.ascontiguousarray(scene.regions)
    lights = np.ascontiguousarray(scene.lights)
    par = _abi.LightTraceParams()
    for k, v in scene.params.items():
        setattr(par, k, v)
    ctx.check(ctx.lib.upbp_trace_light_paths(ctx.h, regions.ctypes.data if len(regions) else None, len(regions),
                                             lights.ctypes.data, len(lights), C.byref(par), path_count, C.byref(lp),
                                             _abi.MEM_DEVICE if device else _abi.MEM_HOST))
    nv, nb = int(lp.vertex_count), int(lp.beam_count)
    if not device:
        return scenes.trim_light_paths(arr, nv, nb, path_count)
    host = scenes.light_path_arrays(1, 1, 1)
    out = {}
    for k, t in arr.items():
        dt = host[k].dtype
        n = path_count if k == "path_ends" else (nv if k in ("vertices", "positions") else nb)
        per = dt.itemsize * (3 if host[k].ndim == 2 else 1)
        out[k] = t[: n * per]
    out["_dtypes"] = {k: (host[k].dtype, host[k].shape[1:]) for k in host}
    return out


def light_paths_to_host(d: dict) -> dict:
    """The device result of trace_light_paths as typed numpy arrays."""
    out = {}
    for k, (dt, tail) in d["_dtypes"].items():
        raw = d[k].cpu().numpy()
        out[k] = raw.view(dt).reshape((-1,) + tuple(tail)) if len(raw) else np.zeros((0,) + tuple(tail), dt)
    return out


class HashGrid:
    """Vertex merging through the reference's hashed grid (SmallUPBP/src/Misc/HashGrid.hxx:38-230)
    with RangeQuery::Process (SmallUPBP/src/Renderers/UPBP.hxx:63-133): PP3D for camera vertices in
    media, SURF for camera vertices on surfaces."""

    QUERY_BYTES = {_abi.PP3D: 48, _abi.SURF: 128}

    def __init__(self, ctx: Context):
        self.ctx = ctx
        self.h = None
        self.technique = None

    def build(self, positions, light_vertices, radius: float, cell_count: int, technique: int = _abi.PP3D):
        """HashGrid::Reserve(cell_count) + Build(light vertices, radius, technique) (UPBP.hxx:689-703).
        positions [n,3] float32 and light_vertices (LIGHT_VERTEX_DTYPE) as host arrays, or torch
        CUDA tensors (the light vertices as a uint8 tensor of n * 112 bytes)."""
        assert self.h is None, "destroy() before building again"
        pos = _arr(positions, np.float32)
        lv = light_vertices if _is_torch(light_vertices) else np.ascontiguousarray(light_vertices)
        dev = _is_torch(pos) and pos.is_cuda
        h = C.c_void_p()
        self.ctx.check(self.ctx.lib.upbp_hash_grid_build(self.ctx.h, _abi.ptr(pos), _abi.ptr(lv), int(pos.shape[0]),
                                                         float(radius), int(cell_count), int(technique),
                                                         _abi.MEM_DEVICE if dev else _abi.MEM_HOST, C.byref(h)))
        self.h = h
        self.technique = int(technique)
        return self

    @property
    def count(self) -> int:
        return int(self.ctx.lib.upbp_hash_grid_count(self.h))

    def eval(self, queries, mis, flags: int = 0, out=None, signatures: bool = False, stats: Optional[dict] = None):
        """RangeQuery::GetContrib per camera vertex ([Q,3]); with signatures=True, (Process calls,
        hash sum) per camera vertex ([Q,2] uint32) instead. `queries` is a PP3D_QUERY_DTYPE /
        SURF_QUERY_DTYPE array (host) or a uint8 CUDA tensor of those records; `mis` a
        scenes.MisInputs or an _abi.MisInputs (e.g. from device_mis)."""
        dev = _is_torch(queries)
        qb = self.QUERY_BYTES[self.technique]
        q_count = (int(queries.numel()) // qb) if dev else len(queries)
        q = queries if dev else np.ascontiguousarray(queries)
        ms = mis if isinstance(mis, _abi.MisInputs) else _abi.mis_inputs(mis)
        if dev:
            import torch

            res = out if out is not None else torch.empty((q_count, 2 if signatures else 3),
                                                          dtype=torch.int32 if signatures else torch.float32,
                                                          device=queries.device)
        else:
            res = out if out is not None else np.zeros((q_count, 2 if signatures else 3),
                                                       np.uint32 if signatures else np.float32)
        st = _abi.EvalStats()
        fn = self.ctx.lib.upbp_pp3d_eval if self.technique == _abi.PP3D else self.ctx.lib.upbp_surf_eval
        self.ctx.check(fn(self.ctx.h, self.h, _abi.ptr(q), q_count, C.byref(ms), int(flags),
                          None if signatures else _abi.ptr(res), _abi.ptr(res) if signatures else None,
                          _abi.MEM_DEVICE if dev else _abi.MEM_HOST, C.byref(st) if stats is not None else None))
        if stats is not None:
            stats.update(st.as_dict())
        return res

    def destroy(self):
        if self.h is not None:
            self.ctx.lib.upbp_hash_grid_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


def device_mis(mis, device="cuda") -> _abi.MisInputs:
    """upbp_mis_inputs over device copies of a scenes.MisInputs (the copies live as long as the result)."""
    import torch

    def dev(a):
        if a is None:
            return None
        return torch.from_numpy(np.frombuffer(np.ascontiguousarray(a).tobytes(), np.uint8).copy()).to(device)

    arrays = {"queries": dev(mis.queries), "camera": dev(mis.camera), "light_vertices": dev(mis.light_vertices),
              "path_ends": dev(np.asarray(mis.path_ends, np.int32)), "beams": dev(mis.beams)}
    return _abi.mis_inputs(mis, _abi.MEM_DEVICE, arrays)


def _out_like(queries: QueryBatchView, given):
    if given is not None:
        return given
    if queries.device:
        import torch

        return torch.empty((queries.count, 3), dtype=torch.float32, device=queries._keep[0].device)
    return np.zeros((queries.count, 3), np.float32)


def eval_batch(ctx: Context, bre: Optional[EmbreeBre], beams: Optional[PhotonBeamsEvaluator],
               queries: QueryBatchView, out_photons=None, out_beams=None, stats: Optional[dict] = None):
    """upbp_eval: returns (photon radiance [Q,3] or None, beam radiance [Q,3] or None)."""
    op = _out_like(queries, out_photons) if bre is not None else None
    ob = _out_like(queries, out_beams) if beams is not None else None
    dev = (op is not None and _is_torch(op)) or (ob is not None and _is_torch(ob))
    st = _abi.EvalStats()
    ctx.check(ctx.lib.upbp_eval(ctx.h, bre.h if bre else None, beams.h if beams else None, C.byref(queries.qb),
                                _abi.ptr(op), _abi.ptr(ob), _abi.MEM_DEVICE if dev else _abi.MEM_HOST,
                                C.byref(st) if stats is not None else None))
    if stats is not None:
        stats.update(st.as_dict())
    return op, ob


def eval_batch_noalloc(ctx: Context, bre: Optional[EmbreeBre], beams: Optional[PhotonBeamsEvaluator],
                       queries: QueryBatchView, out_photons, out_beams, stats: Optional[dict] = None):
    """upbp_eval_noalloc: upbp_eval into caller-provided device outputs ([Q,3] float32 tensors,
    None where the structure is None) with device-resident queries; the library allocates nothing."""
    st = _abi.EvalStats()
    ctx.check(ctx.lib.upbp_eval_noalloc(ctx.h, bre.h if bre else None, beams.h if beams else None,
                                        C.byref(queries.qb), _abi.ptr(out_photons), _abi.ptr(out_beams),
                                        C.byref(st) if stats is not None else None))
    if stats is not None:
        stats.update(st.as_dict())
    return out_photons, out_beams


def eval_upbp(ctx: Context, bre: Optional[EmbreeBre], beams: Optional[PhotonBeamsEvaluator],
              queries: QueryBatchView, mis, stats: Optional[dict] = None, out_photons=None, out_beams=None,
              photons_only: bool = False):
    """upbp_eval_upbp: UPBP-weighted estimates. `mis` is a scenes.MisInputs (host arrays) or an
    _abi.MisInputs already describing its memory (device_mis). photons_only: the beam structure is
    not evaluated (its output is None) and serves only as the photon-beam evaluator whose cell
    selection pdfs the PB2D hits read."""
    ms = mis if isinstance(mis, _abi.MisInputs) else _abi.mis_inputs(mis)
    op = _out_like(queries, out_photons) if bre is not None else None
    ob = _out_like(queries, out_beams) if beams is not None and not photons_only else None
    dev = (op is not None and _is_torch(op)) or (ob is not None and _is_torch(ob))
    st = _abi.EvalStats()
    ctx.check(ctx.lib.upbp_eval_upbp(ctx.h, bre.h if bre else None, beams.h if beams else None, C.byref(queries.qb),
                                     C.byref(ms), _abi.ptr(op), _abi.ptr(ob), _abi.MEM_DEVICE if dev else _abi.MEM_HOST,
                                     C.byref(st) if stats is not None else None))
    if stats is not None:
        stats.update(st.as_dict())
    return op, ob


def hit_signatures(ctx: Context, bre: Optional[EmbreeBre], beams: Optional[PhotonBeamsEvaluator],
                   queries: QueryBatchView):
    """Per-query (hit count, hash sum) for photons and beams, as host arrays."""
    sp = np.zeros((queries.count, 2), np.uint32) if bre is not None else None
    sb = np.zeros((queries.count, 2), np.uint32) if beams is not None else None
    ctx.check(ctx.lib.upbp_hit_signatures(ctx.h, bre.h if bre else None, beams.h if beams else None,
                                          C.byref(queries.qb), _abi.ptr(sp), _abi.ptr(sb), _abi.MEM_HOST))
    return sp, sb


def adjoint(ctx: Context, bre, beams, queries: QueryBatchView, grad_out_photons=None, grad_out_beams=None,
            grad_scattering=None, grad_absorption=None, grad_photon_power=None, grad_beam_power=None,
            stats: Optional[dict] = None):
    """upbp_adjoint. Gradients accumulate into the given arrays (same memory kind as the upstream ones)."""
    arrs = [grad_out_photons, grad_out_beams, grad_scattering, grad_absorption, grad_photon_power, grad_beam_power]
    dev = any(a is not None and _is_torch(a) and a.is_cuda for a in arrs)
    st = _abi.EvalStats()
    ctx.check(ctx.lib.upbp_adjoint(ctx.h, bre.h if bre else None, beams.h if beams else None, C.byref(queries.qb),
                                   *[_abi.ptr(a) for a in arrs], _abi.MEM_DEVICE if dev else _abi.MEM_HOST,
                                   C.byref(st) if stats is not None else None))
    if stats is not None:
        stats.update(st.as_dict())


class BatchPlan:
    """upbp_batch_plan over several independent instances (device-resident queries)."""

    def __init__(self, ctx: Context, instances):
        """instances: list of (EmbreeBre|None, PhotonBeamsEvaluator|None, QueryBatchView, media_base,
        out_photons tensor|None, out_beams tensor|None)."""
        self.ctx = ctx
        arr = (_abi.Instance * len(instances))()
        self._keep = []
        for i, (b, m, q, base, op, ob) in enumerate(instances):
            arr[i].bre = b.h if b else None
            arr[i].beams = m.h if m else None
            arr[i].queries = q.qb
            arr[i].media_base = base
            arr[i].out_photons = _abi.ptr(op)
            arr[i].out_beams = _abi.ptr(ob)
            self._keep.append((b, m, q, op, ob))
        h = C.c_void_p()
        ctx.check(ctx.lib.upbp_batch_plan_create(ctx.h, arr, len(instances), C.byref(h)))
        self.h = h

    def run(self, use_graph: bool = True, stats: Optional[dict] = None):
        st = _abi.EvalStats()
        self.ctx.check(self.ctx.lib.upbp_batch_run(self.ctx.h, self.h, 1 if use_graph else 0,
                                                   C.byref(st) if stats is not None else None))
        if stats is not None:
            stats.update(st.as_dict())

    def destroy(self):
        if getattr(self, "h", None):
            self.ctx.lib.upbp_batch_plan_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


def autotune(ctx: Context, bre, beams, queries: QueryBatchView):
    t = _abi.Tuning()
    ms = C.c_float()
    ctx.check(ctx.lib.upbp_autotune(ctx.h, bre.h if bre else None, beams.h if beams else None,
                                    C.byref(queries.qb), C.byref(t), C.byref(ms)))
    return (t.waves_per_block, t.blocks_per_cu), ms.value
