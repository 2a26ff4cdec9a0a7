"""ctypes front end of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE: only tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg import this module, and only as the checker / CPU baseline.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

from smallupbp_amd import _abi
from smallupbp_amd.scenes import Beams, Photons, Queries

HERE = os.path.dirname(os.path.abspath(__file__))
# UPBP_ORACLE_LIB selects another build of the same sources (the sanitized one of `make asan`)
LIB_PATH = os.environ.get("UPBP_ORACLE_LIB") or os.path.join(HERE, "liboracle.so")
_lib = None


def _threads() -> int:
    """Worker threads for the oracle: the CPUs this process may run on, at most 16."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(n, 16))


def build(force: bool = False) -> str:
    if force or not os.path.exists(LIB_PATH):
        subprocess.check_call(["make", "-s", "-C", HERE] + (["-B"] if force else []))
    return LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        vp, i64, i32, f32 = C.c_void_p, C.c_int64, C.c_int32, C.c_float
        L.or_photons_create.restype = vp
        L.or_photons_create.argtypes = [i64, vp, vp, vp, vp, f32, vp, i32]
        L.or_photons_free.argtypes = [vp]
        L.or_photons_count.restype = i64
        L.or_photons_count.argtypes = [vp]
        L.or_photons_orig_index.argtypes = [vp, vp]
        L.or_bre_eval.restype = C.c_int
        L.or_bre_eval.argtypes = [vp, vp, i32, C.POINTER(_abi.QueryBatch), C.c_int, C.c_int, C.c_int,
                                  vp, vp, vp, vp]
        L.or_beams_create.restype = vp
        L.or_beams_create.argtypes = [i64, vp, vp, vp, vp, vp, vp, vp, vp, f32, i32]
        L.or_beams_free.argtypes = [vp]
        L.or_beams_count.restype = i64
        L.or_beams_count.argtypes = [vp]
        L.or_beams_grid_info.argtypes = [vp, vp]
        L.or_beam_grid_stats.restype = C.c_int
        L.or_bre_eval_brute_force.restype = C.c_int
        L.or_bre_eval_brute_force.argtypes = [vp, vp, i32, vp, vp, vp, vp, vp, i64, C.c_int, vp]
        L.or_beams_grid_geometry.argtypes = [vp, vp]
        L.or_beams_grid_cell_counts.restype = C.c_int
        L.or_beams_grid_cell_counts.argtypes = [vp, vp, i64]
        L.or_beam_grid_stats.argtypes = [vp, vp, i32, C.POINTER(_abi.QueryBatch), C.c_int, vp]
        L.or_beams_set_reduction.restype = C.c_int
        L.or_beams_set_reduction.argtypes = [vp, C.c_uint32, C.c_uint32, i32]
        L.or_beams_selection_pdf.restype = f32
        L.or_beams_selection_pdf.argtypes = [vp, vp]
        L.or_beams_eval.restype = C.c_int
        L.or_beams_eval.argtypes = [vp, vp, i32, C.POINTER(_abi.QueryBatch), C.c_int, C.c_int,
                                    C.c_int, vp, vp, vp, vp]
        L.or_adjoint.restype = C.c_int
        L.or_adjoint.argtypes = [vp, vp, vp, i32, C.POINTER(_abi.QueryBatch), vp, vp, C.c_int, vp, vp,
                                 vp, vp, C.c_int]
        L.or_upbp_eval.restype = C.c_int
        L.or_upbp_eval.argtypes = [vp, vp, vp, i32, C.POINTER(_abi.QueryBatch), C.POINTER(_abi.MisInputs), C.c_int,
                                   C.c_int, C.c_int, C.c_int, vp, vp, vp]
        L.or_exp_cephes.restype = f32
        L.or_trace_light_paths.restype = C.c_int
        L.or_trace_light_paths.argtypes = [vp, i32, vp, i32, vp, i32, C.POINTER(_abi.LightTraceParams), i64,
                                           C.c_int, C.POINTER(_abi.LightPaths)]
        L.or_hgrid_create.restype = vp
        L.or_hgrid_create.argtypes = [i64, vp, vp, f32, i64, C.c_uint32]
        L.or_hgrid_free.argtypes = [vp]
        L.or_hgrid_count.restype = i64
        L.or_hgrid_count.argtypes = [vp]
        L.or_hgrid_info.argtypes = [vp, vp, vp, vp]
        L.or_pp3d_eval.argtypes = [vp, vp, i32, vp, i64, C.POINTER(_abi.MisInputs), C.c_uint32, C.c_int, C.c_int,
                                   C.c_int, vp, vp, vp, vp]
        L.or_surf_eval.argtypes = [vp, vp, i64, C.POINTER(_abi.MisInputs), C.c_uint32, C.c_int, C.c_int, C.c_int,
                                   vp, vp, vp, vp]
        L.or_camera_walk.restype = C.c_double
        L.or_camera_walk.argtypes = [vp, C.c_int] + [C.c_double] * 5 + [C.c_uint32] * 3
        L.or_light_walk.restype = C.c_double
        L.or_light_walk.argtypes = [vp, i64, vp, i64, C.c_int, C.c_int, C.c_double, C.c_double] + [C.c_uint32] * 4 + [vp]
        L.or_knn_d2.argtypes = [vp, i64, vp, i64, i32, C.c_int, vp]
        L.or_knn_d2.restype = None
        L.or_photon_knn_radius.argtypes = [i64, vp, vp, f32, i32, C.c_int, vp]
        L.or_photon_knn_radius.restype = None
        L.or_beam_knn_radius.argtypes = [i64, vp, vp, vp, f32, i32, C.c_int, vp, vp]
        L.or_beam_knn_radius.restype = None
        L.or_exp_cephes.argtypes = [f32]
        L.or_hash32.restype = C.c_uint32
        L.or_hash32.argtypes = [C.c_uint32]
        L.or_phase_pdf.restype = f32
        L.or_phase_pdf.argtypes = [vp, vp, f32, vp]
        L.or_test_bre.restype = C.c_int
        L.or_test_bre.argtypes = [vp, vp, f32, f32, vp, f32, vp, vp]
        L.or_test_beam_beam.restype = C.c_int
        L.or_test_beam_beam.argtypes = [vp, vp, f32, f32, vp, vp, f32, f32, f32, vp, vp, vp, vp]
        _lib = L
    return _lib


def _c(a, dtype):
    return None if a is None else np.ascontiguousarray(a, dtype=dtype)


def query_batch(q: Queries):
    """Host-memory upbp_query_batch over q's arrays (keeps references alive on the struct)."""
    o = _c(q.origin, np.float32)
    d = _c(q.direction, np.float32)
    off = _c(q.seg_offsets, np.int64)
    seg = np.ascontiguousarray(q.segments)
    qb = _abi.QueryBatch(q.count, _abi.ptr(o), _abi.ptr(d), _abi.ptr(off), _abi.ptr(seg), q.beam_type,
                         1 if q.lite else 0, 0, _abi.MEM_HOST)
    qb._keep = (o, d, off, seg)
    return qb


class OraclePhotons:
    def __init__(self, ph: Photons, in_medium=None, leaf_size=4, build_bvh=True):
        self._arrays = (_c(ph.position, np.float32), _c(ph.inc_dir, np.float32), _c(ph.power, np.float32),
                        _c(ph.radius, np.float32), _c(in_medium, np.uint8))
        a = self._arrays
        self.h = lib().or_photons_create(ph.count, _abi.ptr(a[0]), _abi.ptr(a[1]), _abi.ptr(a[2]),
                                         _abi.ptr(a[3]), ph.const_radius, _abi.ptr(a[4]),
                                         leaf_size if build_bvh else -1)
        self.n_in = ph.count

    @property
    def count(self):
        return lib().or_photons_count(self.h)

    def __del__(self):
        if getattr(self, "h", None):
            lib().or_photons_free(self.h)
            self.h = None


class OracleBeams:
    def __init__(self, bm: Beams, grid_res=256, start_radius=None, end_radius=None):
        self._arrays = (_c(bm.origin, np.float32), _c(bm.direction, np.float32), _c(bm.length, np.float32),
                        _c(bm.throughput, np.float32), _c(bm.flags, np.uint32), _c(bm.medium_id, np.int32),
                        _c(start_radius, np.float32), _c(end_radius, np.float32))
        a = self._arrays
        self.h = lib().or_beams_create(bm.count, *[_abi.ptr(x) for x in a], bm.radius, grid_res)
        self.n = bm.count

    def set_reduction(self, max_beams_in_cell, reduction_type=0, seed=1234):
        """Per-cell beam reduction (Grid::reduceBeams): or_beams_eval then applies it, with accel 1
        per crowded cell of the DDA as the reference does, with accel 0 by the cell of each hit's
        query point."""
        r = lib().or_beams_set_reduction(self.h, max_beams_in_cell, reduction_type, seed)
        if r:
            raise RuntimeError("or_beams_set_reduction failed: %s" % _abi.STATUS_NAMES.get(r, r))
        return self

    def selection_pdf(self, pos):
        p = np.ascontiguousarray(pos, np.float32)
        return float(lib().or_beams_selection_pdf(self.h, p.ctypes.data))

    def grid_info(self):
        info = np.zeros(6, np.int64)
        lib().or_beams_grid_info(self.h, info.ctypes.data)
        return info

    def grid_geometry(self):
        """(lo [3], hi [3], cell size [3]) of the reference grid."""
        g = np.zeros(9, np.float32)
        lib().or_beams_grid_geometry(self.h, g.ctypes.data)
        return g[:3], g[3:6], g[6:]

    def grid_cell_counts(self):
        """Beams registered per cell, [res_z, res_y, res_x]."""
        res = self.grid_info()[:3]
        out = np.zeros(int(np.prod(res)), np.uint32)
        if lib().or_beams_grid_cell_counts(self.h, out.ctypes.data, out.size):
            raise RuntimeError("or_beams_grid_cell_counts failed")
        return out.reshape(int(res[2]), int(res[1]), int(res[0]))

    def __del__(self):
        if getattr(self, "h", None):
            lib().or_beams_free(self.h)
            self.h = None


def _media_ptr(media):
    m = np.ascontiguousarray(media)
    return m, m.ctypes.data, len(m)


def bre_eval(oph: OraclePhotons, media, q: Queries, accel=1, precision=32, threads=None,
             signatures=False):
    """Returns (out [Q,3] f32 or f64, sig [Q,2] u32 or None, counters [hits, candidates])."""
    threads = threads or _threads()
    qb = query_batch(q)
    m, mp, nm = _media_ptr(media)
    out = np.zeros((q.count, 3), np.float64 if precision == 64 else np.float32)
    sig = np.zeros((q.count, 2), np.uint32) if signatures else None
    cnt = np.zeros(2, np.uint64)
    r = lib().or_bre_eval(oph.h, mp, nm, C.byref(qb), accel, precision, threads,
                          out.ctypes.data if precision != 64 else None,
                          out.ctypes.data if precision == 64 else None, _abi.ptr(sig), cnt.ctypes.data)
    if r:
        raise RuntimeError("or_bre_eval failed: %s" % _abi.STATUS_NAMES.get(r, r))
    return out, sig, cnt


def bre_eval_brute_force(oph: OraclePhotons, media, origin, direction, min_t, max_t, medium_id, threads=None):
    """EmbreeBre::evalBreBruteForce (Bre.cxx:980-1029) per ray -> [n,3] f32."""
    threads = threads or _threads()
    o = np.ascontiguousarray(origin, np.float32).reshape(-1, 3)
    d = np.ascontiguousarray(direction, np.float32).reshape(-1, 3)
    n = o.shape[0]
    t0 = np.ascontiguousarray(min_t, np.float32).reshape(n)
    t1 = np.ascontiguousarray(max_t, np.float32).reshape(n)
    m = np.ascontiguousarray(medium_id, np.int32).reshape(n)
    mm, mp, nm = _media_ptr(media)
    out = np.zeros((n, 3), np.float32)
    r = lib().or_bre_eval_brute_force(oph.h, mp, nm, o.ctypes.data, d.ctypes.data, t0.ctypes.data, t1.ctypes.data,
                                      m.ctypes.data, n, threads, out.ctypes.data)
    if r:
        raise RuntimeError("or_bre_eval_brute_force failed: %s" % _abi.STATUS_NAMES.get(r, r))
    return out


def beam_grid_stats(obm: OracleBeams, media, q: Queries, threads=None):
    """GridStats per query ([Q,3] u32: beams tested, cells crossed, overfull cells crossed) of the
    reference grid's walk over the query's scattering segments; needs an OracleBeams built with
    grid_res > 0 (and set_reduction for overfull cells)."""
    threads = threads or _threads()
    qb = query_batch(q)
    m, mp, nm = _media_ptr(media)
    out = np.zeros((q.count, 3), np.uint32)
    r = lib().or_beam_grid_stats(obm.h, mp, nm, C.byref(qb), threads, out.ctypes.data)
    if r:
        raise RuntimeError("or_beam_grid_stats failed: %s" % _abi.STATUS_NAMES.get(r, r))
    return out


def beams_eval(obm: OracleBeams, media, q: Queries, accel=0, precision=32, threads=None, signatures=False):
    """accel 0 = brute force (every beam), 1 = the reference uniform grid."""
    threads = threads or _threads()
    qb = query_batch(q)
    m, mp, nm = _media_ptr(media)
    out = np.zeros((q.count, 3), np.float64 if precision == 64 else np.float32)
    sig = np.zeros((q.count, 2), np.uint32) if signatures else None
    cnt = np.zeros(2, np.uint64)
    r = lib().or_beams_eval(obm.h, mp, nm, C.byref(qb), accel, precision, threads,
                            out.ctypes.data if precision != 64 else None,
                            out.ctypes.data if precision == 64 else None, _abi.ptr(sig), cnt.ctypes.data)
    if r:
        raise RuntimeError("or_beams_eval failed: %s" % _abi.STATUS_NAMES.get(r, r))
    return out, sig, cnt


def adjoint(oph, obm, media, q: Queries, w_photons, w_beams, threads=None, fp32_hits=False):
    """FP64 analytic gradient. Returns dict of grad_sigma_s, grad_sigma_a [M,3],
    grad_photon_power [N_in,3], grad_beam_power [B,3]. fp32_hits: the hit sets (and hit
    parameters) of the reference's FP32 predicates, as the FP32 forward and the GPU have them,
    and the kernels' radial factors 1 - d^2 / r^2 in the reference's float arithmetic (near the
    radius their relative accuracy in float is about 2^-24 / (1 - d^2 / r^2)); else those of the
    FP64 forward."""
    threads = threads or _threads()
    qb = query_batch(q)
    m, mp, nm = _media_ptr(media)
    gs = np.zeros((nm, 3))
    ga = np.zeros((nm, 3))
    gp = np.zeros((oph.n_in if oph else 0, 3))
    gb = np.zeros((obm.n if obm else 0, 3))
    wp = None if w_photons is None else np.ascontiguousarray(w_photons, np.float64)
    wb = None if w_beams is None else np.ascontiguousarray(w_beams, np.float64)
    r = lib().or_adjoint(oph.h if oph else None, obm.h if obm else None, mp, nm, C.byref(qb),
                         _abi.ptr(wp), _abi.ptr(wb), threads, gs.ctypes.data, ga.ctypes.data,
                         gp.ctypes.data if oph else None, gb.ctypes.data if obm else None, 1 if fp32_hits else 0)
    if r:
        raise RuntimeError("or_adjoint failed: %s" % _abi.STATUS_NAMES.get(r, r))
    return {"grad_sigma_s": gs, "grad_sigma_a": ga, "grad_photon_power": gp, "grad_beam_power": gb}


def upbp_eval(oph, obm, media, q: Queries, mi, which: int, precision=32, threads=None, accel=0):
    """UPBP-mode estimate (MIS weights) per query; which = 0 photons, 1 beams. accel 0 = brute force
    over items (exact hit sets), 1 = photon BVH / reference beam grid. Returns (out, hits)."""
    threads = threads or _threads()
    qb = query_batch(q)
    m, mp, nm = _media_ptr(media)
    ms = _abi.mis_inputs(mi)
    out = np.zeros((q.count, 3), np.float64 if precision == 64 else np.float32)
    cnt = np.zeros(2, np.uint64)
    r = lib().or_upbp_eval(oph.h if oph else None, obm.h if obm else None, mp, nm, C.byref(qb), C.byref(ms), which,
                           accel, precision, threads, out.ctypes.data if precision != 64 else None,
                           out.ctypes.data if precision == 64 else None, cnt.ctypes.data)
    if r:
        raise RuntimeError("or_upbp_eval failed: %s" % _abi.STATUS_NAMES.get(r, r))
    return out, int(cnt[0])


class OracleHashGrid:
    """HashGrid::Build over the light vertices of a scenes.MergeScene (HashGrid.hxx:46-128): the
    in-medium vertices for PP3D, the others for SURF."""

    def __init__(self, sc):
        self.technique = int(sc.technique)
        self.pos = np.ascontiguousarray(sc.positions, np.float32)
        self.lv = np.ascontiguousarray(sc.mis.light_vertices)
        self.cell_count = int(sc.cell_count)
        self.h = lib().or_hgrid_create(len(self.pos), self.pos.ctypes.data, self.lv.ctypes.data, float(sc.radius),
                                       self.cell_count, self.technique)
        if not self.h:
            raise RuntimeError("or_hgrid_create rejected its arguments")

    def count(self):
        return int(lib().or_hgrid_count(self.h))

    def set_reduction(self, max_beams_in_cell, reduction_type=0, seed=1234):
        """Per-cell beam reduction (Grid::reduceBeams): or_beams_eval then applies it, with accel 1
        per crowded cell of the DDA as the reference does, with accel 0 by the cell of each hit's
        query point."""
        r = lib().or_beams_set_reduction(self.h, max_beams_in_cell, reduction_type, seed)
        if r:
            raise RuntimeError("or_beams_set_reduction failed: %s" % _abi.STATUS_NAMES.get(r, r))
        return self

    def selection_pdf(self, pos):
        p = np.ascontiguousarray(pos, np.float32)
        return float(lib().or_beams_selection_pdf(self.h, p.ctypes.data))

    def grid_info(self):
        """(bounding box (6,), mIndices, mCellEnds) of the built grid."""
        bb = np.zeros(6, np.float32)
        idx = np.zeros(max(self.count(), 1), np.int32)
        ends = np.zeros(self.cell_count, np.int32)
        lib().or_hgrid_info(self.h, bb.ctypes.data, idx.ctypes.data, ends.ctypes.data)
        return bb, idx[:self.count()], ends

    def __del__(self):
        if getattr(self, "h", None):
            lib().or_hgrid_free(self.h)
            self.h = None


def vm_eval(og: OracleHashGrid, sc, flags=0, accel=1, precision=32, threads=None, queries=None):
    """RangeQuery::GetContrib per camera vertex (UPBP.hxx:84-129), PP3D or SURF by the scene's
    technique. Returns (out, sig, counters)."""
    threads = threads or _threads()
    q = np.ascontiguousarray(sc.queries if queries is None else queries)
    ms = _abi.mis_inputs(sc.mis)
    out = np.zeros((len(q), 3), np.float64 if precision == 64 else np.float32)
    sig = np.zeros((len(q), 2), np.uint32)
    cnt = np.zeros(2, np.uint64)
    o32 = out.ctypes.data if precision != 64 else None
    o64 = out.ctypes.data if precision == 64 else None
    if sc.technique == _abi.PP3D:
        m, mp, nm = _media_ptr(sc.media)
        r = lib().or_pp3d_eval(og.h, mp, nm, q.ctypes.data, len(q), C.byref(ms), int(flags), int(accel),
                               int(precision), threads, o32, o64, sig.ctypes.data, cnt.ctypes.data)
    else:
        r = lib().or_surf_eval(og.h, q.ctypes.data, len(q), C.byref(ms), int(flags), int(accel), int(precision),
                               threads, o32, o64, sig.ctypes.data, cnt.ctypes.data)
    if r:
        raise RuntimeError("vertex merging oracle failed: %s" % _abi.STATUS_NAMES.get(r, r))
    return out, sig, cnt


def camera_walk(cam, length, last_rev, ntl, tech, last_sin=0.0, last_ray_rev_inv=0.0, last_ray_rev_ratio=0.0,
                qbt=_abi.LONG_BEAM, pbt=_abi.SHORT_BEAM) -> float:
    """AccumulateCameraPathWeight (PathWeight.hxx:33-142) over cam[0 .. length], FP64."""
    cam = np.ascontiguousarray(cam)
    return lib().or_camera_walk(cam.ctypes.data, int(length), float(last_rev), float(last_sin),
                                float(last_ray_rev_inv), float(last_ray_rev_ratio), float(ntl), qbt, pbt, int(tech))


def light_walk(lv, path_ends, path_idx, path_length, last_rev, ntl, evaluated, tech, qbt=_abi.LONG_BEAM,
               pbt=_abi.SHORT_BEAM) -> float:
    """AccumulateLightPathWeight (PathWeight.hxx:145-270, camera connection false), FP64."""
    lv = np.ascontiguousarray(lv)
    ends = np.ascontiguousarray(path_ends, np.int32)
    err = C.c_int(0)
    w = lib().or_light_walk(lv.ctypes.data, len(lv), ends.ctypes.data, len(ends), int(path_idx), int(path_length),
                            float(last_rev), float(ntl), int(evaluated), qbt, pbt, int(tech), C.byref(err))
    if err.value:
        raise RuntimeError("light walk read outside the light vertices")
    return w


def hash32(i: np.ndarray) -> np.ndarray:
    """Vectorised copy of or_hash32 (lowbias32 of i + 1)."""
    x = (np.asarray(i, np.uint64) + 1) & 0xFFFFFFFF
    x ^= x >> 16
    x = (x * 0x7FEB352D) & 0xFFFFFFFF
    x ^= x >> 15
    x = (x * 0x846CA68B) & 0xFFFFFFFF
    x ^= x >> 16
    return x.astype(np.uint32)


def knn_d2(points: np.ndarray, queries: np.ndarray, k: int, threads=None) -> np.ndarray:
    """k-th smallest squared distance from each query to the points (KdTmpl.hxx:600-792 semantics)."""
    p = np.ascontiguousarray(points, np.float32).reshape(-1, 3)
    q = np.ascontiguousarray(queries, np.float32).reshape(-1, 3)
    out = np.zeros(len(q), np.float32)
    lib().or_knn_d2(p.ctypes.data, len(p), q.ctypes.data, len(q), int(k), threads or _threads(), out.ctypes.data)
    return out


def photon_knn_radius(position, radius: float, knn: int, in_medium=None, threads=None) -> np.ndarray:
    """Per-photon radius of EmbreeBre::build with KNN_RADIUS (Bre.cxx:562-602); 0 for dropped photons."""
    p = np.ascontiguousarray(position, np.float32).reshape(-1, 3)
    m = None if in_medium is None else np.ascontiguousarray(in_medium, np.uint8)
    out = np.zeros(len(p), np.float32)
    lib().or_photon_knn_radius(len(p), p.ctypes.data, None if m is None else m.ctypes.data, float(radius), int(knn),
                               threads or _threads(), out.ctypes.data)
    return out


def beam_knn_radius(bm: Beams, radius: float, knn: int, threads=None):
    """(start, end) radius per beam of PhotonBeamsEvaluator::build with KNN_RADIUS (PhBeams.cxx:77-117)."""
    o = np.ascontiguousarray(bm.origin, np.float32)
    d = np.ascontiguousarray(bm.direction, np.float32)
    ln = np.ascontiguousarray(bm.length, np.float32)
    r0 = np.zeros(bm.count, np.float32)
    r1 = np.zeros(bm.count, np.float32)
    lib().or_beam_knn_radius(bm.count, o.ctypes.data, d.ctypes.data, ln.ctypes.data, float(radius), int(knn),
                             threads or _threads(), r0.ctypes.data, r1.ctypes.data)
    return r0, r1


def light_trace_params(scene) -> _abi.LightTraceParams:
    p = _abi.LightTraceParams()
    for k, v in scene.params.items():
        setattr(p, k, v)
    return p


def trace_light_paths(scene, path_count: int, threads=None, margins=False):
    """upbp_trace_light_paths restated (or_trace_light_paths): dict of numpy arrays trimmed to
    the produced vertices and beams. margins=True adds "margins" [path_count] float32: per path
    the closest call of its branch decisions (relative distance of the compared values)."""
    from smallupbp_amd import scenes

    threads = threads or _threads()
    cap_v, cap_b = scenes.light_path_capacity(scene, path_count)
    arr = scenes.light_path_arrays(cap_v, cap_b, path_count)
    lp = _abi.light_paths(arr, cap_v, cap_b)
    media = np.ascontiguousarray(scene.media)
    regions = np.ascontiguousarray(scene.regions)
    lights = np.ascontiguousarray(scene.lights)
    par = light_trace_params(scene)
    args = (media.ctypes.data, len(media), regions.ctypes.data if len(regions) else None, len(regions),
            lights.ctypes.data, len(lights), C.byref(par), path_count, threads, C.byref(lp))
    mg = None
    if margins:
        mg = np.empty(max(path_count, 1), np.float32)
        f = lib().or_trace_light_paths_margins
        f.restype = C.c_int
        f.argtypes = [C.c_void_p, C.c_int32, C.c_void_p, C.c_int32, C.c_void_p, C.c_int32,
                      C.POINTER(_abi.LightTraceParams), C.c_int64, C.c_int, C.POINTER(_abi.LightPaths), C.c_void_p]
        r = f(*args, mg.ctypes.data)
    else:
        r = lib().or_trace_light_paths(*args)
    if r:
        raise RuntimeError("or_trace_light_paths failed: %s" % _abi.STATUS_NAMES.get(r, r))
    out = scenes.trim_light_paths(arr, lp.vertex_count, lp.beam_count, path_count)
    if margins:
        out["margins"] = mg[:path_count]
    return out
