"""ctypes mirror of include/upbp.h (structs and constants only; no library loading)."""
from __future__ import annotations

import ctypes as C

import numpy as np

UPBP_OK = 0
UPBP_EINVAL = 1
UPBP_ENOMEM = 2
UPBP_EDEVICE = 3
UPBP_ENONFINITE = 4
UPBP_EUNSUPPORTED = 5
UPBP_ENOTBUILT = 6
STATUS_NAMES = {
    0: "UPBP_OK", 1: "UPBP_EINVAL", 2: "UPBP_ENOMEM", 3: "UPBP_EDEVICE",
    4: "UPBP_ENONFINITE", 5: "UPBP_EUNSUPPORTED", 6: "UPBP_ENOTBUILT",
}

SHORT_BEAM = 1
LONG_BEAM = 2
REDUCE_PRESAMPLE, REDUCE_OFFSET, REDUCE_RESAMPLE_FIXED, REDUCE_RESAMPLE = 0, 1, 2, 3
BPT, SURF, PP3D, PB2D, BB1D = 4, 8, 16, 32, 64
NO_SINE_IN_WEIGHTS, PREVIOUS, COMPATIBLE, BB1D_PREVIOUS, SPECULAR_ONLY = 128, 256, 512, 1024, 2048
ORIGIN_IN_MEDIUM, END_IN_MEDIUM = 1, 2
CONSTANT_RADIUS, KNN_RADIUS = 0, 1
MEM_HOST, MEM_DEVICE = 0, 1
ORDER_SORTED, ORDER_AS_GIVEN = 0, 1

c_float_p = C.POINTER(C.c_float)


class Medium(C.Structure):
    _fields_ = [("absorption", C.c_float * 3), ("scattering", C.c_float * 3),
                ("mean_cosine", C.c_float), ("continuation_prob", C.c_float)]


class VolumeSegment(C.Structure):
    _fields_ = [("dist_min", C.c_float), ("dist_max", C.c_float), ("ray_sample_pdf", C.c_float),
                ("ray_sample_rev_pdf", C.c_float), ("attenuation", C.c_float * 3),
                ("medium_id", C.c_int32)]


class BuildParams(C.Structure):
    _fields_ = [("radius_calculation", C.c_int32), ("radius", C.c_float), ("knn", C.c_int32),
                ("leaf_size", C.c_int32), ("beam_piece_factor", C.c_float),
                ("max_pieces_per_beam", C.c_int32), ("max_beams_in_cell", C.c_uint32),
                ("reduction_type", C.c_uint32), ("grid_resolution", C.c_int32), ("reduction_seed", C.c_int32)]


class PhotonInput(C.Structure):
    _fields_ = [("count", C.c_int64), ("position", C.c_void_p), ("inc_dir", C.c_void_p),
                ("power", C.c_void_p), ("radius", C.c_void_p), ("in_medium", C.c_void_p),
                ("memory", C.c_int)]


class BeamInput(C.Structure):
    _fields_ = [("count", C.c_int64), ("origin", C.c_void_p), ("direction", C.c_void_p),
                ("length", C.c_void_p), ("throughput", C.c_void_p), ("flags", C.c_void_p),
                ("medium_id", C.c_void_p), ("start_radius", C.c_void_p), ("end_radius", C.c_void_p),
                ("memory", C.c_int)]


class QueryBatch(C.Structure):
    _fields_ = [("count", C.c_int64), ("origin", C.c_void_p), ("direction", C.c_void_p),
                ("segment_offsets", C.c_void_p), ("segments", C.c_void_p),
                ("beam_type", C.c_uint32), ("lite_segments", C.c_uint32), ("order", C.c_uint32),
                ("memory", C.c_int)]


class EvalStats(C.Structure):
    _fields_ = [("photon_hits", C.c_uint64), ("beam_hits", C.c_uint64),
                ("photon_candidates", C.c_uint64), ("beam_candidates", C.c_uint64),
                ("nodes_visited", C.c_uint64), ("kernel_ms_photons", C.c_float),
                ("kernel_ms_beams", C.c_float), ("kernel_ms_other", C.c_float),
                ("launches", C.c_uint32)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class Instance(C.Structure):
    _fields_ = [("bre", C.c_void_p), ("beams", C.c_void_p), ("queries", QueryBatch),
                ("media_base", C.c_int32), ("out_photons", C.c_void_p), ("out_beams", C.c_void_p)]


class Tuning(C.Structure):
    _fields_ = [("waves_per_block", C.c_int32), ("blocks_per_cu", C.c_int32),
                ("photon_waves_per_block", C.c_int32), ("frustum_leaf", C.c_int32)]


assert C.sizeof(Medium) == 32
assert C.sizeof(VolumeSegment) == 32


def ptr(a) -> int:
    """Address of a numpy array or a torch tensor (None -> 0)."""
    if a is None:
        return 0
    if isinstance(a, np.ndarray):
        assert a.flags["C_CONTIGUOUS"], "arrays passed to the C ABI must be C-contiguous"
        return a.ctypes.data
    return int(a.data_ptr())


def media_array(media: np.ndarray):
    arr = (Medium * len(media))()
    C.memmove(arr, np.ascontiguousarray(media).ctypes.data, 32 * len(media))
    return arr


class MisParams(C.Structure):
    _fields_ = [("techniques", C.c_uint32), ("query_beam_type", C.c_uint32), ("photon_beam_type", C.c_uint32),
                ("min_path_length", C.c_uint32), ("max_path_length", C.c_uint32), ("surf_factor", C.c_float),
                ("pp3d_factor", C.c_float), ("pb2d_factor", C.c_float), ("bb1d_factor", C.c_float),
                ("bb1d_min_mfp", C.c_float), ("min_dist_to_med", C.c_float), ("has_bb1d_beams", C.c_int32)]


class Region(C.Structure):
    _fields_ = [("shape", C.c_uint32), ("medium_id", C.c_int32), ("p0", C.c_float * 3), ("p1", C.c_float * 3)]


class PointLight(C.Structure):
    _fields_ = [("position", C.c_float * 3), ("intensity", C.c_float * 3)]


class LightTraceParams(C.Structure):
    _fields_ = [("seed", C.c_uint64), ("photon_beam_type", C.c_uint32), ("max_path_length", C.c_uint32),
                ("global_medium", C.c_int32), ("beam_path_count", C.c_int32), ("merge_pb2d", C.c_int32),
                ("surf_factor", C.c_float), ("pp3d_factor", C.c_float), ("pb2d_factor", C.c_float),
                ("bb1d_factor", C.c_float), ("bb1d_min_mfp", C.c_float), ("pad", C.c_uint32)]


class LightPaths(C.Structure):
    _fields_ = [("vertices", C.c_void_p), ("positions", C.c_void_p), ("vertex_capacity", C.c_int64),
                ("path_ends", C.c_void_p), ("beam_origin", C.c_void_p), ("beam_direction", C.c_void_p),
                ("beam_length", C.c_void_p), ("beam_throughput", C.c_void_p), ("beam_flags", C.c_void_p),
                ("beam_medium", C.c_void_p), ("beam_mis", C.c_void_p), ("beam_capacity", C.c_int64),
                ("vertex_count", C.c_int64), ("beam_count", C.c_int64)]


LIGHT_PATH_FIELDS = ("vertices", "positions", "path_ends", "beam_origin", "beam_direction", "beam_length",
                     "beam_throughput", "beam_flags", "beam_medium", "beam_mis")


def light_paths(arrays: dict, vertex_capacity: int, beam_capacity: int) -> LightPaths:
    """upbp_light_paths over output arrays (numpy or torch, one memory kind)."""
    lp = LightPaths()
    for f in LIGHT_PATH_FIELDS:
        setattr(lp, f, ptr(arrays[f]))
    lp.vertex_capacity = vertex_capacity
    lp.beam_capacity = beam_capacity
    return lp


class MisInputs(C.Structure):
    _fields_ = [("queries", C.c_void_p), ("camera", C.c_void_p), ("camera_count", C.c_int64),
                ("light_vertices", C.c_void_p), ("light_vertex_count", C.c_int64), ("path_ends", C.c_void_p),
                ("path_count", C.c_int64), ("beams", C.c_void_p), ("params", MisParams), ("memory", C.c_int)]


def mis_inputs(mi, memory=MEM_HOST, arrays=None):
    """upbp_mis_inputs over a scenes.MisInputs (host numpy) or over `arrays` (dict of torch tensors)."""
    src = arrays if arrays is not None else {
        "queries": None if mi.queries is None else np.ascontiguousarray(mi.queries),
        "camera": np.ascontiguousarray(mi.camera),
        "light_vertices": np.ascontiguousarray(mi.light_vertices),
        "path_ends": np.ascontiguousarray(mi.path_ends, np.int32),
        "beams": None if mi.beams is None else np.ascontiguousarray(mi.beams)}
    p = MisParams(**mi.params)
    s = MisInputs(ptr(src["queries"]), ptr(src["camera"]), len(mi.camera), ptr(src["light_vertices"]),
                  len(mi.light_vertices), ptr(src["path_ends"]), len(mi.path_ends), ptr(src["beams"]), p, memory)
    s._keep = src
    return s
