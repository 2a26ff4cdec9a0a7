"""CPU tests of the boundary and host logic (no GPU compute calls).

- libupbp_hip.so loads and exports every function declared in include/upbp.h;
- a context cannot be created without a gfx950 device (no CPU fallback path exists);
- ctypes mirrors match the C struct layouts; the synthetic generators keep their invariants.
"""
import ctypes as C
import os
import re

import numpy as np
import pytest

from smallupbp_amd import _abi, native, scenes

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(ROOT, "include", "upbp.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(upbp_[a-z_0-9]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    if not os.path.exists(native.LIB_PATH):
        pytest.skip("libupbp_hip.so not built")
    lib = C.CDLL(native.LIB_PATH)
    names = declared_functions()
    assert len(names) >= 25
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert sorted(native.EXPORTED) == names


def test_no_device_means_loud_failure():
    if not os.path.exists(native.LIB_PATH):
        pytest.skip("libupbp_hip.so not built")
    import torch

    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    with pytest.raises(native.UpbpError) as e:
        native.Context(0)
    assert e.value.status == _abi.UPBP_EDEVICE


def test_struct_layouts():
    assert C.sizeof(_abi.Medium) == 32 == scenes.MEDIUM_DTYPE.itemsize
    assert C.sizeof(_abi.VolumeSegment) == 32 == scenes.SEGMENT_DTYPE.itemsize
    assert C.sizeof(_abi.QueryBatch) == 8 + 4 * 8 + 4 * 4
    assert C.sizeof(_abi.EvalStats) == 5 * 8 + 3 * 4 + 4
    # constants equal to the reference enums (Defs.hxx:146-194, Media.hxx:45-49)
    assert (_abi.SHORT_BEAM, _abi.LONG_BEAM) == (1, 2)
    assert (_abi.BPT, _abi.SURF, _abi.PP3D, _abi.PB2D, _abi.BB1D) == (4, 8, 16, 32, 64)
    assert (_abi.NO_SINE_IN_WEIGHTS, _abi.PREVIOUS, _abi.COMPATIBLE, _abi.BB1D_PREVIOUS) == (128, 256, 512, 1024)
    assert (_abi.ORIGIN_IN_MEDIUM, _abi.END_IN_MEDIUM) == (1, 2)


def test_segments_cover_rays_and_csr_is_consistent():
    w = scenes.bre_beams(64, 36, 100, 100, seed=3)
    q = w.queries
    off = q.seg_offsets
    assert off[0] == 0 and np.all(np.diff(off) >= 1) and off[-1] == len(q.segments)
    s = q.segments
    # consecutive segments of a query touch, and dist_min <= dist_max
    assert np.all(s["dist_min"] <= s["dist_max"])
    for i in range(0, q.count, 97):
        seg = s[off[i]:off[i + 1]]
        assert np.all(seg["dist_max"][:-1] == seg["dist_min"][1:])
    # attenuation = exp(-sigma_t * length) in float32
    st = (w.media["absorption"] + w.media["scattering"])[s["medium_id"]]
    assert np.allclose(s["attenuation"], np.exp(-st * (s["dist_max"] - s["dist_min"])[:, None]), rtol=1e-6)
    # rays that cross the sphere carry box / sphere / box after the leading vacuum segment
    counts = np.diff(off)
    assert set(np.unique(counts)) <= {2, 4}
    four = np.where(counts == 4)[0][0]
    assert list(s[off[four]:off[four + 1]]["medium_id"]) == [1, 0, 2, 0]


def test_subset_and_beams_stay_in_their_media():
    w = scenes.bre_beams(32, 18, 10, 5000, seed=4)
    sub = w.queries.subset(np.array([5, 3, 200]))
    assert sub.count == 3 and sub.seg_offsets[-1] == len(sub.segments)
    b = w.beams
    end = b.origin + b.direction * b.length[:, None]
    c = np.array(scenes.SPHERE[0])
    inside0 = np.linalg.norm(b.origin - c, axis=1) < scenes.SPHERE[1]
    inside1 = np.linalg.norm(end - c, axis=1) < scenes.SPHERE[1] + 1e-4
    assert np.all(inside1[inside0])  # a beam born in the sphere ends at or before its boundary
    assert np.all((end > -1e-4) & (end < 1 + 1e-4))
    assert np.all(b.medium_id[inside0] == 2)


def test_generators_are_deterministic():
    a = scenes.plain_bre(16, 16, 1000, 0.01, seed=9)
    b = scenes.plain_bre(16, 16, 1000, 0.01, seed=9)
    assert np.array_equal(a.photons.position, b.photons.position)
    assert np.array_equal(a.queries.segments, b.queries.segments)
    c = scenes.plain_bre(16, 16, 1000, 0.01, seed=10)
    assert not np.array_equal(a.photons.position, c.photons.position)


def test_cpp_facade_compiles(tmp_path):
    """include/upbp.hpp (the C++ facade a renderer includes) compiles against upbp.h."""
    import shutil
    import subprocess

    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not available")
    src = tmp_path / "use.cpp"
    src.write_text('#include "upbp.hpp"\n'
                   "int use(upbp::Context& c, const float* p, const float* w) {\n"
                   "  upbp::EmbreeBre bre(c);\n"
                   "  int kept = bre.build(p, nullptr, w, nullptr, 4, UPBP_CONSTANT_RADIUS, 0.01f, 0);\n"
                   "  upbp_volume_segment s{0.f, 1.f, 1.f, 1.f, {1.f, 1.f, 1.f}, 0};\n"
                   "  float o[3] = {0, 0, 0}, d[3] = {0, 0, 1};\n"
                   "  auto rgb = bre.evalBre(UPBP_LONG_BEAM, o, d, &s, 1);\n"
                   "  upbp::QueryBatch qb; qb.add(o, d, &s, 1);\n"
                   "  upbp::PhotonBeamsEvaluator pb(c);\n"
                   "  return kept + (int)rgb[0] + (int)qb.size() + (int)pb.getBeamSelectionPdf(o);\n"
                   "}\n")
    r = subprocess.run([gxx, "-std=c++17", "-fsyntax-only", "-Wall", "-Werror", "-I",
                        os.path.join(ROOT, "include"), str(src)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    c = subprocess.run(["gcc", "-std=c99", "-fsyntax-only", "-Wall", "-Werror", "-x", "c",
                        os.path.join(ROOT, "include", "upbp.h")], capture_output=True, text=True)
    assert c.returncode == 0, c.stderr
