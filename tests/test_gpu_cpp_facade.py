"""The drop-in run from C++: tests/cpp/facade_golden (built by __graft_entry__.build()) includes
only include/upbp.hpp, links libupbp_hip.so and evaluates committed golden fixtures through
upbp::EmbreeBre and upbp::PhotonBeamsEvaluator, the facade classes with the reference's method
shapes (SmallUPBP/src/Bre/Bre.hxx:40-184, SmallUPBP/src/Beams/PhBeams.hxx:46-160). Its outputs,
batched and one query per call, must equal the stored oracle output to REL_TOL and its hit
counts the stored signatures' counts.
"""
import os
import subprocess

import numpy as np
import pytest

from gpu_helpers import REL_TOL, print_case, rel_err

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "tests", "cpp", "facade_golden")
SCENE_CONST_RADIUS = ("bre_cube_24", "bre_short_hg_16")  # photon fixtures with one radius (the reference's build)


def _dump(z, path, photons: bool, beams: bool):
    with open(path, "wb") as f:
        media = np.ascontiguousarray(z["media"])
        f.write(np.int64(len(media)).tobytes() + media.tobytes())
        if photons:
            n = len(z["ph_position"])
            f.write(np.int64(n).tobytes())
            for k in ("ph_position", "ph_inc_dir", "ph_power"):
                f.write(np.ascontiguousarray(z[k], np.float32).tobytes())
            f.write(np.float32(z["ph_const_radius"]).tobytes())
        else:
            f.write(np.int64(0).tobytes())
        if beams:
            n = len(z["bm_origin"])
            f.write(np.int64(n).tobytes())
            for k, t in (("bm_origin", np.float32), ("bm_direction", np.float32), ("bm_length", np.float32),
                         ("bm_throughput", np.float32), ("bm_flags", np.uint32), ("bm_medium", np.int32)):
                f.write(np.ascontiguousarray(z[k], t).tobytes())
            f.write(np.float32(z["bm_radius"]).tobytes())
        else:
            f.write(np.int64(0).tobytes())
        q = len(z["q_origin"])
        f.write(np.int64(q).tobytes() + np.uint32(int(z["q_beam_type"])).tobytes())
        for k, t in (("q_origin", np.float32), ("q_direction", np.float32), ("q_offsets", np.int64)):
            f.write(np.ascontiguousarray(z[k], t).tobytes())
        f.write(np.ascontiguousarray(z["q_segments"]).tobytes())


def _read(path):
    b = open(path, "rb").read()
    Q, K, kept = np.frombuffer(b, np.int64, 3)
    hp, hb = np.frombuffer(b, np.uint64, 2, 24)
    o = 40
    arrs = []
    for n in (Q, Q, K, K):
        arrs.append(np.frombuffer(b, np.float32, 3 * n, o).reshape(-1, 3))
        o += 12 * n
    pdf = float(np.frombuffer(b, np.float32, 1, o)[0])
    return int(kept), int(hp), int(hb), arrs, pdf


@pytest.mark.parametrize("name", ["bre_cube_24", "bre_short_hg_16", "bre_beams_nested_16x9", "beams_mix_400"])
def test_cpp_caller_reproduces_golden_fixture(name, tmp_path):
    from golden import fixtures

    print_case("cpp_facade", name=name)
    assert os.path.exists(EXE), "tests/cpp/facade_golden is not built (run __graft_entry__.build())"
    _, z = fixtures.load(name)
    photons = "ph_position" in z and name in SCENE_CONST_RADIUS
    beams = "bm_origin" in z
    inp, out = tmp_path / "in.bin", tmp_path / "out.bin"
    _dump(z, inp, photons, beams)
    r = subprocess.run([EXE, str(inp), str(out)], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0, r.stderr
    kept, hp, hb, (bp, bb, sp, sb), pdf = _read(out)
    K = len(sp)
    if photons:
        assert kept == len(z["ph_position"])
        assert hp == int(z["sig_photons"][:, 0].astype(np.int64).sum())
        assert rel_err(bp, z["out_photons_f32"]).max() <= REL_TOL
        assert rel_err(sp, z["out_photons_f32"][:K]).max() <= REL_TOL
    if beams:
        assert hb == int(z["sig_beams"][:, 0].astype(np.int64).sum())
        assert rel_err(bb, z["out_beams_f32"]).max() <= REL_TOL
        assert rel_err(sb, z["out_beams_f32"][:K]).max() <= REL_TOL
        assert pdf == 1.0  # no per-cell reduction: getBeamSelectionPdf is 1
