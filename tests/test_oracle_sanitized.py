"""The oracle under AddressSanitizer and UndefinedBehaviorSanitizer.

`make -C oracle asan` builds the same sources with -fsanitize=address,undefined (host code;
the oracle has no GPU code). A child Python process preloads libasan, loads that build
through UPBP_ORACLE_LIB and runs every oracle entry point on small inputs: photon and beam
evaluation (brute force, BVH and grid, FP32 and FP64, with and without per-cell reduction),
the UPBP weights, the adjoint, the hashed-grid merging, k-NN radii and light tracing, plus
malformed inputs that must be rejected (offsets out of range, medium ids outside the table).
Any out-of-bounds access or undefined behaviour aborts the child."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

CHILD = r'''
import numpy as np
from oracle import pyoracle as po
from smallupbp_amd import scenes

w = scenes.bre_beams(24, 14, 3000, 400, seed=3)
oph = po.OraclePhotons(w.photons)
for accel in (0, 1):
    for prec in (32, 64):
        po.bre_eval(oph, w.media, w.queries, accel=accel, precision=prec, signatures=prec == 32)
for res, accel in ((0, 0), (16, 1)):
    obm = po.OracleBeams(w.beams, grid_res=res)
    po.beams_eval(obm, w.media, w.queries, accel=accel, signatures=True)
for rtype in (scenes.REDUCE_PRESAMPLE, scenes.REDUCE_OFFSET, scenes.REDUCE_RESAMPLE_FIXED, scenes.REDUCE_RESAMPLE):
    red = po.OracleBeams(w.beams, grid_res=8).set_reduction(2, rtype, 9)
    for accel in (0, 1):
        po.beams_eval(red, w.media, w.queries, accel=accel, signatures=True)
    red.selection_pdf([0.5, 0.5, 0.5])

u, inm, mi = scenes.upbp_scene(16, 10, 800, 200, seed=4)
uph = po.OraclePhotons(u.photons, in_medium=inm)
ubm = po.OracleBeams(u.beams, grid_res=8).set_reduction(2, scenes.REDUCE_RESAMPLE_FIXED, 3)
for accel in (0, 1):
    po.upbp_eval(uph, ubm, u.media, u.queries, mi, 0, accel=accel)
    po.upbp_eval(None, ubm, u.media, u.queries, mi, 1, accel=accel)
po.upbp_eval(uph, None, u.media, u.queries, mi, 0, precision=64)

g = np.ones((w.queries.count, 3))
po.adjoint(po.OraclePhotons(w.photons, build_bvh=False), po.OracleBeams(w.beams, grid_res=0), w.media, w.queries, g, g)

for ms in (scenes.pp3d_scene(n_paths=300, n_queries=400, seed=5), scenes.surf_scene(n_paths=300, n_queries=400, seed=6)):
    og = po.OracleHashGrid(ms)
    po.vm_eval(og, ms)
    po.vm_eval(og, ms, accel=0, precision=64)

pts = np.random.default_rng(1).uniform(size=(300, 3)).astype(np.float32)
po.knn_d2(pts, pts[:50], 7)

sc = scenes.light_scene(max_path_length=6)
po.trace_light_paths(sc, 300, margins=True)

# malformed inputs are rejected, never read out of bounds
bad = scenes.Queries(w.queries.origin[:2].copy(), w.queries.direction[:2].copy(), np.array([0, 5, 1], np.int64),
                     w.queries.segments[:1].copy(), w.queries.beam_type)
for call in (lambda: po.bre_eval(oph, w.media, bad), lambda: po.beams_eval(obm, w.media, bad)):
    try:
        call()
        raise SystemExit("malformed offsets accepted")
    except RuntimeError:
        pass
seg = w.queries.segments.copy()
seg["medium_id"][0] = 99
badm = scenes.Queries(w.queries.origin, w.queries.direction, w.queries.seg_offsets, seg, w.queries.beam_type)
try:
    po.bre_eval(oph, w.media, badm)
    raise SystemExit("bad medium id accepted")
except RuntimeError:
    pass
print("SANITIZED_OK")
'''


def test_oracle_runs_clean_under_asan_and_ubsan():
    try:
        subprocess.check_call(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "asan"])
    except (OSError, subprocess.CalledProcessError) as e:
        pytest.skip("sanitized oracle build unavailable: %s" % e)
    libasan = subprocess.check_output(["gcc", "-print-file-name=libasan.so"], text=True).strip()
    if not os.path.isabs(libasan) or not os.path.exists(libasan):
        pytest.skip("libasan not found")
    env = dict(os.environ)
    env.update(LD_PRELOAD=libasan, UPBP_ORACLE_LIB=os.path.join(ROOT, "oracle", "_asan", "liboracle.so"),
               ASAN_OPTIONS="detect_leaks=0:abort_on_error=1", UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1",
               PYTHONPATH=ROOT, OMP_NUM_THREADS="4")
    r = subprocess.run([sys.executable, "-c", CHILD], cwd=ROOT, env=env, capture_output=True, text=True, timeout=600)
    assert r.returncode == 0 and "SANITIZED_OK" in r.stdout, r.stdout[-3000:] + r.stderr[-6000:]
