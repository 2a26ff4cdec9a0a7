"""Writes the golden fixtures of tests/golden/*.npz from the CPU oracle.

Each fixture holds the inputs of a small seeded workload (media, photons, beams,
query rays and segments) and the oracle's expected outputs: FP32 radiance in the
reference operation order, the FP64 re-evaluation, and per-query hit signatures.
The reference publishes no vectors for this path, so these pin the oracle (and
the GPU against it) to the values that the closed-form tests in
tests/test_oracle_math.py already validate.

Run: python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import pyoracle as po  # noqa: E402
from smallupbp_amd import scenes  # noqa: E402


def cases():
    yield "bre_cube_24", scenes.plain_bre(24, 24, 2_000, 0.03, seed=101)
    w = scenes.plain_bre(16, 16, 1_500, 0.04, seed=102, mean_cosine=0.5, beam_type=scenes.SHORT_BEAM)
    yield "bre_short_hg_16", w
    yield "bre_beams_nested_16x9", scenes.bre_beams(16, 9, 3_000, 800, seed=103)
    yield "beams_mix_400", scenes.beams_stress(n_queries=400, n_beams=3_000, seed=104)


def pack(name, w):
    d = {"media": w.media, "q_origin": w.queries.origin, "q_direction": w.queries.direction,
         "q_offsets": w.queries.seg_offsets, "q_segments": w.queries.segments,
         "q_beam_type": np.array(w.queries.beam_type)}
    if w.photons is not None:
        oph = po.OraclePhotons(w.photons, build_bvh=False)
        o32, s32, c = po.bre_eval(oph, w.media, w.queries, accel=0, precision=32, signatures=True)
        o64, _, _ = po.bre_eval(oph, w.media, w.queries, accel=0, precision=64)
        d.update(ph_position=w.photons.position, ph_inc_dir=w.photons.inc_dir, ph_power=w.photons.power,
                 ph_radius=w.photons.radius if w.photons.radius is not None else np.zeros(0, np.float32),
                 ph_const_radius=np.array(w.photons.const_radius, np.float32),
                 out_photons_f32=o32, out_photons_f64=o64, sig_photons=s32)
    if w.beams is not None:
        obm = po.OracleBeams(w.beams, grid_res=0)
        o32, s32, c = po.beams_eval(obm, w.media, w.queries, accel=0, precision=32, signatures=True)
        o64, _, _ = po.beams_eval(obm, w.media, w.queries, accel=0, precision=64)
        b = w.beams
        d.update(bm_origin=b.origin, bm_direction=b.direction, bm_length=b.length, bm_throughput=b.throughput,
                 bm_flags=b.flags, bm_medium=b.medium_id, bm_radius=np.array(b.radius, np.float32),
                 out_beams_f32=o32, out_beams_f64=o64, sig_beams=s32)
    np.savez_compressed(os.path.join(HERE, name + ".npz"), **d)


if __name__ == "__main__":
    for name, w in cases():
        pack(name, w)
        print("wrote", name)
