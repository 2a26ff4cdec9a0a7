"""Loads the golden fixtures written by make_golden.py back into workload objects."""
import glob
import os

import numpy as np

from smallupbp_amd import scenes

HERE = os.path.dirname(os.path.abspath(__file__))


def names():
    return sorted(os.path.basename(p)[:-4] for p in glob.glob(os.path.join(HERE, "*.npz")))


def load(name):
    z = np.load(os.path.join(HERE, name + ".npz"), allow_pickle=False)
    q = scenes.Queries(z["q_origin"], z["q_direction"], z["q_offsets"], z["q_segments"], int(z["q_beam_type"]))
    ph = bm = None
    if "ph_position" in z:
        r = z["ph_radius"]
        ph = scenes.Photons(z["ph_position"], z["ph_inc_dir"], z["ph_power"], r if r.size else None,
                            float(z["ph_const_radius"]))
    if "bm_origin" in z:
        bm = scenes.Beams(z["bm_origin"], z["bm_direction"], z["bm_length"], z["bm_throughput"], z["bm_flags"],
                          z["bm_medium"], float(z["bm_radius"]))
    w = scenes.Workload(name, z["media"], ph, bm, q)
    return w, z
