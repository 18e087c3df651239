"""Writes tests/golden/column_4x8x4.npz: a 4x8x4-particle column in a 0.16 x 0.3 x 0.16 m box (the
8192-particle column scene at reduced size) and the CPU oracle's state after its neighbour build
and after 3 steps.

    python tests/golden/make_golden.py

The fixture pins the oracle and the HIP path against a fixed, committed set of numbers: any later
change of either that moves a neighbour, a density bit or a velocity beyond float32 noise shows
up against it.  The reference itself cannot produce these (its runtime, NVIDIA Warp, is not
available); they are the oracle's outputs for fixed inputs.
"""

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from implicitvelocityprojection_amd.config import SimConfig  # noqa: E402
from implicitvelocityprojection_amd.scenes import load_particles, synthetic_scene  # noqa: E402
from oracle import oracle as O  # noqa: E402


def make():
    cfg = SimConfig.from_dict(synthetic_scene(0, scale=(4, 8, 4), box=(0.16, 0.3, 0.16)))
    ps = load_particles(cfg)
    sim = O.OracleSim(cfg, ps)
    sim.rebuild()
    out = dict(x0=ps.x, v0=ps.v, is_dynamic=ps.is_dynamic, material=ps.material, density0=ps.density0,
               count0=sim.count, ids0=sim.ids, gradW0=sim.gradW, density_0=sim.density.copy(), k_0=sim.k.copy())
    outer, pcg = [], []
    for _ in range(3):
        sim.step()
        outer.append(sim.last_ip.outer_iters)
        pcg.append(sum(sim.last_ip.pcg_iters))
    out.update(x3=sim.x, v3=sim.v, density_3=sim.density, outer3=np.array(outer), pcg3=np.array(pcg))
    np.savez_compressed(os.path.join(HERE, "column_4x8x4.npz"), **out)


if __name__ == "__main__":
    make()
