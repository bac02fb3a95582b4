"""Writes tests/golden/golden.npz: small inputs of the benchmark stencils and
the oracle's state after a few steps (regression fixtures for the oracle
and for device tests).  Run: python tests/golden/make_golden.py"""

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [os.path.join(ROOT, "cutile-stencil-dsl_amd"), ROOT]

import oracle  # noqa: E402
from cutile import workloads as w  # noqa: E402


def main():
    heat_in = w.heat2d_init(24, (1, 1), seed=0)
    heat_out = oracle.time_march(heat_in, w.heat2d()._fn, 2, (1, 1), 10, w.dirichlet_spec(2))
    star_in = w.sinusoid_init((12, 10, 14), (4, 4, 4), seed=0)
    oracle.apply_boundary(star_in, w.periodic_spec(3), (4, 4, 4))
    star_out = oracle.time_march(star_in, w.star25()._fn, 3, (4, 4, 4), 2, w.periodic_spec(3))
    gu, gv = w.gray_scott()
    u, v = w.gray_scott_init(32, (1, 1), seed=0, squares=3)
    oracle.apply_boundary(u, w.periodic_spec(2), (1, 1))
    oracle.apply_boundary(v, w.periodic_spec(2), (1, 1))
    res = oracle.time_march_multi({"u": u, "v": v}, [gu._fn, gv._fn], ["u", "v"], 2, (1, 1), 5,
                                  w.periodic_spec(2))
    np.savez_compressed(os.path.join(HERE, "golden.npz"), heat_in=heat_in, heat_out=heat_out,
                        star_in=star_in, star_out=star_out, gs_u_in=u, gs_v_in=v,
                        gs_u_out=res["u"], gs_v_out=res["v"])


if __name__ == "__main__":
    main()
