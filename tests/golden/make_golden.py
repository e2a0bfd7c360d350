"""Writes tests/golden/golden_2d_n64.npz from the FP64 oracle: a 2D N = 64
fractional-Brownian state (seed 7, Leray-projected), its RHS, and the state and
dt sequence after 5 SSP-RK3 steps at C = 0.2 with SmoothCutoff1D(0.05/N, 1, 2 pi N/4)."""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import azeban_oracle as O  # noqa: E402
from azeban_amd import init as I  # noqa: E402


def main():
    N = 64
    u0 = O.initialize(I.fractional_brownian(2, N, seed=7))
    eq = O.IncompressibleEuler(O.Grid(2, N), O.baseline_visc(N))
    dudt = eq.dudt(u0)
    sim = O.Simulation(eq, u0, 0.2)
    sim.simulate_until(1.0, max_steps=5)
    np.savez_compressed(os.path.join(HERE, "golden_2d_n64.npz"), u0=u0, dudt=dudt, u5=sim.u,
                        dts=np.array(sim.dts))


if __name__ == "__main__":
    main()
