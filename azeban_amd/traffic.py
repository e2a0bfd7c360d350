"""Algorithmic HBM bytes per launch of each kernel family of one RHS evaluation.

Bytes = logical elements a kernel must read + elements it must write, at storage
precision, each counted once; implicit zero padding and truncated modes are not
counted.  Complex state (u_hat, RK buffers): 16 B; intermediates: 16 B (FP64) or
8 B (FP32-mixed).  N: grid, M = 3N/2, nh = N/2 + 1, c = dim, nB = dim(dim+1)/2.
"""
from __future__ import annotations


def family_bytes(dim: int, N: int, precision: str = "fp64") -> dict:
    M = 3 * N // 2
    nh = N // 2 + 1
    cs = 16 if precision == "fp64" else 8
    S = 16
    if dim == 3:
        nm = N * N * nh
        fam = {
            "inv_ax0": 3 * nm * S + 3 * M * N * nh * cs,
            "inv_ax1": 3 * M * N * nh * cs + 3 * M * M * nh * cs,
            "row_fused": 3 * M * M * nh * cs + 6 * M * M * nh * cs,
            "fwd_ax1": 6 * M * M * nh * cs + 6 * M * N * nh * cs,
            "fwd_ax0": 6 * M * N * nh * cs + 6 * nm * cs,
        }
        nb = 6
    else:
        nm = N * nh
        fam = {
            "inv_ax0": 2 * nm * S + 2 * M * nh * cs,
            "row_fused": 2 * M * nh * cs + 3 * M * nh * cs,
            "fwd_ax0": 3 * M * nh * cs + 3 * nm * cs,
        }
        nb = 3
    # k-space + RK stage: B_hat, stage input, RK base (stages 2 and 3), output;
    # averaged over the three stages of a step
    fam["kspace_rk"] = nb * nm * cs + dim * nm * S * (2 + 2.0 / 3.0)
    return fam


def step_bytes(dim: int, N: int, precision: str = "fp64") -> float:
    """Algorithmic bytes of one SSP-RK3 step (three RHS evaluations)."""
    return 3 * sum(family_bytes(dim, N, precision).values())
