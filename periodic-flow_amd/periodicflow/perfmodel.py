"""Algorithmic HBM bytes of every pass: the roofline model of the time step.

A pass's algorithmic bytes are the elements it is specified to read once plus
the elements it writes once, times the element size.  Re-reads served by
cache are not counted; measured traffic above this figure is waste.

Per member, with nk = N0 [N1] nh stored modes (nh = N_last/2 + 1), padded
lengths M_i (M = 3N/2 under the 3/2 rule, N under the 2/3 rule) and nr the
stored last-axis entries the row pass reads (floor(N/3) + 1 under the 2/3
rule, nh otherwise):

3D                                  read                 write
  inv_axis0_curl  u_hat -> A        3 nk                 6 M0 N1 nh
  inv_axis1       A -> B            6 M0 N1 nh           6 M0 M1 nh
  row             B -> C            6 M0 M1 nr           3 M0 M1 nh
  fwd_axis1       C -> D            3 M0 M1 nh           3 M0 N1 nh
  fwd_axis0       D -> R            3 M0 N1 nh           3 nk
  stage           u_hat, R, u0      9 nk                 3 nk
  stage (Stokes)  u_hat, u0         6 nk                 3 nk
2D
  inv_axis0_curl                    2 nk                 3 M0 nh
  row                               3 M0 nr              2 M0 nh
  fwd_axis0                         2 M0 nh              2 nk
  stage                             6 nk                 2 nk
  stage (Stokes)                    4 nk                 2 nk
Noise drawn on the device costs no bytes; injected raw draws add
2 nc nk reads per stage.
"""
from __future__ import annotations

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E


def pass_bytes(n, dealias="3/2", precision="fp32", batch=1, nonlinear=True):
    """{pass name: bytes per launch} for one plan (all members)."""
    d = len(n)
    c = 16 if precision == "fp64" else 8
    direct = dealias in ("2/3", "truncate")
    M = [n_i if direct else (3 * n_i) // 2 for n_i in n]
    nh = n[-1] // 2 + 1
    nr = n[-1] // 3 + 1 if direct else nh
    out = {}
    if d == 3:
        N0, N1, _ = n
        M0, M1, _ = M
        nk = N0 * N1 * nh
        if nonlinear:
            out["inv_axis0_curl"] = 3 * nk + 6 * M0 * N1 * nh
            out["inv_axis1"] = 6 * M0 * N1 * nh + 6 * M0 * M1 * nh
            out["row_c2r_cross_r2c"] = 6 * M0 * M1 * nr + 3 * M0 * M1 * nh
            out["fwd_axis1"] = 3 * M0 * M1 * nh + 3 * M0 * N1 * nh
            out["fwd_axis0"] = 3 * M0 * N1 * nh + 3 * nk
            out["stage"] = 12 * nk
        else:
            out["stage"] = 9 * nk  # u_hat and u0 in, u_hat out (no nonlinear term)
    else:
        N0, _ = n
        M0, _ = M
        nk = N0 * nh
        if nonlinear:
            out["inv_axis0_curl"] = 2 * nk + 3 * M0 * nh
            out["row_c2r_cross_r2c"] = 3 * M0 * nr + 2 * M0 * nh
            out["fwd_axis0"] = 2 * M0 * nh + 2 * nk
            out["stage"] = 8 * nk
        else:
            out["stage"] = 6 * nk
    return {k: v * c * batch for k, v in out.items()}


def step_bytes(n, dealias, precision, n_stages, batch=1, nonlinear=True):
    """Algorithmic bytes of one full time step (all stages)."""
    return n_stages * sum(pass_bytes(n, dealias, precision, batch, nonlinear).values())
