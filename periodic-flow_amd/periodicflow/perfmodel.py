"""Algorithmic HBM bytes of every pass: the roofline model of the time step.

A pass's algorithmic bytes are the elements it is specified to read once plus
the elements it writes once, times the element size.  Re-reads served by
cache are not counted; measured traffic above this figure is waste.

Per member, with nk = N0 [N1] nh stored modes (nh = N_last/2 + 1) and padded
lengths M_i (M = 3N/2 under the 3/2 rule, N under the 2/3 rule).  Under the
2/3 rule the inverse passes touch only retained modes: a0 = 2 floor(N0/3)+1
rows of axis 0, a1 = 2 floor(N1/3)+1 rows of axis 1 and a2 = floor(N2/3)+1
entries of the last axis (3/2 rule: a0 = N0, a1 = N1, a2 = nh).

3D                          read                          write
  inv_axis0_curl            3 a0 a1 a2                    5 M0 a1 a2   (u, omega_y, omega_z)
  inv_axis1                 5 M0 a1 a2                    6 M0 M1 a2
  row_c2r_cross_r2c         6 M0 M1 a2                    3 M0 M1 nh
  fwd_axis1                 3 M0 M1 nh                    3 M0 N1 nh
  fwd_axis0_stage (fused)   3 M0 N1 nh + u_hat, u0: 6 nk  u_hat: 3 nk
  unfused: fwd_axis0        3 M0 N1 nh                    3 nk
           stage            u_hat, R, u0: 9 nk            3 nk
  stage (Stokes)            u_hat, u0: 6 nk               3 nk
2D
  inv_axis0_curl            2 a0 a1                       3 M0 a1      (a1 = floor(N1/3)+1 or nh)
  row_c2r_cross_r2c         3 M0 a1                       2 M0 nh
  fwd_axis0_stage           2 M0 nh + 4 nk                2 nk
  stage (Stokes)            4 nk                          2 nk
(Stage 0 writes u0 instead of reading it: same count.)  Noise drawn on the
device costs no bytes; injected raw draws add 2 nc nk reads per stage.
"""
from __future__ import annotations

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E


def pass_bytes(n, dealias="3/2", precision="fp32", batch=1, nonlinear=True, fused=True):
    """{pass name: bytes per launch} for one plan (all members)."""
    d = len(n)
    c = 16 if precision == "fp64" else 8
    direct = dealias in ("2/3", "truncate")
    M = [n_i if direct else (3 * n_i) // 2 for n_i in n]
    nh = n[-1] // 2 + 1
    out = {}
    if d == 3:
        N0, N1, N2 = n
        M0, M1, _ = M
        nk = N0 * N1 * nh
        a0 = 2 * (N0 // 3) + 1 if direct else N0
        a1 = 2 * (N1 // 3) + 1 if direct else N1
        a2 = N2 // 3 + 1 if direct else nh
        if nonlinear:
            out["inv_axis0_curl"] = 3 * a0 * a1 * a2 + 5 * M0 * a1 * a2
            out["inv_axis1"] = 5 * M0 * a1 * a2 + 6 * M0 * M1 * a2
            out["row_c2r_cross_r2c"] = 6 * M0 * M1 * a2 + 3 * M0 * M1 * nh
            out["fwd_axis1"] = 3 * M0 * M1 * nh + 3 * M0 * N1 * nh
            if fused:
                out["fwd_axis0_stage"] = 3 * M0 * N1 * nh + 9 * nk
            else:
                out["fwd_axis0"] = 3 * M0 * N1 * nh + 3 * nk
                out["stage"] = 12 * nk
        else:
            out["stage"] = 9 * nk
    else:
        N0, N1 = n
        M0, _ = M
        nk = N0 * nh
        a0 = 2 * (N0 // 3) + 1 if direct else N0
        a1 = N1 // 3 + 1 if direct else nh
        if nonlinear:
            out["inv_axis0_curl"] = 2 * a0 * a1 + 3 * M0 * a1
            out["row_c2r_cross_r2c"] = 3 * M0 * a1 + 2 * M0 * nh
            if fused:
                out["fwd_axis0_stage"] = 2 * M0 * nh + 6 * nk
            else:
                out["fwd_axis0"] = 2 * M0 * nh + 2 * nk
                out["stage"] = 8 * nk
        else:
            out["stage"] = 6 * nk
    return {k: v * c * batch for k, v in out.items()}


def step_bytes(n, dealias, precision, n_stages, batch=1, nonlinear=True, fused=True):
    """Algorithmic bytes of one full time step (all stages)."""
    return n_stages * sum(pass_bytes(n, dealias, precision, batch, nonlinear, fused).values())
