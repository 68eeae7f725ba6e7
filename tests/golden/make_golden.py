"""Writes tests/golden/golden_small.npz from the numpy restatement
(tests/golden/numpy_restatement.py) -- an implementation independent of the C++
oracle.  The reference ships no fixtures for this path, so these vectors are the
committed known answers the oracle is checked against.

Run:  python tests/golden/make_golden.py
"""
import pathlib
import sys

import numpy as np

HERE = pathlib.Path(__file__).resolve().parent
sys.path.insert(0, str(HERE))
import numpy_restatement as R  # noqa: E402

OUT = HERE / "golden_small.npz"


def main():
    rng = np.random.default_rng(20261203)
    g = {}
    # case A: 32 x 32 x 4 planes, 1000 um + 50 um steps, 2 um pixels, 0.633 um
    W, H, Nz = 32, 32, 4
    optA = dict(W=W, H=H, nz=Nz, z0=1000.0, dz=50.0, wavelength=0.633, resolution=2.0)
    x = ((rng.standard_normal((Nz, H, W)) + 1j * rng.standard_normal((Nz, H, W)))
         * (rng.random((Nz, H, W)) < 0.1)).astype(np.complex64).astype(np.complex128)
    r = rng.standard_normal((H, W)).astype(np.float32).astype(np.float64) + 0j
    g["A_x"] = x
    g["A_r"] = r
    for rnd in (False, True):
        op = R.Operator(round_args=rnd, **optA)
        tag = "round" if rnd else "f64"
        g[f"A_fwd_{tag}"] = op.forward(x)
        g[f"A_adj_{tag}"] = op.adjoint(r)
    g["A_kappa"] = R.kappa(W, H, 0.633, 2.0)
    g["A_planes"] = R.plane_list(1000.0, 50.0, Nz)
    # prox inputs
    x0 = (0.5 * (rng.standard_normal((Nz, H, W)) + 1j * rng.standard_normal((Nz, H, W)))
          * (rng.random((Nz, H, W)) < 0.3)).astype(np.complex64).astype(np.complex128)
    gr = (rng.standard_normal((Nz, H, W)) + 1j * rng.standard_normal((Nz, H, W))).astype(np.complex64).astype(np.complex128)
    g["A_prox_x0"] = x0
    g["A_prox_g"] = gr
    g["A_prox_fl2d"] = R.prox_fl2d(x0, gr, 0.25, 0.05, 0.02, 5)
    g["A_prox_l1"] = R.prox_l1(x0, gr, 0.25, 0.05)
    # small solve: a hologram made by the forward operator from a sparse volume + noise
    truth = np.zeros((Nz, H, W), np.complex128)
    truth[1, 10:13, 8:11] = -0.8
    truth[3, 20:22, 20:24] = -0.6
    op = R.Operator(round_args=False, **optA)
    holo = op.forward(truth) + 0.01 * rng.standard_normal((H, W))
    holo = holo.real.astype(np.float32).astype(np.float64) + 0j
    g["A_holo"] = holo
    x1, est, res, log = R.solve(op, holo, iters=6, reg=0.01, reg_tv=0.01, n_tv=5)
    g["A_solve_x"] = x1
    g["A_solve_est"] = est
    g["A_solve_f"] = np.array(log["f"])
    g["A_solve_f_initial"] = np.array(log["f_initial"])
    g["A_solve_restarts"] = np.array(log["restarts"])
    g["A_solve_backtracks"] = np.array(log["backtracks"])
    # the same solve with float-rounded phase arguments (the FP32 oracle's operator)
    opr = R.Operator(round_args=True, **optA)
    xr, _, _, logr = R.solve(opr, holo, iters=6, reg=0.01, reg_tv=0.01, n_tv=5)
    g["A_solve_x_round"] = xr
    g["A_solve_f_round"] = np.array(logr["f"])
    # case B: 48 x 32, the 1024^2 workload's optics (large z: float rounding of the
    # phase argument is visible)
    W, H, Nz = 48, 32, 3
    optB = dict(W=W, H=H, nz=Nz, z0=2500.0, dz=20.0, wavelength=0.632, resolution=1.1)
    xb = ((rng.standard_normal((Nz, H, W)) + 1j * rng.standard_normal((Nz, H, W)))
          * (rng.random((Nz, H, W)) < 0.2)).astype(np.complex64).astype(np.complex128)
    rb = rng.standard_normal((H, W)).astype(np.float32).astype(np.float64) + 0j
    g["B_x"] = xb
    g["B_r"] = rb
    for rnd in (False, True):
        op = R.Operator(round_args=rnd, **optB)
        tag = "round" if rnd else "f64"
        g[f"B_fwd_{tag}"] = op.forward(xb)
        g[f"B_adj_{tag}"] = op.adjoint(rb)
    np.savez_compressed(OUT, **g)
    print(f"wrote {OUT} ({OUT.stat().st_size} bytes, {len(g)} arrays)")


if __name__ == "__main__":
    main()
