"""CPU tests that pin the oracle's FISTA control loop (oracle ``Solver::solve``, the
restatement of src/core/sparse_compressive_holo.cu:173-633) without restating it.

The loop is composed here from oracle functions that are pinned elsewhere -- A and
A^T (closed forms K1-K4, K8, golden vectors), the prox (golden vectors, closed forms,
ROF optimality) and the restart weights (single-voxel cases) -- following the
textbook FISTA of Beck & Teboulle (SIAM J. Imaging Sci. 2, 2009) with the adaptive
gradient restart of O'Donoghue & Candes (Found. Comput. Math. 15, 2015):

    g_k     = A^T (A x_k - b)
    y_k     = prox_t(x_k - t g_k)                      t = 1 / Nz
    restart = <x_k - y_k, y_k - y_{k-1}> > 0           (with the reference's weights)
    alpha   = 1 after a restart; alpha' = (1 + sqrt(1 + 4 alpha^2)) / 2
    x_{k+1} = y_k + (alpha - 1) / alpha' (y_k - y_{k-1})

The first momentum steps of an unrestarted FISTA are the textbook constants
0, (phi - 1) / alpha_3 = 0.2817535, 0.4340428, 0.5310638 (phi the golden ratio).
A misreading of the momentum step, the alpha recurrence, the returned point (the
momentum point x_K, Q8), the gradient's sign or the restart bookkeeping makes the
composed iterate differ from the oracle's.  A long mode-5 solve is checked at the
fixed point x = soft(x - t A^T(A x - b), t lambda) that any correct proximal
gradient loop must reach, with the soft threshold in its closed form.
"""
import pathlib
import sys

import numpy as np
import pytest

HERE = pathlib.Path(__file__).resolve().parent
sys.path.insert(0, str(HERE / "golden"))
import numpy_restatement as R  # noqa: E402


def cfg(W, H, Nz, its, reg, tv, ntv=5, mode=7, z0=1000.0, dz=50.0, lam=0.633, dx=2.0):
    return dict(width=W, height=H, num_planes=Nz, start_plane=z0, plane_stepsize=dz, wavelength=lam,
                resolution=dx, num_inverse_iterations=its, regularization_param=reg,
                regularization_TV_param=tv, num_tv_iterations=ntv, mode=mode)


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(np.asarray(b)), 1e-300))


# textbook FISTA momentum steps (alpha_1 = 1, no restart): (alpha_k - 1) / alpha_{k+1}
TEXTBOOK_STEPS = [0.0, 0.2817535251, 0.4340427828, 0.5310638054]


def compose(orc, p, holo, iters):
    """FISTA composed from the oracle's pinned A, A^T, prox and the restart weights.
    Returns the momentum point after `iters` iterations, the objective after each
    iteration, the restart decisions, the restart sums and the momentum steps used."""
    Nz = p.num_planes
    t = 1.0 / Nz
    reg, reg_tv = float(np.float32(p.regularization_param)), float(np.float32(p.regularization_TV_param))
    b = holo.astype(np.complex128)
    x = np.zeros((Nz, p.height, p.width), np.complex128)
    y_prev = x.copy()
    alpha = 1.0
    fs, restarts, rsums, steps = [], [], [], []
    for _ in range(iters):
        g = orc.adjoint(p, orc.forward(p, x, "f64") - b, "f64")
        y = orc.prox(p, x, g, t, "f64")
        # the reference's restart test: sum of weighted (x_k - y_k)(y_k - y_{k-1}) > 0
        rs = R.restart_sum(x, y, y_prev)
        restart = rs > 0
        if restart:
            alpha = 1.0
        alpha_next = (1.0 + np.sqrt(1.0 + 4.0 * alpha * alpha)) / 2.0
        s = (alpha - 1.0) / alpha_next
        alpha = alpha_next
        x = y + s * (y - y_prev)
        y_prev = y
        r = orc.forward(p, x, "f64") - b
        f = 0.5 * float(np.sum(np.abs(r) ** 2)) + reg * float(np.sum(np.abs(x)))
        if p.mode == 7 and reg_tv != 0:
            f += reg_tv * R.tv2d(x)
        fs.append(f)
        restarts.append(int(restart))
        rsums.append(rs)
        steps.append(s)
    return x, np.array(fs), restarts, np.array(rsums), steps


@pytest.fixture(scope="module")
def small_scene(scenes):
    c = cfg(32, 32, 2, its=45, reg=0.01, tv=0.01)
    holo, _ = scenes.disks_scene(scenes.optics_of(c), 3, 2.0, 4.0, 3, 0.01)
    return c, holo


def test_first_iterations_are_textbook_fista(orc, small_scene):
    """Iterations 1-4 of the oracle equal FISTA composed from pinned primitives, with the
    textbook momentum steps 0, 0.2818, 0.4340, 0.5311 (no restart in this window)."""
    c, holo = small_scene
    for K in (1, 2, 3, 4):
        p = orc.make_params(**dict(c, num_inverse_iterations=K))
        vol, est, res, st = orc.solve(p, holo, None, "f64")
        x, fs, restarts, rsums, steps = compose(orc, p, holo, K)
        np.testing.assert_allclose(steps, TEXTBOOK_STEPS[:K], rtol=0, atol=1e-10)
        assert restarts == [0] * K and list(st["restarts"]) == [0] * K
        assert list(st["backtracks"]) == [0] * K
        assert rel(vol, x) < 1e-10, (K, rel(vol, x))
        np.testing.assert_allclose(st["f"], fs, rtol=1e-10)
    # the start: x = 0, so f_initial = |b|^2 / 2
    assert st["f_initial"] == pytest.approx(0.5 * float(np.sum(np.abs(holo.astype(np.complex128)) ** 2)),
                                            rel=1e-12)


def test_restart_resets_momentum_like_textbook_fista(orc, small_scene):
    """45 iterations on the same scene: the restart test fires (at iteration 41), the oracle's
    decisions, restart sums, objective and returned momentum point follow the composition
    through and after the restart."""
    c, holo = small_scene
    p = orc.make_params(**c)
    vol, est, res, st = orc.solve(p, holo, None, "f64")
    x, fs, restarts, rsums, steps = compose(orc, p, holo, c["num_inverse_iterations"])
    assert sum(restarts) >= 1, "the scene must restart"
    assert list(st["restarts"]) == restarts
    k = restarts.index(1)
    assert steps[k] == 0.0 and steps[k + 1] == pytest.approx(TEXTBOOK_STEPS[1], abs=1e-10)
    big = np.abs(rsums) > 1e-12 * np.abs(rsums).max()
    np.testing.assert_allclose(st["restart_sum"][big], rsums[big], rtol=1e-6)
    assert np.all(np.sign(st["restart_sum"]) == np.sign(rsums))
    assert rel(vol, x) < 1e-8, rel(vol, x)
    np.testing.assert_allclose(st["f"], fs, rtol=1e-9)
    # the hologram momentum by linearity (Q8) equals A of the returned point
    np.testing.assert_allclose(est, orc.forward(p, vol, "f64"), atol=1e-9)


def test_mode5_solve_reaches_the_proximal_fixed_point(orc, scenes):
    """A correct proximal-gradient loop converges to x = soft(x - t A^T(A x - b), t lambda)
    (per component, the reference's separate re/im threshold); soft is the closed form
    sign(v) max(|v| - T, 0)."""
    c = cfg(32, 32, 2, its=1000, reg=0.01, tv=0.0, mode=5)
    holo, _ = scenes.disks_scene(scenes.optics_of(c), 3, 2.0, 4.0, 3, 0.01)
    p = orc.make_params(**c)
    x, est, res, st = orc.solve(p, holo, None, "f64")
    t = 0.5
    T = t * float(np.float32(0.01))
    v = x - t * orc.adjoint(p, orc.forward(p, x, "f64") - holo, "f64")
    soft = lambda a: np.sign(a) * np.maximum(np.abs(a) - T, 0.0)  # noqa: E731
    y = soft(v.real) + 1j * soft(v.imag)
    assert rel(y, x) < 1e-7, rel(y, x)
    assert 0 < st["nnz"] < x.size
    # the objective's L1 term is the modulus (Q9): f = |Ax - b|^2 / 2 + lambda sum |x|
    f = 0.5 * float(np.sum(np.abs(orc.forward(p, x, "f64") - holo) ** 2)) + float(np.float32(0.01)) * float(
        np.sum(np.abs(x)))
    assert st["f_final"] == pytest.approx(f, rel=1e-10)


def test_mode7_solve_reaches_the_prox_fixed_point(orc, scenes):
    """Mode 7 after 200 iterations: x = prox(x - t A^T(A x - b)) with the oracle's pinned
    fused-lasso prox, to 1e-6."""
    c = cfg(32, 32, 2, its=200, reg=0.01, tv=0.01)
    holo, _ = scenes.disks_scene(scenes.optics_of(c), 3, 2.0, 4.0, 3, 0.01)
    p = orc.make_params(**c)
    x, est, res, st = orc.solve(p, holo, None, "f64")
    y = orc.prox(p, x, orc.adjoint(p, orc.forward(p, x, "f64") - holo, "f64"), 0.5, "f64")
    assert rel(y, x) < 1e-6, rel(y, x)


def test_planted_loop_mistakes_fail_the_composition(orc, small_scene):
    """The composition discriminates: a momentum step taken from the wrong alpha, a gradient
    of the wrong sign or the prox output returned instead of the momentum point all move
    the iterate by far more than the tolerance used above."""
    c, holo = small_scene
    K = 4
    p = orc.make_params(**dict(c, num_inverse_iterations=K))
    vol, _, _, _ = orc.solve(p, holo, None, "f64")
    Nz = p.num_planes
    t = 1.0 / Nz
    b = holo.astype(np.complex128)

    def run(step_rule, grad_sign=1.0, return_y=False):
        x = np.zeros((Nz, 32, 32), np.complex128)
        y_prev = x.copy()
        a = 1.0
        for _ in range(K):
            g = grad_sign * orc.adjoint(p, orc.forward(p, x, "f64") - b, "f64")
            y = orc.prox(p, x, g, t, "f64")
            a_next = (1.0 + np.sqrt(1.0 + 4.0 * a * a)) / 2.0
            x = y + step_rule(a, a_next) * (y - y_prev)
            y_prev, a = y, a_next
        return y if return_y else x

    good = run(lambda a, an: (a - 1) / an)
    assert rel(vol, good) < 1e-10
    assert rel(vol, run(lambda a, an: (an - 1) / an)) > 1e-3      # alpha index off by one
    assert rel(vol, run(lambda a, an: (a - 1) / an, grad_sign=-1.0)) > 1e-3
    assert rel(vol, run(lambda a, an: (a - 1) / an, return_y=True)) > 1e-3
