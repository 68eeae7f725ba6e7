"""Captured iterations (HIP graphs of the two halves of a queued iteration, replayed with the
momentum step read from device memory) against the eager launches of the same iterations:
identical volumes and identical per-iteration objective, restart sums and decisions, bit for
bit, over whole solves (every replay runs between the solver's own eager work: the scalar
copies, the host's decisions, the step writes), two solves in a row on one context (the
second replays the first's graphs), the fused and multi-pass mode 7, modes 5 and 6, a
run-time-plan size and forced backtracks (a new step size captures new graphs)."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def cfg(W, H, Nz, its=12, reg=0.01, tv=0.01, ntv=5, mode=7):
    return dict(width=W, height=H, num_planes=Nz, start_plane=1000.0, plane_stepsize=50.0, wavelength=0.633,
                resolution=2.0, num_inverse_iterations=its, regularization_param=reg,
                regularization_TV_param=tv, num_tv_iterations=ntv, mode=mode)


def _solves(pkg, c, holos, graphs, backtracks=None):
    old = {k: os.environ.get(k) for k in ("RIHVR_GRAPHS", "RIHVR_TEST_BACKTRACKS")}
    os.environ["RIHVR_GRAPHS"] = "1" if graphs else "0"
    if backtracks is not None:
        os.environ["RIHVR_TEST_BACKTRACKS"] = str(backtracks)
    try:
        out = []
        with pkg.SparseCompressiveHolo(pkg.make_params(**c)) as rec:
            for h in holos:
                st = rec.inverseReconstruct(h)
                out.append((rec.getVolume(), st))
        return out
    finally:
        for k, v in old.items():
            if v is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = v


CASES = [
    ("mode7", cfg(64, 64, 8)),
    ("c1", None),
    ("mode5", cfg(64, 48, 6, tv=0.0, reg=0.02)),
    ("mode6", cfg(64, 64, 6, mode=6)),
    ("ntv12", cfg(96, 64, 4, ntv=12)),
    ("runtime-plan", cfg(272, 208, 3)),
]


@pytest.mark.parametrize("name", [n for n, _ in CASES])
def test_captured_iterations_equal_eager(pkg, scenes, gpu_available, name):
    c = dict(CASES[[n for n, _ in CASES].index(name)][1] or dict(scenes.C1, mode=7))
    print(f"graphs vs eager: {name} {c['width']}x{c['height']}x{c['num_planes']}", flush=True)
    holos = [scenes.disks_scene(scenes.optics_of(c), 6, 2.0, 4.0, 40 + k, 0.01)[0] for k in range(2)]
    eager = _solves(pkg, c, holos, graphs=False)
    graph = _solves(pkg, c, holos, graphs=True)
    for (ve, se), (vg, sg) in zip(eager, graph):
        assert np.array_equal(vg.view(np.uint32), ve.view(np.uint32))
        for k in ("f", "restart_sum", "restarts", "backtracks", "ratio", "lim"):
            assert np.array_equal(np.asarray(sg[k]), np.asarray(se[k])), k
        assert sg["f_final"] == se["f_final"] and sg["nnz"] == se["nnz"]


def test_captured_iterations_with_forced_backtracks(pkg, scenes, gpu_available):
    """Iterations 2 and 5 reject their first attempt: the step halves, so new graphs are
    captured for the new step, and the queued work of the rejected attempt is discarded."""
    c = cfg(64, 64, 8, its=10)
    holos = [scenes.disks_scene(scenes.optics_of(c), 6, 2.0, 4.0, 51, 0.01)[0]]
    mask = (1 << 2) | (1 << 5)
    eager = _solves(pkg, c, holos, graphs=False, backtracks=mask)
    graph = _solves(pkg, c, holos, graphs=True, backtracks=mask)
    (ve, se), (vg, sg) = eager[0], graph[0]
    assert list(sg["backtracks"]) == list(se["backtracks"]) and sum(se["backtracks"]) == 2
    assert np.array_equal(vg.view(np.uint32), ve.view(np.uint32))
    assert np.array_equal(np.asarray(sg["f"]), np.asarray(se["f"]))
