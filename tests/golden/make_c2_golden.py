"""Writes tests/golden/c2_oracle_f64.npz: the FP64 oracle's solve of the seeded
1024^2 x 128 workload (plain FISTA + 2D TV, mode 7, 100 FISTA x 5 TV iterations, the
500 tracers + 20 rods hologram of scenes.c2_hologram), stored sparse: the flat indices of
the non-zero voxels of the returned volume (int32), their complex128 values, the
per-iteration objective and decisions, and the hologram's sum as a check that a test
regenerated the same input.  Run from the repository root after building the oracle
(make -C oracle); it takes about half an hour on 8 host cores.
"""
import pathlib
import sys
import time

import numpy as np

ROOT = pathlib.Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

import __graft_entry__ as entry  # noqa: E402
import oracle  # noqa: E402


def main(out=ROOT / "tests" / "golden" / "c2_oracle_f64.npz"):
    scenes = entry._load_submodule("scenes")
    holo = scenes.c2_hologram()
    p = oracle.make_params(**dict(scenes.C2, mode=7))
    t0 = time.time()
    vol, est, res, st = oracle.solve(p, holo, None, "f64")
    idx = np.flatnonzero((vol.real != 0) | (vol.imag != 0))
    np.savez_compressed(out, idx=idx.astype(np.int32), val=vol.ravel()[idx], f=st["f"], restarts=st["restarts"],
                        backtracks=st["backtracks"], f_final=st["f_final"], nnz=st["nnz"],
                        holo_sum=np.sum(holo.astype(np.complex128)))
    print(f"{out}: {len(idx)} voxels, {time.time() - t0:.0f} s")


if __name__ == "__main__":
    main()
