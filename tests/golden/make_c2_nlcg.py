"""Write tests/golden/c2_nlcg.json: the oracle's NLCG trajectory on the GRASP workload.

Workload: synth.config_c2() (256^2 x 34 frames, 12 coils, 21 golden-angle spokes x 512 per
frame, KB width 6, grid 512^2, temporal TV 0.02 + spatial TV 0.002), k-space y = the FP64
oracle's forward of the phantom rounded to complex64 plus 1 % complex Gaussian noise (seed 202),
start x0 = the FP64 oracle's adjoint of y.  Runs (oracle/nlcg.py, reference semantics of
graspcg/solvers/cg.py:162-246 with Armijo, line_search.py:24-79; x evaluated at x + t d
without the add/subtract drift of objective.py:256-266, as the device engine does):

* PR+ with the beta restart (restart_c = 0), 30 iterations;
* FR with the reference's first beta = 1 (directions.py:145-155), which stops after one
  iteration with "not_descent" on this workload.

Per iteration the fixture keeps f at the start, the accepted step, f at the step and beta,
plus checksums of y and x0 so that a consumer can confirm it rebuilt the same inputs.
Takes about 20 minutes on 8 CPU cores.  Usage: python tests/golden/make_c2_nlcg.py
"""
from __future__ import annotations

import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from mrirecon_amd import synth  # noqa: E402
from oracle.kbnufft import KBNufft  # noqa: E402
from oracle.nlcg import TV, nlcg  # noqa: E402

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "c2_nlcg.json")


def c2_inputs():
    """(workload, oracle, y, x0) of the fixture, rebuilt identically by the GPU test."""
    wl = synth.config_c2()
    ora = KBNufft(wl.maps, wl.traj, wl.k_offsets, wl.dcf, wl.W, wl.osf)
    y = synth.measure(wl, ora.forward)
    x0 = ora.adjoint(y)
    return wl, ora, y, x0


def checksums(y: np.ndarray, x0: np.ndarray) -> dict:
    y64, x64 = y.astype(np.complex128), x0.astype(np.complex128)
    return {"y_norm2": float(np.vdot(y64, y64).real), "y_first": [float(y64[0].real), float(y64[0].imag)],
            "x0_norm2": float(np.vdot(x64, x64).real)}


def main() -> None:
    t0 = time.time()
    wl, ora, y, x0 = c2_inputs()
    regs = [TV(wl.lam_t, (0,)), TV(wl.lam_s, (1, 2))]
    out = {"workload": "config_c2", "checks": checksums(y, x0), "runs": {}}
    for name, kw in (("prplus_restart", dict(direction="prplus", restart_c=0.0, max_iter=30)),
                     ("fr", dict(direction="fr", max_iter=30))):
        r = nlcg(ora.forward, ora.adjoint, y.astype(np.complex128), x0.astype(np.complex128), regs,
                 mutate_x=False, **kw)
        out["runs"][name] = {
            "n_iter": r.n_iter, "stop": r.stop, "f_final": r.f_final,
            "f": [h["f"] for h in r.history], "step": [h["step"] for h in r.history],
            "f_step": [h["f_step"] for h in r.history], "beta": [h.get("beta") for h in r.history]}
        print(f"{name}: {r.n_iter} iterations, stop {r.stop}, f {r.f_final:.10e} ({time.time() - t0:.0f} s)",
              flush=True)
    with open(OUT, "w") as fh:
        json.dump(out, fh, indent=1)
    print(f"wrote {OUT}")


if __name__ == "__main__":
    main()
