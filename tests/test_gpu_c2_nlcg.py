"""Full-size NLCG parity on the GRASP workload (256^2 x 34 frames, 12 coils, 21 spokes x 512 per
frame, temporal TV 0.02 + spatial TV 0.002): the device engine against the oracle trajectory
committed in tests/golden/c2_nlcg.json (written by tests/golden/make_c2_nlcg.py with the
reference semantics of graspcg/solvers/cg.py:162-246 and graspcg/numerics/line_search.py:24-79).

Bar: the same accepted steps, every f_k within 1e-4 relative, the same stop reason.
"""
import json
import os
import sys

import numpy as np
import pytest
import torch

from mrirecon_amd import NLCGEngine, RegManager, TVND, TVParams
from mrirecon_amd.nufft.backends.hip import HipNUFFTAdapter
from oracle.nlcg import TV, objective_value

pytestmark = pytest.mark.gpu
DEV = torch.device("cuda:0")
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.fixture(scope="module")
def c2():
    sys.path.insert(0, GOLDEN)
    try:
        from make_c2_nlcg import c2_inputs, checksums
    finally:
        sys.path.remove(GOLDEN)
    with open(os.path.join(GOLDEN, "c2_nlcg.json")) as fh:
        ref = json.load(fh)
    wl, ora, y, x0 = c2_inputs()
    chk = checksums(y, x0)
    for k in ("y_norm2", "x0_norm2"):                 # the inputs the fixture was made from
        assert abs(chk[k] - ref["checks"][k]) <= 1e-10 * abs(ref["checks"][k]), k
    be = HipNUFFTAdapter(ndim=2, traj_units="rad", kb_width=wl.W, osf=wl.osf)
    be.prepare((wl.n_coils, wl.S, wl.S), torch.from_numpy(wl.traj), torch.from_numpy(wl.dcf),
               torch.from_numpy(wl.maps).to(DEV), device=DEV, k_offsets=wl.k_offsets)
    regm = RegManager([TVND("tv_t", TVParams(weight=wl.lam_t, axes="temporal")),
                       TVND("tv_s", TVParams(weight=wl.lam_s, axes="spatial"))])
    return wl, ora, y, x0, be, regm, ref


@pytest.mark.parametrize("name,kw", [("prplus_restart", dict(direction="prplus", restart_c=0.0)),
                                     ("fr", dict(direction="fr"))])
def test_c2_nlcg_matches_the_oracle_trajectory(c2, name, kw):
    wl, ora, y, x0, be, regm, ref = c2
    r = ref["runs"][name]
    eng = NLCGEngine(be.plan, torch.from_numpy(y).to(DEV), problems=1, frames=wl.B, regm=regm, max_hist=64, **kw)
    eng.prime(torch.from_numpy(x0.astype(np.complex64)).to(DEV))
    eng.run(30)
    torch.cuda.synchronize()
    hist = eng.history()[0]
    stop = eng.stop_reasons()[0]
    f = np.array([h["f"] for h in hist])
    fr = np.array(r["f"])
    print(f"[c2 {name}] {len(hist)} it (oracle {r['n_iter']}), stop {stop} (oracle {r['stop']}), "
          f"max rel f diff {np.max(np.abs(f - fr[:len(f)]) / np.abs(fr[:len(f)])):.2e}")
    assert len(hist) == r["n_iter"]
    assert [h["step"] for h in hist] == r["step"]
    assert np.max(np.abs(f - fr) / np.abs(fr)) <= 1e-4
    assert np.max(np.abs(np.array([h["f_step"] for h in hist]) - r["f_step"]) / np.abs(r["f_step"])) <= 1e-4
    expect_stop = {"max_iter": "active", "not_descent": "not_descent"}[r["stop"]]
    assert stop == expect_stop
    regs = [TV(wl.lam_t, (0,)), TV(wl.lam_s, (1, 2))]
    f_gpu = objective_value(ora.forward, y, eng.x.cpu().numpy().astype(np.complex128), regs)
    print(f"[c2 {name}] final f {f_gpu:.10e} (oracle {r['f_final']:.10e})")
    assert abs(f_gpu - r["f_final"]) <= 1e-4 * abs(r["f_final"])
    eng.close()
