"""mrh_tv_energy_grad_noalloc (partials in a static device array, no allocation) equals
mrh_tv_energy_grad (caller's workspace) on two calls in a row, the second on a smaller input."""
import ctypes

import numpy as np
import pytest
import torch

from mrirecon_amd import RegManager, TVND, TVParams, _lib

pytestmark = pytest.mark.gpu
DEV = torch.device("cuda:0")


def _case(P, F, S0, S1, seed):
    rng = np.random.default_rng(seed)
    x = (rng.standard_normal((P * F, S0, S1)) + 1j * rng.standard_normal((P * F, S0, S1))).astype(np.complex64)
    g0 = (rng.standard_normal(x.shape) + 1j * rng.standard_normal(x.shape)).astype(np.complex64)
    regm = RegManager([TVND("t", TVParams(weight=0.02, axes="temporal")),
                       TVND("s", TVParams(weight=0.005, axes="spatial"))])
    return regm.tv_desc(P, F, S0, S1), torch.from_numpy(x).to(DEV), torch.from_numpy(g0).to(DEV)


def test_noalloc_tv_equals_workspace_entry_point():
    lib = _lib.load()
    work = torch.empty(int(lib.mrh_reduce_workspace_len()), dtype=torch.float64, device=DEV)
    for (P, F, S0, S1, seed) in ((3, 34, 256, 256, 1), (2, 5, 48, 40, 2)):   # second call smaller
        print(f"[noalloc] P={P} F={F} {S0}x{S1}", flush=True)
        desc, x, g0 = _case(P, F, S0, S1, seed)
        st = _lib.stream_of(x)
        g_ref, g_new = g0.clone(), g0.clone()
        e_ref = torch.full((P,), -1.0, dtype=torch.float64, device=DEV)
        e_new = torch.full((P,), -2.0, dtype=torch.float64, device=DEV)
        _lib.check(lib.mrh_tv_energy_grad(ctypes.byref(desc), x.data_ptr(), g_ref.data_ptr(), work.data_ptr(),
                                          e_ref.data_ptr(), st))
        _lib.check(lib.mrh_tv_energy_grad_noalloc(ctypes.byref(desc), x.data_ptr(), g_new.data_ptr(),
                                                  e_new.data_ptr(), st))
        # energy only (g = NULL) through the static array as well
        e_only = torch.zeros((P,), dtype=torch.float64, device=DEV)
        _lib.check(lib.mrh_tv_energy_grad_noalloc(ctypes.byref(desc), x.data_ptr(), None, e_only.data_ptr(), st))
        torch.cuda.synchronize()
        assert torch.equal(e_new, e_ref) and torch.equal(e_only, e_ref)
        assert torch.equal(g_new, g_ref)
        assert not torch.equal(g_ref, g0)
    bad = _case(1, 2, 8, 8, 3)[0]
    bad.P = 70000
    with pytest.raises(ValueError):
        _lib.check(lib.mrh_tv_energy_grad_noalloc(ctypes.byref(bad), x.data_ptr(), None, e_only.data_ptr(), st))
