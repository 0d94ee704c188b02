"""The C-ABI library loads, exports every entry point include/mrirecon_hip.h declares,
and its ctypes mirror has the same struct layouts as the C compiler (no GPU needed)."""
import ctypes
import os
import re
import subprocess

import pytest

from mrirecon_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "mrirecon_hip.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(mrh_\w+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = _lib.load()
    names = declared_functions()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert sorted(names) == _lib.exported_symbols()


def test_abi_version_and_workspace():
    lib = _lib.load()
    assert lib.mrh_abi_version() == _lib.ABI_VERSION
    assert lib.mrh_reduce_workspace_len() >= 1024


def test_errors_are_reported_without_a_gpu():
    lib = _lib.load()
    rc = lib.mrh_nufft_plan_create(None, None)
    assert rc == 1 and b"null" in lib.mrh_last_error()
    with pytest.raises(ValueError):
        _lib.check(rc)
    d = _lib.NufftDesc()
    d.ndim = 3
    h = ctypes.c_void_p()
    rc = lib.mrh_nufft_plan_create(ctypes.byref(d), ctypes.byref(h))
    assert rc == 4
    with pytest.raises(NotImplementedError):
        _lib.check(rc)
    # temporal basis: sizes are checked before any pointer or launch; N = 0 is a no-op
    rc = lib.mrh_tbasis_forward(None, 4, 17, None, 10, None, None)
    assert rc == 1 and b"K = 17" in lib.mrh_last_error()
    rc = lib.mrh_tbasis_adjoint(None, 0, 2, None, 10, None, 1, None)
    assert rc == 1 and b"T = 0 frames" in lib.mrh_last_error()
    assert lib.mrh_tbasis_diag(None, 3, 2, None, 1, 0, None, None) == 0
    rc = lib.mrh_tbasis_diag(None, 3, 2, None, 1, 5, None, None)
    assert rc == 1 and b"null" in lib.mrh_last_error()


C_PROBE = r"""
#include <stddef.h>
#include <stdio.h>
#include "mrirecon_hip.h"
#define P(T, f) printf(#T "." #f " %zu\n", offsetof(T, f));
int main(void) {
  printf("mrh_nufft_desc %zu\nmrh_nufft_info %zu\nmrh_tv_term %zu\nmrh_tv_desc %zu\nmrh_cg_desc %zu\nmrh_cg_hist %zu\n"
         "mrh_nufft3_desc %zu\n",
         sizeof(mrh_nufft_desc), sizeof(mrh_nufft_info), sizeof(mrh_tv_term), sizeof(mrh_tv_desc),
         sizeof(mrh_cg_desc), sizeof(mrh_cg_hist), sizeof(mrh_nufft3_desc));
  P(mrh_nufft_desc, beta) P(mrh_nufft_desc, k_offsets) P(mrh_nufft_desc, maps) P(mrh_nufft_desc, strip_cols)
  P(mrh_nufft_info, beta) P(mrh_nufft_info, workspace_bytes) P(mrh_nufft_info, coil_chunk)
  P(mrh_tv_desc, term) P(mrh_cg_desc, k_prob) P(mrh_cg_desc, c1) P(mrh_cg_desc, max_hist) P(mrh_cg_desc, restart_c)
  P(mrh_cg_desc, ls_kind) P(mrh_cg_desc, c2) P(mrh_cg_desc, diag) P(mrh_cg_desc, gl_weight) P(mrh_cg_desc, gl_F)
  P(mrh_cg_desc, gl_L) P(mrh_cg_desc, gl_prof)
  P(mrh_nufft3_desc, beta) P(mrh_nufft3_desc, k_offsets) P(mrh_nufft3_desc, dcf_mode) P(mrh_nufft3_desc, maps)
  return 0;
}
"""


def test_struct_layouts_match_c(tmp_path):
    src = tmp_path / "probe.c"
    src.write_text(C_PROBE)
    exe = tmp_path / "probe"
    subprocess.run(["gcc", "-std=c11", f"-I{os.path.join(ROOT, 'include')}", str(src), "-o", str(exe)],
                   check=True)
    out = dict(line.rsplit(" ", 1) for line in subprocess.run([str(exe)], capture_output=True, text=True,
                                                             check=True).stdout.split("\n") if line)
    sizes = {"mrh_nufft_desc": _lib.NufftDesc, "mrh_nufft_info": _lib.NufftInfo, "mrh_tv_term": _lib.TVTerm,
             "mrh_tv_desc": _lib.TVDesc, "mrh_cg_desc": _lib.CGDesc, "mrh_cg_hist": _lib.CGHist,
             "mrh_nufft3_desc": _lib.Nufft3Desc}
    for cname, cls in sizes.items():
        assert int(out[cname]) == ctypes.sizeof(cls), cname
    offs = {"mrh_nufft_desc": _lib.NufftDesc, "mrh_nufft_info": _lib.NufftInfo, "mrh_tv_desc": _lib.TVDesc,
            "mrh_cg_desc": _lib.CGDesc, "mrh_nufft3_desc": _lib.Nufft3Desc}
    for key, val in out.items():
        if "." in key:
            cname, field = key.split(".")
            assert getattr(offs[cname], field).offset == int(val), key


def test_graph_laplacian_host_side():
    """The host builds the graph term's L and majoriser profile as the oracle does (symmetrised W,
    zero diagonal, dense or sparse input); the standalone TV op refuses a manager holding a graph
    term, and a manager holds at most one."""
    import numpy as np
    import torch

    from mrirecon_amd import GraphLaplacian, GraphLapParams, RegManager, TVND, TVParams
    from oracle.nlcg import GraphLap

    rng = np.random.default_rng(3)
    W = rng.random((7, 7))
    W[2, :] = W[:, 2] = 0.0                  # an isolated frame
    for norm in ("none", "rw", "sym"):
        gl = GraphLaplacian("g", torch.from_numpy(W).to_sparse(), GraphLapParams(weight=0.3, normalize=norm))
        ref = GraphLap(0.3, W, normalize=norm)
        assert np.allclose(gl.laplacian(), ref.laplacian(), rtol=0, atol=1e-12)
        assert np.allclose(gl.profile(), ref.profile(), rtol=0, atol=1e-12)
    regm = RegManager([TVND("t", TVParams(weight=0.01, axes="temporal")), gl])
    assert regm.graph() is gl and len(regm.terms()) == 1
    with pytest.raises(NotImplementedError):
        regm.energy_and_grad_tensors(torch.zeros((7, 4, 4), dtype=torch.complex64))
    with pytest.raises(NotImplementedError):
        RegManager([gl, GraphLaplacian("h", W, GraphLapParams(weight=0.1))])
    with pytest.raises(ValueError):
        GraphLaplacian("bad", W, GraphLapParams(weight=0.1, normalize="lap"))


def test_missing_library_fails_loudly(tmp_path):
    """No CPU fallback: with the library absent, the first use raises ImportError."""
    import sys
    env = dict(os.environ, MRH_LIB=str(tmp_path / "absent.so"))
    code = ("from mrirecon_amd import _lib\n"
            "try:\n    _lib.load()\nexcept ImportError as e:\n    print('ImportError:', e)\n")
    out = subprocess.run([sys.executable, "-c", code], cwd=ROOT, env=env, capture_output=True, text=True,
                         timeout=300)
    assert out.returncode == 0, out.stderr
    assert "ImportError:" in out.stdout and "no CPU fallback" in out.stdout
