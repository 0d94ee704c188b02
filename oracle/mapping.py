"""ORACLE (test infrastructure only) -- CPU restatement of the reference's low-rank
temporal-basis mapping and the TV regulariser applied through it.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may use
this module.  numpy complex128 / float64 throughout.

* TemporalBasisOp (graspcg/regularization/mapping.py:112-231): u = U V with U (T, K)
  and V (K, ...) flattened to (K, N) (:141-147); the adjoint adds U^H g_u into g_V and
  returns zeros to the field (:149-164); diag_push adds sum_t |U_tk|^2 k_u[t] into
  diag_V (:170-199); diag_push_profile adds the per-frame profile's sum_t |U_tk|^2 v[t]
  to every element of diag_V[k] (:200-231).
* MappedRegularizer.energy_grad_fixed_axes (mapping.py:321-356): the inner TV on the
  mapped field u, its gradient pushed back through the op; the field gradient is
  untouched (the op is terminal in parameter space).
* Transformed.add_diag (graspcg/regularization/wrappers.py:39-105): (A) when the
  caller's context carries an image diagonal and the TV runs over frames, the frame
  profile 2 lambda w_0^2 x degree (TVND.majorizer_profile, tv_nd.py:187-215) goes
  through diag_push_profile, which is terminal and returns None, so the code falls
  through to (B) as well: the constant k = 2 lambda sum_a w_a^2 (TVND.majorizer_diag,
  tv_nd.py:166-186) broadcast over u goes through diag_push.

Pinning (tests/test_oracle_mapping.py): the reference's own expected values of
test_diag_push_temporal_basis (graspcg/tests/test_reg_diag_mapped.py:92-126); explicit
loops on tiny sizes; the adjoint identity <U V, W> = <V, U^H W>, which a missing
conjugate or a transposed U fails; U = identity (K = T) reducing the mapped TV to the
plain TV of V; central finite differences of E(V) against g_V along random
complex directions.
"""
from __future__ import annotations

from typing import Optional, Sequence, Tuple

import numpy as np

from .nlcg import TV, _degree, tv_energy_grad


def tb_forward(U: np.ndarray, V: np.ndarray) -> np.ndarray:
    """u = U V: (T, K) x (K, ...) -> (T, ...)   (mapping.py:141-147)."""
    K = V.shape[0]
    return (U @ V.reshape(K, -1)).reshape((U.shape[0],) + V.shape[1:])


def tb_adjoint(U: np.ndarray, gu: np.ndarray, gV: Optional[np.ndarray] = None) -> np.ndarray:
    """g_V (+)= U^H g_u   (mapping.py:160-162)."""
    T = gu.shape[0]
    out = (U.conj().T @ gu.reshape(T, -1)).reshape((U.shape[1],) + gu.shape[1:])
    if gV is not None:
        gV += out
        return gV
    return out


def tb_diag_push(U: np.ndarray, ku: np.ndarray, diag_V: np.ndarray) -> np.ndarray:
    """diag_V[k] += sum_t |U_tk|^2 k_u[t]   (mapping.py:192-197)."""
    T = ku.shape[0]
    U2 = (U.conj() * U).real
    diag_V += (U2.T @ ku.reshape(T, -1)).reshape(diag_V.shape)
    return diag_V


def tb_diag_push_profile(U: np.ndarray, v: np.ndarray, diag_V: np.ndarray) -> np.ndarray:
    """diag_V[k, ...] += sum_t |U_tk|^2 v[t]   (mapping.py:224-230)."""
    U2 = (U.conj() * U).real
    DV = U2.T @ np.asarray(v, dtype=np.float64).reshape(-1)
    diag_V += DV.reshape((-1,) + (1,) * (diag_V.ndim - 1))
    return diag_V


def mapped_tv_energy_grad(U: np.ndarray, V: np.ndarray, regs: Sequence[TV],
                          gV: Optional[np.ndarray] = None) -> Tuple[float, np.ndarray]:
    """E = sum of the TV terms at u = U V (frames on axis 0); g_V += U^H grad_u E.
    (MappedRegularizer.energy_grad_fixed_axes, mapping.py:321-356.)"""
    u = tb_forward(U, V)
    gu = np.zeros_like(u)
    E = 0.0
    for r in regs:
        E += float(tv_energy_grad(u, r, gu))
    gV = np.zeros_like(V) if gV is None else gV
    tb_adjoint(U, gu, gV)
    return E, gV


def mapped_tv_add_diag(U: np.ndarray, reg: TV, diag_V: np.ndarray, image_diag: bool = False) -> np.ndarray:
    """Transformed.add_diag for one TV term through a temporal basis (wrappers.py:39-105)."""
    if reg.weight == 0.0 or len(reg.axes) == 0:
        return diag_V
    if reg.axis_weights is not None and len(reg.axis_weights) != len(reg.axes):
        return diag_V           # no profile and k = 0 for mismatched weights (tv_nd.py:176-179,197-199)
    ws = list(reg.axis_weights) if reg.axis_weights is not None else [1.0] * len(reg.axes)
    T = U.shape[0]
    if image_diag and 0 in reg.axes:                                   # (A), then fall through
        w0 = ws[list(reg.axes).index(0)]
        tb_diag_push_profile(U, 2.0 * reg.weight * w0 * w0 * _degree(T), diag_V)
    k = 2.0 * reg.weight * sum(w * w for w in ws)                      # (B)
    tb_diag_push(U, np.full((T,) + diag_V.shape[1:], k), diag_V)
    return diag_V
