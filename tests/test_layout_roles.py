"""Axis grammar of the front end (no GPU): roles tokens (graspcg/core/roles.py:23-65) and the
label layouts of user tensors (graspcg/nufft/layout.py:11-194)."""
import pytest
import torch

from mrirecon_amd.nufft.layout import (AxisSpec, plan_image_layout, plan_image_layout_from_sizes,
                                       plan_kspace_layout, plan_kspace_layout_from_sizes)
from mrirecon_amd.roles import IMAGE_ROLES, Roles


def test_roles_tokens_resolve_to_their_groups():
    r = Roles(unlike=1, like=2, nufft=2)
    assert r.total() == r.ndim == 5
    for tok in ("temporal", "time", "unlike", "TEMPORAL"):
        assert r.resolve_axes(tok) == (0,)
    assert r.resolve_axes("like") == (1, 2)
    for tok in ("spatial", "image", "nufft"):
        assert r.resolve_axes(tok) == (3, 4)
    assert r.resolve_axes("all") == (0, 1, 2, 3, 4)
    assert r.resolve_axes("none") == ()
    assert r.temporal == (0,) and r.like_axes == (1, 2) and r.spatial == (3, 4)
    tm = r.token_map()
    assert tm["spatial"] == (3, 4) and tm["none"] == () and len(tm) == 9


def test_roles_lists_mix_tokens_and_indices_first_occurrence_kept():
    r = IMAGE_ROLES   # (F, 1, S0, S1)
    assert r.resolve_axes(["spatial", 0]) == (2, 3, 0)
    assert r.resolve_axes((0, -1, 3, "temporal")) == (0, 3)
    assert r.resolve_axes([]) == ()


@pytest.mark.parametrize("spec,exc", [("diagonal", KeyError), (["x"], KeyError), (4, TypeError),
                                      ({0}, TypeError), ([1.0], TypeError), ([4], IndexError),
                                      ([-5], IndexError)])
def test_roles_errors(spec, exc):
    with pytest.raises(exc):
        IMAGE_ROLES.resolve_axes(spec)


def test_axis_spec_both_keyword_styles_and_defaults():
    a = AxisSpec(image_labels=("B", "L", "coil", "X", "Y"), kspace_labels=("B", "L", "coil", "K"),
                 image_fft_labels=("X", "Y"))
    b = AxisSpec(image=("B", "L", "coil", "X", "Y"), kspace=("B", "L", "coil", "K"), image_fft=("X", "Y"))
    for s in (a, b):
        assert s.coil == "coil" and s.kspace_fft == ("K",)
        assert s.like_image() == ("L",) and s.like_kspace() == ("L",)
    c = AxisSpec(image=("T", "X", "Y"), kspace=("T", "ch", "S"), image_fft=("X", "Y"), kspace_fft=("S",),
                 coil="ch", image_labels=("B", "T", "X", "Y"))
    assert c.image == ("B", "T", "X", "Y")          # the long keyword wins
    assert c.like_kspace() == ("T",) and c.like_image() == ("T",)
    with pytest.raises(TypeError):
        AxisSpec(image=("B", "X", "Y"), kspace=("B", "coil", "K"))
    with pytest.raises(TypeError):
        AxisSpec(image=("B", "X", "Y"), kspace=("B", "coil", "K"), image_fft=("X", "Y"), spatial=("X",))


def test_layout_plans_from_tensors_and_sizes():
    ax = AxisSpec(image=("X", "B", "E", "coil", "Y", "P"), kspace=("K", "P", "coil", "B", "E"),
                  image_fft=("X", "Y"))
    x = torch.zeros(8, 3, 2, 1, 6, 5)
    ip = plan_image_layout(x, ax)
    assert (ip.B, ip.like_labels, ip.like_sizes, ip.L_other, ip.spatial) == (3, ("E", "P"), (2, 5), 10, (8, 6))
    y = torch.zeros(40, 5, 4, 3, 2)
    kp = plan_kspace_layout(y, ax)
    assert (kp.B, kp.like_labels, kp.like_sizes, kp.L_other, kp.C, kp.K) == (3, ("P", "E"), (5, 2), 10, 4, 40)
    with pytest.raises(ValueError):
        plan_image_layout(torch.zeros(8, 3, 2, 4, 6, 5), ax)      # coil axis of size 4
    with pytest.raises(ValueError):
        plan_image_layout(torch.zeros(8, 3, 6), ax)               # wrong rank
    ip2 = plan_image_layout_from_sizes(3, [2, 5], (8, 6), ax)
    assert ip2.like_sizes == (2, 5) and ip2.L_other == 10
    ip3 = plan_image_layout_from_sizes(3, [10], (8, 6), ax)       # one total: the first label takes it
    assert ip3.like_sizes == (10, 1) and ip3.L_other == 10
    kp2 = plan_kspace_layout_from_sizes(3, [7], 4, 40, ax)
    assert kp2.like_sizes == (7, 1) and kp2.L_other == 7 and (kp2.C, kp2.K) == (4, 40)
    with pytest.raises(ValueError):
        plan_image_layout_from_sizes(3, [2, 5, 1], (8, 6), ax)
    with pytest.raises(ValueError):
        plan_kspace_layout_from_sizes(3, [], 4, 40, ax)
    plain = AxisSpec(image=("B", "coil", "X", "Y"), kspace=("B", "coil", "K"), image_fft=("X", "Y"))
    ip4 = plan_image_layout_from_sizes(2, [9, 9], (4, 4), plain)  # no Like labels: sizes ignored
    assert ip4.like_sizes == () and ip4.L_other == 1
