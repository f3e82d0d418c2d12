"""The backward kernel differentiates the function the forward renders.

The backward pass recomputes each fragment's sampling state (LOD,
anisotropy, probe axis) from the plane equations and exports it
(swr_grad_desc.sample_state).  Here that export must equal, bit for bit,
the state the oracle's forward pass used for the same fragment
(swo_draw_state) -- on trilinear, 8x anisotropic and 4x MSAA scenes, and on
views at the fitting workloads' full size.  With tests/test_grad_oracle.py
(the FP64 model at the oracle's forward state reproduces the oracle's image)
and the forward parity tests (GPU image = oracle image), this closes the
chain: the gradients the GPU computes are those of the image it renders.
On the grid scenes the FP64 model is also evaluated directly at the GPU's
exported state and compared with the GPU's image.
"""
import numpy as np
import pytest
import torch

from grad_cases import G, params_of, view_of
from oracle_runner import oracle_render_state
from swiftshader_amd import diff, scenes
from test_grad_oracle import STATE_CASES, _grid_scene_mips

pytestmark = pytest.mark.gpu
DEV = "cuda:0"


def _t(a):
    return torch.from_numpy(np.ascontiguousarray(a)).to(DEV)


def _gpu_forward_state(d):
    view = diff.View(d, DEV)
    pos, col, uv = _t(d.positions), _t(d.colors), _t(d.texcoords)
    levels = [_t(l) for l in d.texture.levels]
    with torch.no_grad():
        color = diff.draw(view, pos, col, uv, levels)
    S, H, W = view.shape
    state = torch.zeros((S, H, W, 4), dtype=torch.float32, device=DEV)
    d_color = torch.zeros((S, H, W, 4), dtype=torch.float32, device=DEV)
    diff.backward_raw(view, pos, col, uv, levels, view.ids, d_color, state)
    torch.cuda.synchronize()
    return color.cpu().numpy(), view.ids.cpu().numpy().view(np.uint32), state.cpu().numpy()


def _check_state(name, d):
    color, ids, state = _gpu_forward_state(d)
    ocol, _, oids, ostate, _ = oracle_render_state(d, threads=4)
    assert np.array_equal(ids, oids), f"{name}: primitive ids differ"
    assert np.array_equal(color.view(np.uint32), ocol.view(np.uint32)), f"{name}: image differs"
    cov = ids != 0
    bad = (state.view(np.uint32) != ostate.view(np.uint32)).any(-1) & cov
    assert not bad.any(), f"{name}: {int(bad.sum())} sampling states differ, first at {np.argwhere(bad)[0]}"
    return color, ids, state


@pytest.mark.parametrize("case", range(len(STATE_CASES)))
def test_backward_state_equals_forward_state_and_shades_the_image(case):
    d = _grid_scene_mips(70 + case, **STATE_CASES[case])
    print(f"[gpu] state case {case} {STATE_CASES[case]}", flush=True)
    color, ids, state = _check_state(f"grid {case}", d)
    frag = G.fragments(ids, state)
    got = G.shade(view_of(d), params_of(d), frag)
    for q in range(d.samples):
        sel = frag.smask[:, q]
        want = color[q][frag.y[sel], frag.x[sel]].astype(np.float64)
        assert np.max(np.abs(got[sel] - want)) < 2e-5


@pytest.mark.parametrize("which", ["diff_view_512", "msaa_aniso_view"])
def test_backward_state_equals_forward_state_full_size(which):
    if which == "diff_view_512":
        d = scenes.diff_views(seed=4, views=1)[0]
    else:
        d = scenes.msaa_aniso(seed=5, views=1, tri_lo=20000, tri_hi=20001)[0]
    print(f"[gpu] full-size state {which} {d.width}x{d.height} s={d.samples}", flush=True)
    _, ids, state = _check_state(which, d)
    lods = state[ids != 0][:, 0]
    assert lods.max() - lods.min() > 2.0
