"""Differentiable path on the GPU against the CPU oracle and the FP64 gradient oracle.

* forward with the primitive_id record: colour and ids bit-exact with the oracle;
* swr_generate_mips bit-exact with the box-filter generator; swr_mip_backward
  equal to the FP64 adjoint;
* swr_draw_backward: gradients of every kind within 1e-4 (relative, L2 norm
  over each kind) of the FP64 oracle evaluated on the GPU's own visibility
  record and sampling state, and within 1e-4 of FP64 central finite
  differences on 64 random parameters;
* the torch.autograd.Function path reproduces the direct entry points.
"""
import numpy as np
import pytest
import torch

from grad_cases import G, fd_derivative, float_scene, gradient_rel_error, params_of, view_of
from oracle_runner import oracle_render_ids
from swiftshader_amd import abi, diff, scenes

pytestmark = pytest.mark.gpu
DEV = "cuda:0"
TOL = 1e-4


def _t(a):
    return torch.from_numpy(np.ascontiguousarray(a)).to(DEV)


def gpu_forward(d):
    """Forward through diff.Draw without autograd; returns colour, ids (numpy) and the device inputs."""
    view = diff.View(d, DEV)
    pos, col, uv = _t(d.positions), _t(d.colors) if d.colors is not None else None, \
        _t(d.texcoords) if d.texcoords is not None else None
    levels = [_t(l) for l in d.texture.levels] if d.texture is not None else []
    with torch.no_grad():
        color = diff.draw(view, pos, col, uv, levels)
    torch.cuda.synchronize()
    return view, (pos, col, uv, levels), color.cpu().numpy(), view.ids.cpu().numpy().view(np.uint32)


FWD_CASES = [dict(), dict(samples=4), dict(aniso=True, address=abi.ADDRESS_CLAMP_TO_EDGE),
             dict(shader=abi.SHADER_SAMPLE), dict(shader=abi.SHADER_INTERPOLATE)]


@pytest.mark.parametrize("case", range(len(FWD_CASES)))
def test_forward_with_primitive_ids_is_bit_exact(case):
    d = float_scene(300 + case, size=96, triangles=300, **FWD_CASES[case])
    print(f"[gpu] forward ids case {case}", flush=True)
    _, _, color, ids = gpu_forward(d)
    ocol, _, oids, _ = oracle_render_ids(d)
    assert np.array_equal(ids, oids)
    assert np.array_equal(color.view(np.uint32), ocol.view(np.uint32))
    assert (ids != 0).sum() > 1000


@pytest.mark.parametrize("shape,u8", [((64, 64), False), ((20, 12), False), ((1, 9), False), ((37, 50), True)])
def test_generate_mips_bit_exact(shape, u8):
    rng = np.random.default_rng(5)
    if u8:
        lvl0 = rng.integers(0, 256, shape + (4,), dtype=np.uint8)
    else:
        lvl0 = rng.normal(0.5, 0.15, shape + (4,)).astype(np.float32)
    want = scenes.box_mip_chain(lvl0)
    levels = [_t(lvl0)] + [torch.empty(w.shape, dtype=_t(lvl0).dtype, device=DEV) for w in want[1:]]
    diff.generate_mips(levels)
    torch.cuda.synchronize()
    for a, b in zip(levels, want):
        assert np.array_equal(a.cpu().numpy(), b)


@pytest.mark.parametrize("shape", [(64, 64), (20, 12), (1, 9), (9, 1)])
def test_mip_backward_is_the_adjoint(shape):
    rng = np.random.default_rng(6)
    chain = G.box_chain(np.zeros(shape + (4,)))
    gs = [rng.normal(size=l.shape).astype(np.float32) for l in chain]
    gs[0][:] = 0
    bufs = [_t(g) for g in gs]
    t = diff._texture_struct(bufs, abi.FORMAT_R32G32B32A32_SFLOAT)
    import ctypes
    from swiftshader_amd import runtime
    runtime.check(runtime.lib().swr_mip_backward(ctypes.byref(t), runtime._stream_handle(None)))
    torch.cuda.synchronize()
    want = G.box_chain_adjoint([g.astype(np.float64) for g in gs])
    assert np.max(np.abs(bufs[0].cpu().numpy() - want)) <= 1e-6 * max(1.0, np.abs(want).max())


def run_backward(d, seed):
    view, (pos, col, uv, levels), color, ids = gpu_forward(d)
    S, H, W = view.shape
    rng = np.random.default_rng(seed)
    d_color = rng.normal(size=(S, H, W, 4)).astype(np.float32)
    state = torch.zeros((S, H, W, 4), dtype=torch.float32, device=DEV)
    d_pos, d_col, d_uv, d_lv = diff.backward_raw(view, pos, col, uv, levels, view.ids, _t(d_color), state)
    torch.cuda.synchronize()
    frag = G.fragments(ids, state.cpu().numpy())
    g = G.upstream(frag, d_color)
    gpu = {"positions": d_pos.cpu().numpy()}
    if d_col is not None:
        gpu["colors"] = d_col.cpu().numpy()
    if d_uv is not None:
        gpu["texcoords"] = d_uv.cpu().numpy()
    if d_lv:
        gpu["texture_levels"] = [t.cpu().numpy() for t in d_lv]
    return gpu, frag, g


BWD_CASES = [dict(), dict(samples=4), dict(aniso=True), dict(aniso=True, address=abi.ADDRESS_CLAMP_TO_EDGE, samples=4),
             dict(shader=abi.SHADER_SAMPLE, address=abi.ADDRESS_CLAMP_TO_EDGE), dict(shader=abi.SHADER_INTERPOLATE),
             dict(aniso=True, singular_uv=True), dict(aniso=True, singular_uv=True, samples=4)]


@pytest.mark.parametrize("case", range(len(BWD_CASES)))
def test_backward_matches_fp64_oracle(case):
    kw = BWD_CASES[case]
    d = float_scene(400 + case, size=128, triangles=600, tex=64, **kw)
    print(f"[gpu] backward case {case} {kw}", flush=True)
    gpu, frag, g = run_backward(d, 40 + case)
    view = view_of(d)
    want = G.gradient(view, params_of(d), frag, g)
    assert len(frag.t) > 5000
    for k in ("positions", "colors", "texcoords"):
        if k in want:
            assert gradient_rel_error(gpu[k], want[k]) < TOL, k
    if "texture_levels" in want:
        for lvl, (a, b) in enumerate(zip(gpu["texture_levels"], want["texture_levels"])):
            if np.abs(b).max() > 0:
                assert gradient_rel_error(a, b) < TOL, ("level", lvl)


def test_backward_against_central_differences_64_params():
    """64 random parameters of a 2-view-sized scene: FP64 central differences
    of the forward (frozen visibility and sampling state) against the GPU."""
    d = float_scene(500, size=256, triangles=4000, tex=256, area_hi=2000.0, uv_hi=4.0)
    print("[gpu] finite differences", flush=True)
    gpu, frag, g = run_backward(d, 50)
    view = view_of(d)
    params = params_of(d)
    rng = np.random.default_rng(51)
    used = np.unique(G._vertex_ids(view, frag.t, len(d.positions)))
    # Texture gradient of level 0 from the GPU: its per-level gradients pushed through the FP64 adjoint.
    gpu_tex0 = G.box_chain_adjoint([l.astype(np.float64) for l in gpu["texture_levels"]])
    nz = np.argwhere(np.abs(gpu_tex0).sum(-1) > 0)
    checks = []
    for i in range(64):
        kind = ("positions", "colors", "texcoords", "texture")[i % 4]
        if kind == "texture":
            y, x = nz[rng.integers(len(nz))]
            checks.append((kind, (int(y), int(x), int(rng.integers(0, 4)))))
        elif kind == "positions":
            checks.append((kind, (int(rng.choice(used)), int(rng.choice([0, 1, 3])))))
        else:
            checks.append((kind, (int(rng.choice(used)), int(rng.integers(0, 2 if kind == "texcoords" else 4)))))
    got_all = {"positions": gpu["positions"], "colors": gpu["colors"], "texcoords": gpu["texcoords"],
               "texture": gpu_tex0}
    worst = 0.0
    for kind, idx in checks:
        fd = fd_derivative(view, params, frag, g, kind, idx)
        got = float(got_all[kind][idx])
        scale = np.abs(got_all[kind]).max()
        err = abs(got - fd) / max(abs(fd), 1e-3 * scale)
        worst = max(worst, err)
        assert err < TOL, (kind, idx, got, fd)
    print(f"[gpu] worst relative FD error {worst:.2e}", flush=True)


def test_autograd_path_matches_direct_calls():
    d = float_scene(600, size=96, triangles=300, tex=64)
    view = diff.View(d, DEV)
    pos = _t(d.positions).requires_grad_(True)
    col = _t(d.colors).requires_grad_(True)
    uv = _t(d.texcoords).requires_grad_(True)
    tex0 = _t(d.texture.levels[0]).requires_grad_(True)
    levels = diff.texture_chain(tex0, len(d.texture.levels))
    img = diff.draw(view, pos, col, uv, levels)
    rng = np.random.default_rng(61)
    w = _t(rng.normal(size=img.shape).astype(np.float32))
    (img * w).sum().backward()
    torch.cuda.synchronize()
    _, _, color, ids = gpu_forward(d)
    assert np.array_equal(img.detach().cpu().numpy().view(np.uint32), color.view(np.uint32))
    # Direct entry points with the same upstream gradient.
    d_pos, d_col, d_uv, d_lv = diff.backward_raw(view, pos.detach(), col.detach(), uv.detach(),
                                                 [l.detach() for l in levels], _t(ids.view(np.int32)), w)
    torch.cuda.synchronize()
    for a, b in ((pos.grad, d_pos), (col.grad, d_col), (uv.grad, d_uv)):
        assert gradient_rel_error(a.cpu().numpy(), b.cpu().numpy()) < 1e-6
    want0 = G.box_chain_adjoint([l.cpu().numpy().astype(np.float64) for l in d_lv])
    assert gradient_rel_error(tex0.grad.cpu().numpy(), want0) < 1e-6


@pytest.mark.parametrize("which", ["diff_view_512", "msaa_aniso_view"])
def test_backward_at_full_view_size(which):
    """One view at the sizes the fitting workloads use: 512^2, 20k triangles,
    1024^2 trilinear texture; and a 4x MSAA view with 8x anisotropic
    filtering of a 2048^2 texture.  Every gradient kind within 1e-4 of the
    FP64 oracle; 16 parameters also against central differences."""
    if which == "diff_view_512":
        d = scenes.diff_views(seed=4, views=1)[0]
    else:
        d = scenes.msaa_aniso(seed=5, views=1, tri_lo=20000, tri_hi=20001)[0]
    print(f"[gpu] full-size backward {which} {d.width}x{d.height} s={d.samples} tris={d.triangle_count}", flush=True)
    gpu, frag, g = run_backward(d, 70)
    view = view_of(d)
    params = params_of(d)
    want = G.gradient(view, params, frag, g)
    errs = {k: gradient_rel_error(gpu[k], want[k]) for k in ("positions", "colors", "texcoords")}
    lv = [gradient_rel_error(a, b) for a, b in zip(gpu["texture_levels"], want["texture_levels"])
          if np.abs(b).max() > 0]
    print(f"[gpu] relative errors {errs} texture levels max {max(lv):.2e}", flush=True)
    assert max(errs.values()) < TOL and max(lv) < TOL
    rng = np.random.default_rng(71)
    used = np.unique(G._vertex_ids(view, frag.t, len(d.positions)))
    for i in range(16):
        kind = ("positions", "colors", "texcoords")[i % 3]
        idx = (int(rng.choice(used)), int(rng.choice([0, 1, 3]) if kind == "positions" else rng.integers(0, 2)))
        fd = fd_derivative(view, params, frag, g, kind, idx)
        got = float(gpu[kind][idx])
        assert abs(got - fd) / max(abs(fd), 1e-3 * np.abs(gpu[kind]).max()) < TOL, (kind, idx, got, fd)
