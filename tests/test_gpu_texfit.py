"""Multi-view fitting on the GPU: the batched, graph-captured step against
per-draw launches.

* the batch plan's backward (one launch over all views) gives the same
  gradients as swr_draw_backward per view (float atomics: equal up to
  summation order);
* a captured step replays to the same loss and gradients as the eager step,
  replay after replay;
* Adam on the texture lowers the loss against targets rendered from a
  second texture.
"""
import numpy as np
import pytest
import torch

from grad_cases import gradient_rel_error
from swiftshader_amd import abi, scenes, texfit
from swiftshader_amd.scenes import TextureData

pytestmark = pytest.mark.gpu
DEV = "cuda:0"


def small_views(seed, samples=1, aniso=False, views=3):
    rng = np.random.default_rng(seed)
    tex = TextureData(format=abi.FORMAT_R32G32B32A32_SFLOAT, levels=scenes.float_texture(rng, 64),
                      anisotropy_enable=1 if aniso else 0, max_anisotropy=8.0 if aniso else 1.0)
    out = []
    for v in range(views):
        w, h = 96 + 16 * v, 80 + 8 * v
        pos, uv, col = scenes.screen_triangles(rng, 400, w, h, area_hi=400)
        out.append(scenes.DrawData(width=w, height=h, positions=pos, colors=col, texcoords=uv, samples=samples,
                                   shader=abi.SHADER_MODULATE, texture=tex,
                                   color_format=abi.FORMAT_R32G32B32A32_SFLOAT))
    return out, tex


def _grads(fit):
    out = [l.detach().cpu().numpy().copy() for l in fit.d_levels]
    if fit.geometry_grads:
        for vb in fit.views:
            out += [vb.d_pos.cpu().numpy().copy(), vb.d_col.cpu().numpy().copy(), vb.d_uv.cpu().numpy().copy()]
    return out


@pytest.mark.parametrize("samples,aniso", [(1, False), (4, True)])
def test_batched_backward_matches_per_view(samples, aniso):
    draws, tex = small_views(31 + samples, samples, aniso)
    rng = np.random.default_rng(3)
    targets = [torch.from_numpy(rng.normal(0.5, 0.2, (d.height, d.width, 4)).astype(np.float32)) for d in draws]
    fit = texfit.MultiViewFit(draws, torch.from_numpy(tex.levels[0]), targets, geometry_grads=True, device=DEV)
    print(f"[gpu texfit] plain vs batched samples={samples} aniso={aniso}", flush=True)
    fit.step_plain(update=False)
    torch.cuda.synchronize()
    loss_p = fit.current_loss()
    gp = _grads(fit)
    imgs_p = [vb.color.cpu().numpy().copy() for vb in fit.views]
    fit.step_batched(update=False)
    torch.cuda.synchronize()
    assert fit.current_loss() == loss_p
    for a, vb in zip(imgs_p, fit.views):
        assert np.array_equal(a.view(np.uint32), vb.color.cpu().numpy().view(np.uint32))
    for a, b in zip(gp, _grads(fit)):
        assert gradient_rel_error(b, a) < 1e-5


def test_graph_replays_match_eager_step():
    draws, tex = small_views(41, 4, True)
    rng = np.random.default_rng(4)
    targets = [torch.from_numpy(rng.normal(0.5, 0.2, (d.height, d.width, 4)).astype(np.float32)) for d in draws]
    fit = texfit.MultiViewFit(draws, torch.from_numpy(tex.levels[0]), targets, geometry_grads=True, device=DEV)
    fit.step_batched(update=False)
    torch.cuda.synchronize()
    loss_e = fit.current_loss()
    ge = _grads(fit)
    fit.capture(update=False)
    for rep in range(3):
        print(f"[gpu texfit] replay {rep}", flush=True)
        (torch.randn(1 << 18, device=DEV) * 3).sum().item()
        fit.step_graph()
        torch.cuda.synchronize()
        assert fit.current_loss() == loss_e
        for a, b in zip(ge, _grads(fit)):
            assert gradient_rel_error(b, a) < 1e-5


def test_adam_fit_lowers_the_loss():
    draws, tex = small_views(51, 1, False, views=4)
    rng = np.random.default_rng(52)
    target_tex = scenes.float_texture(rng, 64)
    ref = texfit.MultiViewFit(draws, torch.from_numpy(target_tex[0]), device=DEV)
    ref.forward_plain()
    torch.cuda.synchronize()
    targets = [vb.color[0].clone() for vb in ref.views]
    fit = texfit.MultiViewFit(draws, torch.from_numpy(tex.levels[0]), targets, lr=2e-2, device=DEV)
    fit.capture(update=True)
    fit.step_graph()
    torch.cuda.synchronize()
    first = fit.current_loss()
    for _ in range(60):
        fit.step_graph()
    torch.cuda.synchronize()
    last = fit.current_loss()
    print(f"[gpu texfit] loss {first:.5f} -> {last:.5f}", flush=True)
    assert last < 0.3 * first


def test_resolve_of_sample_planes_apart_equals_contiguous():
    """runtime.resolve takes its pitches from the tensor: sample planes that lie
    apart (as in MultiViewFit's sample-major layout) resolve bit for bit like a
    contiguous copy."""
    from swiftshader_amd import runtime
    rng = np.random.default_rng(53)
    H, W, off = 37, 29, 1000
    planes = torch.from_numpy(rng.normal(0.5, 0.3, (4, 4 * (off + 3 * H * W))).astype(np.float32)).to(DEV)
    src = planes[:, 4 * off:4 * (off + H * W)].view(4, H, W, 4)
    got = torch.zeros((1, H, W, 4), dtype=torch.float32, device=DEV)
    want = torch.zeros_like(got)
    runtime.resolve(src, got)
    runtime.resolve(src.contiguous(), want)
    torch.cuda.synchronize()
    assert np.array_equal(got.cpu().numpy().view(np.uint32), want.cpu().numpy().view(np.uint32))


def test_msaa_fit_resolves_every_view_in_one_launch():
    """4x views: the fit's one-launch resolve over all views' sample planes
    gives each view the image a per-view resolve of its colour buffer gives,
    and an Adam fit against resolved targets lowers the loss."""
    from swiftshader_amd import runtime
    draws, tex = small_views(54, 4, False, views=4)
    rng = np.random.default_rng(55)
    ref = texfit.MultiViewFit(draws, torch.from_numpy(scenes.float_texture(rng, 64)[0]), device=DEV)
    assert ref.planar
    ref.forward_plain()
    ref._resolve_and_grad()
    targets = []
    for vb in ref.views:
        want = torch.zeros_like(vb.resolved)
        runtime.resolve(vb.color.contiguous(), want)
        torch.cuda.synchronize()
        assert np.array_equal(vb.resolved.cpu().numpy().view(np.uint32), want.cpu().numpy().view(np.uint32))
        targets.append(want[0].clone())
    fit = texfit.MultiViewFit(draws, torch.from_numpy(tex.levels[0]), targets, lr=2e-2, device=DEV)
    fit.capture(update=True)
    fit.step_graph()
    torch.cuda.synchronize()
    first = fit.current_loss()
    for _ in range(60):
        fit.step_graph()
    torch.cuda.synchronize()
    last = fit.current_loss()
    print(f"[gpu texfit] 4x loss {first:.5f} -> {last:.5f}", flush=True)
    assert last < 0.3 * first


def test_checkpoint_resume_continues_the_fit(tmp_path):
    """A fit checkpointed after 5 captured steps and resumed in a fresh
    MultiViewFit (its own graph) holds the saved state exactly, and after 5
    more steps the same texture and Adam state as the uninterrupted fit
    (within float-atomic summation order: the gradients of two runs agree to
    about 1e-6)."""
    draws, tex = small_views(61, 4, True)
    rng = np.random.default_rng(62)
    targets = [torch.from_numpy(rng.normal(0.5, 0.2, (d.height, d.width, 4)).astype(np.float32)) for d in draws]

    def make():
        f = texfit.MultiViewFit(draws, torch.from_numpy(tex.levels[0]), targets, lr=1e-2, device=DEV)
        f.capture(update=True)
        return f

    a = make()
    for _ in range(5):
        a.step_graph()
    torch.cuda.synchronize()
    path = str(tmp_path / "fit.pt")
    a.save(path)
    for _ in range(5):
        a.step_graph()
    torch.cuda.synchronize()
    print("[gpu texfit] resumed from checkpoint", flush=True)
    b = make()
    b.load(path)
    saved = torch.load(path, weights_only=True)
    for x, key in ((b.tex0, "texture"), (b.m, "adam_m"), (b.v, "adam_v"), (b.t, "adam_t")):
        assert torch.equal(x.cpu(), saved[key])
    for _ in range(5):
        b.step_graph()
    torch.cuda.synchronize()
    assert torch.equal(a.t, b.t)
    for x, y in ((a.tex0, b.tex0), (a.m, b.m), (a.v, b.v)):
        assert torch.allclose(x, y, rtol=1e-4, atol=1e-7), float((x - y).abs().max())
    with pytest.raises(ValueError):
        c = texfit.MultiViewFit(draws, torch.zeros((32, 32, 4)), targets, device=DEV)
        c.load(path)
