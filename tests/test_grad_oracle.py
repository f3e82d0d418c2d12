"""Pins the FP64 gradient oracle (oracle/grad_fp64.py).

* `gradient` (hand-derived adjoint) against central finite differences of
  `shade` (its forward) on random parameters of every kind -- positions x, y,
  z, w, colours, texcoords and level-0 texels -- for the three shaders, 1x and
  4x, REPEAT and CLAMP_TO_EDGE, trilinear and 8x anisotropic;
* the box-chain adjoint by the inner-product identity <Mx, y> = <x, M^T y>;
* `shade` against the CPU oracle's rendered image on scenes whose vertices
  sit exactly on the 28.4 grid (so the reference's snap changes nothing)
  and whose texture has one level (LOD fixed at 0).
"""
import numpy as np
import pytest

from grad_cases import G, fd_derivative, float_scene, params_of, view_of
from oracle_runner import oracle_render_ids
from swiftshader_amd import abi, scenes
from swiftshader_amd.scenes import DrawData, TextureData


def random_state(rng, n, aniso):
    st = np.zeros((n, 4), np.float32)
    st[:, 0] = rng.uniform(0.0, 2.5, n)
    if aniso:
        st[:, 1] = rng.uniform(1.0, 6.0, n)
        st[:, 2] = rng.uniform(-0.05, 0.05, n)
        st[:, 3] = rng.uniform(-0.05, 0.05, n)
    return st


CASES = [
    dict(shader=abi.SHADER_MODULATE),
    dict(shader=abi.SHADER_MODULATE, samples=4),
    dict(shader=abi.SHADER_MODULATE, aniso=True),
    dict(shader=abi.SHADER_MODULATE, aniso=True, address=abi.ADDRESS_CLAMP_TO_EDGE, samples=4),
    dict(shader=abi.SHADER_SAMPLE, address=abi.ADDRESS_CLAMP_TO_EDGE),
    dict(shader=abi.SHADER_INTERPOLATE),
]


@pytest.mark.parametrize("case", range(len(CASES)))
def test_gradient_matches_central_differences(case):
    kw = CASES[case]
    d = float_scene(100 + case, **kw)
    _, _, ids, _ = oracle_render_ids(d)
    frag = G.fragments(ids)
    rng = np.random.default_rng(7 + case)
    frag.state = random_state(rng, len(frag.t), kw.get("aniso", False))
    view = view_of(d)
    params = params_of(d)
    g = rng.normal(size=(len(frag.t), 4))
    an = G.gradient(view, params, frag, g)
    used = np.unique(G._vertex_ids(view, frag.t, len(d.positions)))
    checks = []
    for _ in range(12):
        checks.append(("positions", (int(rng.choice(used)), int(rng.choice([0, 1, 2, 3])))))
    if view.shader != G.SHADER_SAMPLE:
        for _ in range(6):
            checks.append(("colors", (int(rng.choice(used)), int(rng.integers(0, 4)))))
    if view.shader != G.SHADER_INTERPOLATE:
        for _ in range(8):
            checks.append(("texcoords", (int(rng.choice(used)), int(rng.integers(0, 2)))))
        nz = np.argwhere(np.abs(an["texture"]).sum(-1) > 0)
        for k in rng.choice(len(nz), 10):
            checks.append(("texture", (int(nz[k][0]), int(nz[k][1]), int(rng.integers(0, 4)))))
    scale = {k: np.abs(v).max() for k, v in an.items() if k != "texture_levels"}
    for name, idx in checks:
        fd = fd_derivative(view, params, frag, g, name, idx)
        a = an[name][idx]
        assert abs(fd - a) <= 1e-5 * max(abs(a), 1e-3 * scale[name]) + 1e-9, (name, idx, fd, a)
    # z reaches no colour.
    assert np.all(an["positions"][:, 2] == 0)


@pytest.mark.parametrize("shape", [(16, 16), (12, 20), (5, 7), (1, 9), (8, 1)])
def test_box_chain_adjoint(shape):
    rng = np.random.default_rng(3)
    x = rng.normal(size=shape + (4,))
    levels = G.box_chain(x)
    ys = [rng.normal(size=l.shape) for l in levels]
    lhs = sum(float(np.sum(l * y)) for l, y in zip(levels, ys))
    rhs = float(np.sum(x * G.box_chain_adjoint(ys)))
    assert abs(lhs - rhs) <= 1e-12 * max(1.0, abs(lhs))


def test_box_chain_matches_float32_generator():
    rng = np.random.default_rng(4)
    lvl = rng.normal(0.5, 0.15, (20, 12, 4)).astype(np.float32)
    a = G.box_chain(lvl)
    b = scenes.box_mip_chain(lvl)
    assert len(a) == len(b)
    for x, y in zip(a, b):
        assert np.max(np.abs(x - y)) < 1e-6


def _grid_scene(seed, samples, shader):
    """Vertices on the 28.4 grid with w in {0.5, 1, 2} (exact projection) and
    a single-level float texture (the LOD is fixed at 0, as the reference
    does for one-level views)."""
    rng = np.random.default_rng(seed)
    size = 32
    n = 24
    pts = rng.integers(16, (size - 1) * 16, (n * 3, 2)) / 16.0      # screen positions, 1/16 steps
    w = rng.choice([0.5, 1.0, 2.0], n * 3)
    ndc = (pts - (size / 2 - 0.5)) / (size / 2)
    z = rng.uniform(0.1, 0.9, n * 3)
    pos = np.stack([ndc[:, 0] * w, ndc[:, 1] * w, z * w, w], 1).astype(np.float32)
    assert np.all(pos[:, 0].astype(np.float64) / pos[:, 3] == ndc[:, 0])
    col = rng.uniform(0.2, 1.0, (n * 3, 4)).astype(np.float32)
    uv = rng.uniform(-0.5, 1.5, (n * 3, 2)).astype(np.float32)
    tex = TextureData(format=abi.FORMAT_R32G32B32A32_SFLOAT, levels=scenes.float_texture(rng, 16)[:1])
    return DrawData(width=size, height=size, positions=pos, colors=col, texcoords=uv, samples=samples,
                    shader=shader, texture=tex, color_format=abi.FORMAT_R32G32B32A32_SFLOAT,
                    clear_color=(-1, -1, -1, -1))


@pytest.mark.parametrize("samples,shader", [(1, abi.SHADER_MODULATE), (4, abi.SHADER_MODULATE),
                                            (1, abi.SHADER_INTERPOLATE), (1, abi.SHADER_SAMPLE)])
def test_shade_matches_oracle_image_where_snapping_is_exact(samples, shader):
    d = _grid_scene(21 + samples + shader, samples, shader)
    color, _, ids, _ = oracle_render_ids(d)
    frag = G.fragments(ids)
    assert len(frag.t) > 200
    got = G.shade(view_of(d), params_of(d), frag)
    for q in range(samples):
        sel = frag.smask[:, q]
        want = color[q][frag.y[sel], frag.x[sel]].astype(np.float64)
        assert np.max(np.abs(got[sel] - want)) < 2e-5


def _grid_scene_mips(seed, samples, aniso, address=abi.ADDRESS_REPEAT, size=64, tex=64, uv_hi=6.0):
    """_grid_scene with a full mip chain, texture coordinates spread so that
    the LOD varies over several levels, and optional 8x anisotropy."""
    rng = np.random.default_rng(seed)
    n = 40
    pts = rng.integers(16, (size - 1) * 16, (n * 3, 2)) / 16.0
    w = rng.choice([0.5, 1.0, 2.0, 4.0], n * 3)
    ndc = (pts - (size / 2 - 0.5)) / (size / 2)
    z = rng.uniform(0.1, 0.9, n * 3)
    pos = np.stack([ndc[:, 0] * w, ndc[:, 1] * w, z * w, w], 1).astype(np.float32)
    assert np.all(pos[:, 0].astype(np.float64) / pos[:, 3] == ndc[:, 0])
    col = rng.uniform(0.2, 1.0, (n * 3, 4)).astype(np.float32)
    uv = rng.uniform(-2.0, uv_hi, (n * 3, 2)).astype(np.float32)
    t = TextureData(format=abi.FORMAT_R32G32B32A32_SFLOAT, levels=scenes.float_texture(rng, tex),
                    address_u=address, address_v=address, anisotropy_enable=1 if aniso else 0,
                    max_anisotropy=8.0 if aniso else 1.0)
    return DrawData(width=size, height=size, positions=pos, colors=col, texcoords=uv, samples=samples,
                    shader=abi.SHADER_MODULATE, texture=t, color_format=abi.FORMAT_R32G32B32A32_SFLOAT,
                    clear_color=(-1, -1, -1, -1))


STATE_CASES = [dict(samples=1, aniso=False), dict(samples=1, aniso=True), dict(samples=4, aniso=False),
               dict(samples=4, aniso=True, address=abi.ADDRESS_CLAMP_TO_EDGE)]


@pytest.mark.parametrize("case", range(len(STATE_CASES)))
def test_shade_at_the_forward_sampling_state_matches_the_oracle_image(case):
    """The FP64 model evaluated at the sampling state the oracle's forward
    used for each fragment (trilinear, 8x anisotropic, 4x MSAA) reproduces
    the oracle's image where the 28.4 snap is exact: the gradient is that
    of the function the forward renders."""
    from oracle_runner import oracle_render_state
    d = _grid_scene_mips(70 + case, **STATE_CASES[case])
    color, _, ids, state, _ = oracle_render_state(d)
    frag = G.fragments(ids, state)
    assert len(frag.t) > 500
    lods = frag.state[:, 0]
    assert lods.max() - lods.min() > 1.5  # several levels in play
    if STATE_CASES[case]["aniso"]:
        assert (np.rint(frag.state[:, 1]) >= 2).mean() > 0.2
    got = G.shade(view_of(d), params_of(d), frag)
    for q in range(d.samples):
        sel = frag.smask[:, q]
        want = color[q][frag.y[sel], frag.x[sel]].astype(np.float64)
        assert np.max(np.abs(got[sel] - want)) < 2e-5
