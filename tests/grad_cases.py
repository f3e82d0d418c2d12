"""Scenes and adapters shared by the gradient tests."""
import os
import sys

import numpy as np

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import grad_fp64 as G  # noqa: E402  (the FP64 gradient oracle: test infrastructure)

from swiftshader_amd import abi, scenes  # noqa: E402
from swiftshader_amd.scenes import DrawData, TextureData  # noqa: E402


def view_of(d: DrawData) -> "G.View":
    t = d.texture
    filt, mip, au, av, nlev = G.FILT_LINEAR, G.MIP_LINEAR, G.ADDRESS_REPEAT, G.ADDRESS_REPEAT, 0
    if t is not None:
        if t.anisotropy_enable:
            filt = G.FILT_ANISO
        elif t.mag_filter == abi.FILTER_NEAREST and t.min_filter == abi.FILTER_NEAREST:
            filt = G.FILT_POINT
        mip = G.MIP_LINEAR if t.mipmap_mode == abi.MIPMAP_MODE_LINEAR else G.MIP_POINT
        au, av, nlev = t.address_u, t.address_v, len(t.levels)
    vp = d.viewport[:4] if d.viewport is not None else (0, 0, d.width, d.height)
    topo = d.topology if d.topology in (abi.TOPOLOGY_TRIANGLE_STRIP, abi.TOPOLOGY_TRIANGLE_FAN) else 3
    return G.View(d.width, d.height, tuple(vp), d.shader, filt, mip, au, av, d.indices, nlev, topo,
                  d.provoking_vertex)


def params_of(d: DrawData) -> dict:
    p = {"positions": d.positions.astype(np.float64)}
    if d.colors is not None:
        p["colors"] = d.colors.astype(np.float64)
    if d.texcoords is not None:
        p["texcoords"] = d.texcoords.astype(np.float64)
    if d.texture is not None:
        p["texture"] = d.texture.levels[0].astype(np.float64)
    return p


def float_scene(seed, shader=abi.SHADER_MODULATE, samples=1, aniso=False, address=abi.ADDRESS_REPEAT, size=48,
                triangles=40, tex=32, levels=None, uv_hi=2.0, area_hi=300.0, singular_uv=False):
    """A small RGBA32F draw with a float texture: the differentiable form.
    singular_uv gives every third triangle v = u at its vertices: on a square
    texture its UV Jacobian is exactly singular (det = -0), so the
    anisotropic probe count rounds to INT_MIN and the probe loop runs once."""
    rng = np.random.default_rng(seed)
    pos, uv, col = scenes.screen_triangles(rng, triangles, size, size, area_lo=20.0, area_hi=area_hi, uv_hi=uv_hi)
    if singular_uv:
        uv = uv.reshape(-1, 3, 2).copy()
        uv[::3, :, 1] = uv[::3, :, 0]
        uv = uv.reshape(-1, 2)
    chain = scenes.float_texture(rng, tex)
    if levels:
        chain = chain[:levels]
    t = TextureData(format=abi.FORMAT_R32G32B32A32_SFLOAT, levels=chain, address_u=address, address_v=address,
                    anisotropy_enable=1 if aniso else 0, max_anisotropy=8.0 if aniso else 1.0)
    return DrawData(width=size, height=size, positions=pos, colors=col, texcoords=uv, samples=samples,
                    shader=shader, texture=t if shader != abi.SHADER_INTERPOLATE else None,
                    color_format=abi.FORMAT_R32G32B32A32_SFLOAT, clear_color=(0, 0, 0, 0))


def gradient_rel_error(got: np.ndarray, want: np.ndarray) -> float:
    """||got - want|| / ||want|| (0 when both vanish)."""
    got = np.asarray(got, np.float64).ravel()
    want = np.asarray(want, np.float64).ravel()
    den = np.linalg.norm(want)
    num = np.linalg.norm(got - want)
    return 0.0 if den == 0 and num == 0 else num / max(den, 1e-300)


def fd_derivative(view, params, frag, g, kind, idx, rel_h=1e-6):
    """Central finite difference of G.objective with respect to params[kind][idx].

    The objective is piecewise smooth in vertex parameters: bilinear filtering
    is piecewise linear in texel coordinates, so a step that moves a footprint
    across a texel boundary averages two slopes.  The derivative at the point
    is the limit, so the step is shrunk until two estimates (h and h/4)
    agree.  For a vertex parameter only the fragments of primitives using the
    vertex are evaluated (the rest do not change), which keeps rounding small
    at tiny steps.  Texel parameters enter linearly: one step is exact.
    """
    sub = frag
    if kind != "texture":
        vi = G._vertex_ids(view, frag.t, len(params["positions"]))
        m = (vi == idx[0]).any(axis=1)
        sub = G.Fragments(frag.x[m], frag.y[m], frag.t[m], frag.smask[m], frag.state[m])
        g = g[m]

    def central(h):
        p = {k: v.copy() for k, v in params.items()}
        p[kind][idx] += h
        fp = G.objective(view, p, sub, g)
        p[kind][idx] -= 2 * h
        fm = G.objective(view, p, sub, g)
        return (fp - fm) / (2 * h)

    scale = max(1.0, abs(params[kind][idx]))
    if kind == "texture":
        return central(1e-3 * scale)
    h = rel_h * scale
    prev = central(h)
    for _ in range(6):
        cur = central(h / 4)
        if abs(cur - prev) <= 1e-6 * max(abs(cur), 1e-3):
            return cur
        h /= 10
        prev = central(h)
    return prev
