"""FP64 reference of the draw path's gradients.

TEST INFRASTRUCTURE ONLY: the checker the gradient tests compare the HIP
backward kernels against.  Nothing in the swiftshader_amd package imports it.

The reference has no backward pass.  The function differentiated here is the
fixed-function draw (the shader family of src/Pipeline/PixelProgram.cpp:
138-364, the float texture path of src/Pipeline/SamplerCore.cpp) with what is
discrete held fixed:

  * visibility -- each covered sample's colour comes from the primitive in
    the draw's primitive_id record (1 + primitive index, 0 = none);
  * the sampling state of each fragment -- LOD, anisotropic probe count and
    axis -- taken as an input ("state": float32 (lod, aniso, uDelta, vDelta));
    it picks the mip levels, the trilinear weight (frac of the float32 lod)
    and the probe positions u + (B + i*A) * uDelta with the float32 weights
    A = 1/N, B = -(N-1)/(2N) (Constants.cpp:278-286);
  * the tap indices of each filter footprint (floor of the texel coordinate).

Attributes are interpolated at the pixel centre with perspective-correct
barycentrics of the UNSNAPPED projected vertices X = X0 + x/w * Wh (the
reference snaps X to 28.4, which has zero derivative), and the texture is
filtered with the forward's addressing (REPEAT: u - floor(u);
CLAMP_TO_EDGE: clamp to [0, 1 - 2^-24], zero derivative outside).  Mip
levels 1.. are the 2x2 box filter of level 0, ((a + b) + (c + d)) / 4.

Everything runs in float64.  `shade` is the forward; `gradient` is its
hand-derived adjoint.  The tests pin `gradient` against central finite
differences of `shade` (a dropped term or a wrong index fails them) and
`shade` against the CPU oracle's rendered image on scenes where the 28.4
snap is exact.
"""
from dataclasses import dataclass
from typing import List, Optional

import numpy as np

FILT_POINT, FILT_LINEAR, FILT_ANISO = 0, 3, 4
MIP_POINT, MIP_LINEAR = 1, 2
SHADER_INTERPOLATE, SHADER_SAMPLE, SHADER_MODULATE = 0, 1, 2
ADDRESS_REPEAT, ADDRESS_CLAMP_TO_EDGE = 0, 2


@dataclass
class View:
    """Static part of one draw: viewport, shader, sampler, visibility."""
    width: int
    height: int
    viewport: tuple          # (x, y, w, h) as in the draw descriptor
    shader: int
    filter: int = FILT_LINEAR
    mipmap: int = MIP_LINEAR
    address_u: int = ADDRESS_REPEAT
    address_v: int = ADDRESS_REPEAT
    indices: Optional[np.ndarray] = None   # (T*3,) or None for 0,1,2,...
    level_count: int = 0                   # mip levels in the view (0 = the full chain)
    topology: int = 3                      # VkPrimitiveTopology: 3 list, 4 strip, 5 fan
    provoking_vertex: int = 0              # 0 first, 1 last


@dataclass
class Fragments:
    """One entry per (pixel, owning primitive): samples of a pixel owned by
    the same primitive share a fragment, shaded once at the pixel centre."""
    x: np.ndarray       # (n,) int
    y: np.ndarray       # (n,) int
    t: np.ndarray       # (n,) primitive index
    smask: np.ndarray   # (n, S) bool: the samples the fragment covers
    state: np.ndarray   # (n, 4) float32 lod, aniso, uDelta, vDelta


def fragments(ids: np.ndarray, state: Optional[np.ndarray] = None) -> Fragments:
    """ids: (S, H, W) uint32 primitive_id record; state: (S, H, W, 4) float32."""
    S = ids.shape[0]
    xs, ys, ts, ms, st = [], [], [], [], []
    for q in range(S):
        first = ids[q] != 0
        for q2 in range(q):
            first &= ids[q2] != ids[q]
        y, x = np.nonzero(first)
        t1 = ids[q][y, x]
        m = np.stack([ids[q2][y, x] == t1 if q2 >= q else np.zeros(len(y), bool) for q2 in range(S)], 1)
        xs.append(x)
        ys.append(y)
        ts.append(t1.astype(np.int64) - 1)
        ms.append(m)
        st.append(state[q][y, x] if state is not None else np.zeros((len(y), 4), np.float32))
    return Fragments(np.concatenate(xs).astype(np.int64), np.concatenate(ys).astype(np.int64), np.concatenate(ts),
                     np.concatenate(ms), np.concatenate(st).astype(np.float32))


def upstream(frag: Fragments, d_color: np.ndarray) -> np.ndarray:
    """Per-fragment upstream gradient: the sum of d_color over its samples."""
    g = np.zeros((len(frag.t), 4))
    for q in range(frag.smask.shape[1]):
        g += np.where(frag.smask[:, q:q + 1], d_color[q][frag.y, frag.x].astype(np.float64), 0.0)
    return g


def box_chain(level0: np.ndarray) -> List[np.ndarray]:
    levels = [np.asarray(level0, np.float64)]
    cur = levels[0]
    while cur.shape[0] > 1 or cur.shape[1] > 1:
        h, w = cur.shape[:2]
        nh, nw = max(h // 2, 1), max(w // 2, 1)
        y0 = np.minimum(2 * np.arange(nh), h - 1)
        y1 = np.minimum(2 * np.arange(nh) + 1, h - 1)
        x0 = np.minimum(2 * np.arange(nw), w - 1)
        x1 = np.minimum(2 * np.arange(nw) + 1, w - 1)
        cur = ((cur[y0][:, x0] + cur[y0][:, x1]) + (cur[y1][:, x0] + cur[y1][:, x1])) * 0.25
        levels.append(cur)
    return levels


def box_chain_adjoint(grads: List[np.ndarray]) -> np.ndarray:
    """Adjoint of box_chain: per-level gradients -> gradient of level 0."""
    g = [np.array(x, np.float64) for x in grads]
    for k in range(len(g) - 1, 0, -1):
        h, w = g[k - 1].shape[:2]
        nh, nw = g[k].shape[:2]
        y0 = np.minimum(2 * np.arange(nh), h - 1)
        y1 = np.minimum(2 * np.arange(nh) + 1, h - 1)
        x0 = np.minimum(2 * np.arange(nw), w - 1)
        x1 = np.minimum(2 * np.arange(nw) + 1, w - 1)
        for ys in (y0, y1):
            for xs in (x0, x1):
                np.add.at(g[k - 1], (ys[:, None], xs[None, :]), 0.25 * g[k])
    return g[0]


def chain(view: View, level0: np.ndarray) -> List[np.ndarray]:
    levels = box_chain(level0)
    return levels[:view.level_count] if view.level_count else levels


def _f32(v):
    return np.float32(v)


def _screen_constants(view: View):
    # Renderer.cpp:300-345, in float32 as the descriptor's consumer computes them.
    vx, vy, vw, vh = [_f32(v) for v in view.viewport]
    Wf, Hf = _f32(0.5) * vw, _f32(0.5) * vh
    X0xF = (vx + Wf) * _f32(16) - _f32(8)
    Y0xF = (vy + Hf) * _f32(16) - _f32(8)
    WxF, HxF = Wf * _f32(16), Hf * _f32(16)
    return float(X0xF) / 16, float(Y0xF) / 16, float(WxF) / 16, float(HxF) / 16


def _vertex_ids(view: View, t: np.ndarray, V: int):
    """Vertex indices of primitives t (setBatchIndices, Renderer.cpp:50-145)."""
    t = np.asarray(t, np.int64)
    first = view.provoking_vertex == 0
    if view.topology == 4:      # strip
        odd = t & 1
        p = np.stack([t + (0 if first else 2), t + odd + (1 if first else 0), t + (1 - odd) + (1 if first else 0)], 1)
    elif view.topology == 5:    # fan
        a, b, z = t + 1, t + 2, np.zeros_like(t)
        p = np.stack([a, b, z], 1) if first else np.stack([b, z, a], 1)
    else:                       # list
        o = [0, 1, 2] if first else [2, 0, 1]
        p = 3 * t[:, None] + np.array(o)[None, :]
    vi = p if view.indices is None else np.asarray(view.indices, np.int64)[p]
    return np.minimum(vi, V - 1)


class _Tex:
    """Flattened mip chain (float64) with per-level offsets."""

    def __init__(self, levels: List[np.ndarray]):
        self.w = np.array([l.shape[1] for l in levels])
        self.h = np.array([l.shape[0] for l in levels])
        self.off = np.concatenate([[0], np.cumsum(self.w * self.h)])[:-1]
        self.flat = np.concatenate([l.reshape(-1, 4) for l in levels]).astype(np.float64)
        self.n = len(levels)


def _axis(filt, uvw, dim, mode):
    """One axis of the float-path address (SamplerCore.cpp:2463-2638):
    tap indices, fraction and d fraction / d uvw.  `dim` may vary per entry."""
    if mode == ADDRESS_CLAMP_TO_EDGE:
        hi = 1.0 - 2.0 ** -24
        c = np.clip(uvw, 0.0, hi)
        dc = ((uvw > 0.0) & (uvw <= hi)).astype(np.float64)
    else:
        c = uvw - np.floor(uvw)
        dc = np.ones_like(uvw)
    coord = c * dim
    if filt == FILT_POINT:
        x0 = np.minimum(coord.astype(np.int64), dim - 1)
        z = np.zeros_like(uvw)
        return x0, x0, z, z
    coord = coord - 0.5
    fl = np.floor(coord)
    x0 = fl.astype(np.int64)
    f = coord - fl
    x1 = x0 + 1
    if mode == ADDRESS_CLAMP_TO_EDGE:
        x0 = np.maximum(x0, 0)
        x1 = np.minimum(x1, dim - 1)
    else:
        x0 = np.where(x0 < 0, dim - 1, x0)
        x1 = np.where(x1 >= dim, 0, x1)
    return x0, x1, f, dc * dim


def _taps(view: View, tex: _Tex, u, v, state):
    """Yields (weight, level, tap indices into tex.flat, tap weights, dfu, dfv, fu, fv) per probe/level."""
    lod = state[:, 0].astype(np.float32)
    n = len(u)
    if view.filter == FILT_ANISO:
        N = np.maximum(np.rint(state[:, 1].astype(np.float32)).astype(np.int64), 1)
        A = np.where(N > 1, (np.float32(1) / N.astype(np.float32)), np.float32(1)).astype(np.float64)
        B = np.where(N > 1, -(N - 1).astype(np.float32) / (2 * N).astype(np.float32), np.float32(0)).astype(np.float64)
    else:
        N = np.ones(n, np.int64)
        A = np.ones(n)
        B = np.zeros(n)
    if view.mipmap == MIP_LINEAR:
        l0 = np.trunc(lod).astype(np.int64)
        fr = (lod - np.floor(lod)).astype(np.float32)
        fr = np.minimum(fr, np.float32(np.frombuffer(np.uint32(0x3F7FFFFF).tobytes(), np.float32)[0]))
        wl = [(l0, 1.0 - fr.astype(np.float64)), (l0 + 1, fr.astype(np.float64))]
    else:
        wl = [(np.rint(lod).astype(np.int64), np.ones(n))]
    du = state[:, 2].astype(np.float64)
    dv = state[:, 3].astype(np.float64)
    for i in range(int(N.max()) if n else 0):
        act = i < N
        tt = B + i * A
        ui = u + tt * du
        vi = v + tt * dv
        for lvl, w in wl:
            lvl = np.minimum(lvl, tex.n - 1)
            W = tex.w[lvl]
            H = tex.h[lvl]
            x0, x1, fu, dfu = _axis(view.filter, ui, W, view.address_u)
            y0, y1, fv, dfv = _axis(view.filter, vi, H, view.address_v)
            base = tex.off[lvl]
            idx = np.stack([base + y0 * W + x0, base + y0 * W + x1, base + y1 * W + x0, base + y1 * W + x1], 1)
            tw = np.stack([(1 - fu) * (1 - fv), fu * (1 - fv), (1 - fu) * fv, fu * fv], 1)
            yield np.where(act, A * w, 0.0), idx, tw, dfu, dfv, fu, fv


def _interp(view: View, params: dict, frag: Fragments):
    """Perspective-correct barycentrics at the pixel centres, and friends."""
    pos = np.asarray(params["positions"], np.float64)
    vi = _vertex_ids(view, frag.t, len(pos))
    X0, Y0, Wh, Hh = _screen_constants(view)
    P = pos[vi]                                    # (n, 3, 4)
    x, y, w = P[..., 0], P[..., 1], P[..., 3]
    rhw = np.where(w != 0, 1.0 / np.where(w != 0, w, 1.0), 1.0)
    X = X0 + x * rhw * Wh
    Y = Y0 + y * rhw * Hh
    D = (Y[:, 1] - Y[:, 2]) * (X[:, 0] - X[:, 2]) + (X[:, 2] - X[:, 1]) * (Y[:, 0] - Y[:, 2])
    gx = np.stack([Y[:, 1] - Y[:, 2], Y[:, 2] - Y[:, 0], Y[:, 0] - Y[:, 1]], 1) / D[:, None]
    gy = np.stack([X[:, 2] - X[:, 1], X[:, 0] - X[:, 2], X[:, 1] - X[:, 0]], 1) / D[:, None]
    ex = frag.x - X[:, 2]
    ey = frag.y - Y[:, 2]
    l0 = gx[:, 0] * ex + gy[:, 0] * ey
    l1 = gx[:, 1] * ex + gy[:, 1] * ey
    lam = np.stack([l0, l1, 1.0 - l0 - l1], 1)
    r = lam * rhw
    S = r.sum(1)
    b = r / S[:, None]
    return dict(vi=vi, x=x, y=y, w=w, rhw=rhw, lam=lam, gx=gx, gy=gy, S=S, b=b, Wh=Wh, Hh=Hh)


def shade(view: View, params: dict, frag: Fragments) -> np.ndarray:
    """Fragment colours (n, 4), float64."""
    I = _interp(view, params, frag)
    b, vi = I["b"], I["vi"]
    n = len(frag.t)
    col = np.ones((n, 4))
    if view.shader != SHADER_SAMPLE:
        col = np.einsum("nk,nkc->nc", b, np.asarray(params["colors"], np.float64)[vi])
    if view.shader == SHADER_INTERPOLATE:
        return col
    uv = np.einsum("nk,nkc->nc", b, np.asarray(params["texcoords"], np.float64)[vi])
    tex = _Tex(chain(view, params["texture"]))
    val = np.zeros((n, 4))
    for wgt, idx, tw, _, _, _, _ in _taps(view, tex, uv[:, 0], uv[:, 1], frag.state):
        val += wgt[:, None] * np.einsum("nj,njc->nc", tw, tex.flat[idx])
    return val * col if view.shader == SHADER_MODULATE else val


def objective(view: View, params: dict, frag: Fragments, g: np.ndarray) -> float:
    """The linear functional sum_frag g . colour, whose gradient is J^T g."""
    return float(np.sum(shade(view, params, frag) * g))


def gradient(view: View, params: dict, frag: Fragments, g: np.ndarray) -> dict:
    """Adjoint of `shade`: gradients of sum_frag g . colour."""
    I = _interp(view, params, frag)
    b, vi = I["b"], I["vi"]
    pos = np.asarray(params["positions"], np.float64)
    n = len(frag.t)
    out = {"positions": np.zeros_like(pos)}
    Gk = np.zeros((n, 3))
    col = np.ones((n, 4))
    if view.shader != SHADER_SAMPLE:
        C = np.asarray(params["colors"], np.float64)
        col = np.einsum("nk,nkc->nc", b, C[vi])
    dcol = g.copy()
    if view.shader != SHADER_INTERPOLATE:
        T = np.asarray(params["texcoords"], np.float64)
        uv = np.einsum("nk,nkc->nc", b, T[vi])
        tex = _Tex(chain(view, params["texture"]))
        gtex = g * col if view.shader == SHADER_MODULATE else g
        val = np.zeros((n, 4))
        dU = np.zeros(n)
        dV = np.zeros(n)
        gflat = np.zeros_like(tex.flat)
        for wgt, idx, tw, dfu, dfv, fu, fv in _taps(view, tex, uv[:, 0], uv[:, 1], frag.state):
            c = tex.flat[idx]                                   # (n, 4 taps, 4)
            val += wgt[:, None] * np.einsum("nj,njc->nc", tw, c)
            d = np.einsum("nc,njc->nj", gtex, c)                # g . texel per tap
            dU += wgt * dfu * ((1 - fv) * (d[:, 1] - d[:, 0]) + fv * (d[:, 3] - d[:, 2]))
            dV += wgt * dfv * ((1 - fu) * (d[:, 2] - d[:, 0]) + fu * (d[:, 3] - d[:, 1]))
            contrib = (wgt[:, None] * tw)[:, :, None] * gtex[:, None, :]
            for c4 in range(4):
                gflat[:, c4] += np.bincount(idx.reshape(-1), contrib[..., c4].reshape(-1), minlength=len(gflat))
        per_level = [gflat[o:o + h * w].reshape(h, w, 4) for o, h, w in zip(tex.off, tex.h, tex.w)]
        out["texture_levels"] = per_level
        out["texture"] = box_chain_adjoint(per_level)
        dcol = g * val if view.shader == SHADER_MODULATE else None
        duv = np.stack([b * dU[:, None], b * dV[:, None]], -1)       # (n, 3, 2)
        out["texcoords"] = np.zeros_like(T)
        np.add.at(out["texcoords"], vi.reshape(-1), duv.reshape(-1, 2))
        Gk += dU[:, None] * T[vi][..., 0] + dV[:, None] * T[vi][..., 1]
    if dcol is not None:
        C = np.asarray(params["colors"], np.float64)
        out["colors"] = np.zeros_like(C)
        np.add.at(out["colors"], vi.reshape(-1), (b[..., None] * dcol[:, None, :]).reshape(-1, 4))
        Gk += np.einsum("nc,nkc->nk", dcol, C[vi])
    # b = r / sum r, r_k = lambda_k * rhw_k; d lambda_i / d X_j = -lambda_j * d lambda_i / d px.
    lam, rhw, S = I["lam"], I["rhw"], I["S"]
    Abar = (b * Gk).sum(1)
    dr = (Gk - Abar[:, None]) / S[:, None]
    dlam = dr * rhw
    drhw = dr * lam
    Gx = (dlam * I["gx"]).sum(1)
    Gy = (dlam * I["gy"]).sum(1)
    dX = -lam * Gx[:, None]
    dY = -lam * Gy[:, None]
    dx = dX * rhw * I["Wh"]
    dy = dY * rhw * I["Hh"]
    drhw = drhw + dX * I["x"] * I["Wh"] + dY * I["y"] * I["Hh"]
    dw = np.where(I["w"] != 0, -drhw * rhw * rhw, 0.0)
    dp = np.zeros((n, 3, 4))
    dp[..., 0] = dx
    dp[..., 1] = dy
    dp[..., 3] = dw
    np.add.at(out["positions"], vi.reshape(-1), dp.reshape(-1, 4))
    return out
