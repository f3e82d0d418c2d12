"""Multi-view differentiable rendering and texture fitting.

A set of views (draws into RGBA32F targets, optionally 4x MSAA resolved by
averaging) share one float texture whose level 0 is the parameter; its mip
chain is the 2x2 box filter of level 0.  One step:

  mips -> clear -> forward of every view -> resolve -> L2 loss against the
  target images -> backward of every view (texel gradients through the
  filter footprint into every level, optionally vertex gradients) -> mip
  adjoint into level 0 -> Adam update of level 0

Two ways to run the same kernels:
  plain     one launch sequence per draw (swr_draw / swr_draw_backward per
            view, each draw waiting on its own pair count), no graph;
  batched   one batch plan for all views (one persistent raster launch and
            one backward launch per segment), the whole step captured into
            one HIP graph and replayed.
Everything stays on the device (Adam's step count is a device tensor), so a
captured step needs no host value; the loss is computed on request.
save() / load() checkpoint the texture and Adam's state, so that a fit can
stop and resume where it was (also under a captured graph).
"""
import ctypes
from dataclasses import dataclass
from typing import List, Optional

import torch

from . import abi, diff, runtime
from .batch import BatchPlan, grad_desc
from .desc import Buffers, build_desc
from .scenes import DrawData


def _flat(shapes, dtype, device):
    """One zeroed flat tensor and contiguous views of it with the given shapes."""
    sizes = [int(torch.Size(sh).numel()) for sh in shapes]
    flat = torch.zeros(sum(sizes), dtype=dtype, device=device)
    out, off = [], 0
    for sh, n in zip(shapes, sizes):
        out.append(flat[off:off + n].view(sh))
        off += n
    return flat, out


def _planes(samples: int, shapes, channels: int, dtype, device, align: int = 4096):
    """A zeroed (samples, N) tensor and, per (H, W) shape, a (samples, H, W[,
    channels]) view of it: every view's sample plane q lies in row q, so an
    elementwise pass over sample q of all views is one launch.  N is the
    total pixel count rounded up to `align` pixels, times channels."""
    pix = [H * W for H, W in shapes]
    total = -(-sum(pix) // align) * align
    flat = torch.zeros((samples, total * channels), dtype=dtype, device=device)
    out, off = [], 0
    for (H, W), n in zip(shapes, pix):
        v = flat[:, off * channels:(off + n) * channels]
        out.append(v.view(samples, H, W, channels) if channels > 1 else v.view(samples, H, W))
        off += n
    return flat, out, total


@dataclass
class _ViewBuffers:
    draw: DrawData
    pos: torch.Tensor
    col: torch.Tensor
    uv: torch.Tensor
    color: torch.Tensor      # (S, H, W, 4)
    depth: torch.Tensor      # (S, H, W)
    ids: torch.Tensor        # (S, H, W) int32
    resolved: torch.Tensor   # (1, H, W, 4)
    target: torch.Tensor     # (1, H, W, 4)
    d_color: torch.Tensor    # (S, H, W, 4)
    d_pos: Optional[torch.Tensor]
    d_col: Optional[torch.Tensor]
    d_uv: Optional[torch.Tensor]
    desc: abi.DrawDesc = None


class MultiViewFit:
    def __init__(self, draws: List[DrawData], texture0: torch.Tensor, targets: Optional[List[torch.Tensor]] = None,
                 lr: float = 1e-2, geometry_grads: bool = False, device="cuda:0"):
        self.device = torch.device(device)
        self.geometry_grads = geometry_grads
        self.lr = lr
        self.tex0 = texture0.detach().to(self.device, torch.float32).contiguous().clone()
        self.levels = [self.tex0] + diff.mip_levels_like(self.tex0)
        # Every per-view and per-level buffer is a slice of one flat tensor per
        # kind, so a clear or an elementwise update of all views is one launch.
        self._d_levels_flat, self.d_levels = _flat([tuple(l.shape) for l in self.levels], torch.float32, self.device)
        self.m = torch.zeros_like(self.tex0)
        self.v = torch.zeros_like(self.tex0)
        self.t = torch.zeros((), dtype=torch.float32, device=self.device)
        self._b1 = torch.full((), 0.9, dtype=torch.float32, device=self.device)
        self._b2 = torch.full((), 0.999, dtype=torch.float32, device=self.device)
        self.views: List[_ViewBuffers] = []
        self.n_values = 0
        for d in draws:
            if d.color_format != abi.FORMAT_R32G32B32A32_SFLOAT or d.blend_enable or d.texture is None:
                raise ValueError("views render a float texture into RGBA32F targets without blending")
        dev = self.device
        samples = [(d.samples, d.height, d.width) for d in draws]
        self.single_sample = all(d.samples == 1 for d in draws)
        # Views that all have 4 samples keep their sample planes sample-major
        # across views, so that the resolve, the loss gradient and its spread
        # over the samples each take one launch for all views.
        self.planar = all(d.samples == 4 for d in draws)
        if self.planar:
            hw = [(H, W) for _, H, W in samples]
            self._color_flat, colors, self._plane_pixels = _planes(4, hw, 4, torch.float32, dev)
            self._depth_flat, depths, _ = _planes(4, hw, 1, torch.float32, dev)
            self._ids_flat, ids, _ = _planes(4, hw, 1, torch.int32, dev)
            self._dcolor_flat, dcolors, _ = _planes(4, hw, 4, torch.float32, dev)
            self._resolved_flat, resolved, _ = _planes(1, hw, 4, torch.float32, dev)
            self._target_flat, tgts, _ = _planes(1, hw, 4, torch.float32, dev)
            self._g_flat = torch.zeros_like(self._resolved_flat)
        else:
            self._color_flat, colors = _flat([(S, H, W, 4) for S, H, W in samples], torch.float32, dev)
            self._depth_flat, depths = _flat(samples, torch.float32, dev)
            self._ids_flat, ids = _flat(samples, torch.int32, dev)
            self._dcolor_flat, dcolors = _flat([(S, H, W, 4) for S, H, W in samples], torch.float32, dev)
            self._resolved_flat, resolved = _flat([(1, H, W, 4) for _, H, W in samples], torch.float32, dev)
            self._target_flat, tgts = _flat([(1, H, W, 4) for _, H, W in samples], torch.float32, dev)
        geo = [None] * len(draws), [None] * len(draws), [None] * len(draws)
        self._geo_flat = []
        if geometry_grads:
            vcount = [len(d.positions) for d in draws]
            fp, dp = _flat([(n, 4) for n in vcount], torch.float32, dev)
            fc, dc = _flat([(n, 4) for n in vcount], torch.float32, dev)
            fu, du = _flat([(n, 2) for n in vcount], torch.float32, dev)
            self._geo_flat = [fp, fc, fu]
            geo = dp, dc, du
        self._depth_flat.fill_(1.0)
        for i, d in enumerate(draws):
            H, W = d.height, d.width
            if targets is not None:
                tgts[i].copy_(targets[i].to(dev).reshape(1, H, W, 4))
            vb = _ViewBuffers(
                draw=d,
                pos=torch.from_numpy(d.positions).to(dev), col=torch.from_numpy(d.colors).to(dev),
                uv=torch.from_numpy(d.texcoords).to(dev),
                color=colors[i], depth=depths[i], ids=ids[i], resolved=resolved[i], target=tgts[i],
                d_color=dcolors[i], d_pos=geo[0][i], d_col=geo[1][i], d_uv=geo[2][i])
            vb.desc = self._desc(vb)
            self.views.append(vb)
            self.n_values += H * W * 4
        self.plan: Optional[BatchPlan] = None
        self.graph = None

    # ---- descriptors -------------------------------------------------------
    def _desc(self, vb: _ViewBuffers) -> abi.DrawDesc:
        d = vb.draw
        S, H, W = d.samples, d.height, d.width
        buf = Buffers(positions=vb.pos.data_ptr(), colors=vb.col.data_ptr(), texcoords=vb.uv.data_ptr())
        buf.texture_levels = [(l.data_ptr(), l.shape[1], l.shape[0], l.shape[1] * 16) for l in self.levels]
        buf.color, buf.color_pitch, buf.color_sample_pitch = vb.color.data_ptr(), W * 16, vb.color.stride(0) * 4
        buf.depth, buf.depth_pitch, buf.depth_sample_pitch = vb.depth.data_ptr(), W * 4, vb.depth.stride(0) * 4
        desc = build_desc(d, buf)
        desc.primitive_id = abi.Attachment(vb.ids.data_ptr(), abi.FORMAT_R32_UINT, W * 4, vb.ids.stride(0) * 4)
        return desc

    def _grad(self, vb: _ViewBuffers) -> abi.GradDesc:
        return grad_desc(vb.d_color, vb.d_pos, vb.d_col, vb.d_uv, self.d_levels)

    def make_plan(self, autotune: bool = False, headroom_percent: int = 10):
        self.plan = BatchPlan([vb.desc for vb in self.views], headroom_percent)
        if autotune:
            self.tuning = self.plan.autotune()
        self.plan.set_grads([self._grad(vb) for vb in self.views])

    # ---- pieces of a step -----------------------------------------------------
    def _clear(self):
        self._color_flat.zero_()
        self._depth_flat.fill_(1.0)
        self._ids_flat.zero_()

    def _image(self, vb: _ViewBuffers) -> torch.Tensor:
        return vb.resolved if vb.draw.samples > 1 else vb.color

    def _resolve_and_grad(self):
        """Resolve, and dL/d(sample colour) of L = mean over all views' values
        of (image - target)^2 (elementwise ops only: no reductions, whose
        library kernels may record memsets into a captured graph)."""
        inv = 1.0 / self.n_values
        if self.single_sample:
            # Images are the colour buffers themselves, laid out like the targets.
            torch.sub(self._color_flat, self._target_flat, out=self._dcolor_flat)
            self._dcolor_flat.mul_(2.0 * inv)
            return
        if self.planar:
            # All views at once: the sample planes as one image 4096 pixels wide.
            rows = self._plane_pixels // 4096
            r = abi.ResolveDesc(4096, rows, abi.FORMAT_R32G32B32A32_SFLOAT, self._color_flat.data_ptr(), 4096 * 16,
                                self._color_flat.stride(0) * 4, 4, self._resolved_flat.data_ptr(), 4096 * 16)
            runtime.check(runtime.lib().swr_resolve(ctypes.byref(r), runtime._stream_handle(None)), "swr_resolve")
            torch.sub(self._resolved_flat, self._target_flat, out=self._g_flat)
            self._g_flat.mul_(2.0 * inv)
            torch.mul(self._g_flat.expand(4, -1), 0.25, out=self._dcolor_flat)
            return
        for vb in self.views:
            if vb.draw.samples > 1:
                runtime.resolve(vb.color, vb.resolved)
            g = (self._image(vb) - vb.target) * (2.0 * inv)
            if vb.draw.samples > 1:
                vb.d_color.copy_((g * (1.0 / vb.draw.samples)).expand_as(vb.d_color))
            else:
                vb.d_color.copy_(g)

    def current_loss(self) -> float:
        """L of the images the last step rendered (computed eagerly)."""
        total = 0.0
        for vb in self.views:
            dv = (self._image(vb) - vb.target).double()
            total += float((dv * dv).sum().item())
        return total / self.n_values

    def _zero_grads(self):
        self._d_levels_flat.zero_()
        for t in self._geo_flat:
            t.zero_()

    def _mip_adjoint(self):
        t = diff._texture_struct(self.d_levels, abi.FORMAT_R32G32B32A32_SFLOAT)
        runtime.check(runtime.lib().swr_mip_backward(ctypes.byref(t), runtime._stream_handle(None)), "swr_mip_backward")

    def _adam(self, eps=1e-8):
        """Adam (beta1 0.9, beta2 0.999) on level 0, bias corrections from the device step count."""
        g = self.d_levels[0]
        self.t.add_(1.0)
        self.m.mul_(0.9).add_(g, alpha=0.1)
        self.v.mul_(0.999).addcmul_(g, g, value=0.001)
        c1 = 1.0 - torch.pow(self._b1, self.t)
        c2 = 1.0 - torch.pow(self._b2, self.t)
        step = (self.m / c1) / (torch.sqrt(self.v / c2) + eps)
        self.tex0.sub_(step * self.lr)

    # ---- steps -----------------------------------------------------------------
    def forward_plain(self):
        diff.generate_mips(self.levels)
        self._clear()
        lib = runtime.lib()
        for vb in self.views:
            runtime.check(lib.swr_draw(ctypes.byref(vb.desc), runtime._stream_handle(None)), "swr_draw")

    def backward_plain(self):
        lib = runtime.lib()
        self._zero_grads()
        for vb in self.views:
            g = self._grad(vb)
            runtime.check(lib.swr_draw_backward(ctypes.byref(vb.desc), ctypes.byref(g), runtime._stream_handle(None)),
                          "swr_draw_backward")
        self._mip_adjoint()

    def step_plain(self, update: bool = True):
        self.forward_plain()
        self._resolve_and_grad()
        self.backward_plain()
        if update:
            self._adam()

    def step_batched(self, update: bool = True):
        """One step through the batch plan (eager launches of the plan)."""
        if self.plan is None:
            self.make_plan()
        diff.generate_mips(self.levels)
        self._clear()
        self.plan.run()
        self._resolve_and_grad()
        self._zero_grads()
        self.plan.backward()
        self._mip_adjoint()
        if update:
            self._adam()

    def capture(self, update: bool = True):
        """Captures one batched step into a HIP graph (after one eager warm-up step)."""
        if self.plan is None:
            self.make_plan()
        s = torch.cuda.Stream(device=self.device)
        s.wait_stream(torch.cuda.current_stream(self.device))
        with torch.cuda.stream(s):
            self.step_batched(update=False)
        torch.cuda.current_stream(self.device).wait_stream(s)
        torch.cuda.synchronize(self.device)
        self.graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(self.graph):
            self.step_batched(update=update)

    def step_graph(self):
        self.graph.replay()

    def texture(self) -> torch.Tensor:
        return self.tex0

    # ---- checkpoint / resume -------------------------------------------------
    def state_dict(self) -> dict:
        """The fitting state: the texture (level 0), Adam's moments and its
        step count, and the learning rate (host copies)."""
        return {"texture": self.tex0.detach().cpu().clone(), "adam_m": self.m.detach().cpu().clone(),
                "adam_v": self.v.detach().cpu().clone(), "adam_t": self.t.detach().cpu().clone(), "lr": self.lr}

    def load_state_dict(self, state: dict):
        """Restores a state_dict() in place (the tensors a captured graph
        reads keep their addresses, so the graph stays valid)."""
        if tuple(state["texture"].shape) != tuple(self.tex0.shape):
            raise ValueError(f"checkpoint texture {tuple(state['texture'].shape)} != {tuple(self.tex0.shape)}")
        self.tex0.copy_(state["texture"].to(self.device))
        self.m.copy_(state["adam_m"].to(self.device))
        self.v.copy_(state["adam_v"].to(self.device))
        self.t.copy_(state["adam_t"].to(self.device))
        self.lr = float(state["lr"])

    def save(self, path: str):
        """Writes state_dict() to `path` (torch.save of tensors and a float)."""
        torch.save(self.state_dict(), path)

    def load(self, path: str):
        """Restores a checkpoint written by save() (tensors only: loaded with
        weights_only=True)."""
        self.load_state_dict(torch.load(path, map_location="cpu", weights_only=True))
