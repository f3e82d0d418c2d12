"""Line and point states the library refuses, through the C ABI (the line
and point scenes themselves are parity cases beside the triangles').

Depth bias is applied to triangles only by the reference's setup, while its
pixel routine would still add the field (lines and points); a Bresenham line turns
multisampling off in the reference, a mode not built here; widths outside
(0, 64] are invalid; the backward pass differentiates triangles only.
diff.View carries no point sizes, so its point draws use side 1.
"""
import numpy as np
import pytest

from swiftshader_amd import abi
from swiftshader_amd.scenes import DrawData

pytestmark = pytest.mark.gpu


def _lines(**kw):
    pos = np.array([[-0.8, -0.6, 0.5, 1.0], [0.7, 0.4, 0.5, 1.0], [0.1, -0.9, 0.3, 1.0], [-0.2, 0.8, 0.6, 1.0]],
                   np.float32)
    args = dict(width=32, height=24, positions=pos, colors=np.ones((4, 4), np.float32),
                topology=abi.TOPOLOGY_LINE_LIST, color_format=abi.FORMAT_R32G32B32A32_SFLOAT)
    args.update(kw)
    return DrawData(**args)


@pytest.mark.parametrize("kw", [dict(depth_bias=(2.0, 0.0, 0.0)), dict(depth_bias=(0.0, 0.0, 1.0)),
                                dict(line_rasterization_mode=abi.LINE_RASTERIZATION_BRESENHAM, samples=4),
                                dict(line_width=-1.0), dict(line_width=65.0), dict(line_rasterization_mode=4)],
                         ids=["bias_constant", "bias_slope", "bresenham_msaa", "width_negative", "width_large",
                              "mode_unknown"])
def test_refused_line_states(kw):
    import torch
    from swiftshader_amd import runtime
    print(f"[gpu lines] refusal {kw}", flush=True)
    dd = runtime.prepare(_lines(**kw), "cuda:0")
    with pytest.raises(runtime.SwrError):
        runtime.draw(dd)
    torch.cuda.synchronize()


def _points(**kw):
    pos = np.array([[-0.5, -0.5, 0.5, 1.0], [0.4, 0.3, 0.5, 1.0]], np.float32)
    args = dict(width=32, height=24, positions=pos, colors=np.ones((2, 4), np.float32),
                point_sizes=np.array([3.0, 2.0], np.float32), topology=abi.TOPOLOGY_POINT_LIST,
                color_format=abi.FORMAT_R32G32B32A32_SFLOAT)
    args.update(kw)
    return DrawData(**args)


def test_refused_point_states():
    import torch
    from swiftshader_amd import diff, runtime
    dd = runtime.prepare(_points(depth_bias=(1.0, 0.0, 0.0)), "cuda:0")
    with pytest.raises(runtime.SwrError):
        runtime.draw(dd)
    d = _points()
    view = diff.View(d, "cuda:0")
    pos = torch.from_numpy(d.positions).cuda()
    col = torch.from_numpy(d.colors).cuda()
    color = diff.draw(view, pos, col)
    torch.cuda.synchronize()
    assert float(color.sum()) > 0
    with pytest.raises(runtime.SwrError):
        diff.backward_raw(view, pos, col, None, [], view.ids, torch.ones_like(color))
    torch.cuda.synchronize()


def test_accepted_line_states():
    """Depth bias enabled with both factors zero adds nothing and is accepted."""
    import torch
    from swiftshader_amd import runtime
    dd = runtime.prepare(_lines(depth_bias=(0.0, 0.5, 0.0)), "cuda:0")
    runtime.draw(dd)
    torch.cuda.synchronize()
    assert int(dd.targets.occlusion.cpu().item()) > 0


def test_backward_of_lines_is_refused():
    import torch
    from swiftshader_amd import diff, runtime
    d = _lines()
    view = diff.View(d, "cuda:0")
    pos = torch.from_numpy(d.positions).cuda()
    col = torch.from_numpy(d.colors).cuda()
    color = diff.draw(view, pos, col)
    torch.cuda.synchronize()
    assert float(color.sum()) > 0
    with pytest.raises(runtime.SwrError):
        diff.backward_raw(view, pos, col, None, [], view.ids, torch.ones_like(color))
    torch.cuda.synchronize()
