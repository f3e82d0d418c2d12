"""Address modes the library refuses, through the C ABI (the mirrored modes
themselves are parity cases beside the other scenes).

CLAMP_TO_BORDER (VkSamplerAddressMode 3) is not built; the backward pass
differentiates REPEAT and CLAMP_TO_EDGE only.
"""
import numpy as np
import pytest

from swiftshader_amd import abi
from swiftshader_amd.scenes import DrawData, TextureData

pytestmark = pytest.mark.gpu


def _quad(address_u):
    tex = np.linspace(0.0, 1.0, 8 * 8 * 4, dtype=np.float32).reshape(8, 8, 4)
    texture = TextureData(format=abi.FORMAT_R32G32B32A32_SFLOAT, levels=[tex], address_u=address_u,
                          address_v=abi.ADDRESS_CLAMP_TO_EDGE)
    xy = np.array([[-1, -1], [1, -1], [1, 1], [-1, -1], [1, 1], [-1, 1]], np.float32)
    pos = np.concatenate([xy, np.full((6, 1), 0.5, np.float32), np.ones((6, 1), np.float32)], axis=1)
    return DrawData(width=16, height=16, positions=pos, colors=np.ones((6, 4), np.float32), texcoords=xy * 1.5,
                    shader=abi.SHADER_SAMPLE, texture=texture, color_format=abi.FORMAT_R32G32B32A32_SFLOAT)


def test_clamp_to_border_is_refused():
    import torch
    from swiftshader_amd import runtime
    print("[gpu address] CLAMP_TO_BORDER refusal", flush=True)
    dd = runtime.prepare(_quad(3), "cuda:0")
    with pytest.raises(runtime.SwrError):
        runtime.draw(dd)
    torch.cuda.synchronize()


def test_mirrored_backward_is_refused():
    import torch
    from swiftshader_amd import diff, runtime
    print("[gpu address] mirrored backward refusal", flush=True)
    d = _quad(abi.ADDRESS_MIRRORED_REPEAT)
    view = diff.View(d, "cuda:0")
    pos = torch.from_numpy(d.positions).cuda()
    col = torch.from_numpy(d.colors).cuda()
    uv = torch.from_numpy(d.texcoords).cuda()
    levels = [torch.from_numpy(np.ascontiguousarray(lv)).cuda() for lv in d.texture.levels]
    color = diff.draw(view, pos, col, uv, levels)
    torch.cuda.synchronize()
    assert float(color.sum()) > 0
    with pytest.raises(runtime.SwrError):
        diff.backward_raw(view, pos, col, uv, levels, view.ids, torch.ones_like(color))
    torch.cuda.synchronize()
