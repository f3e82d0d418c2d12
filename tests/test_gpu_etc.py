"""swr_decode_etc on the GPU against the oracle, bit for bit: every input
type (EAC R11 / RG11 signed and unsigned, ETC2 RGB8, RGB8 with punch-through
alpha, RGBA8) over random blocks, which reach the individual, differential,
T, H and planar modes and every modifier table; ragged and large images,
destination bytes outside the image untouched; the refused cases."""
import ctypes

import numpy as np
import pytest
import torch

from oracle_runner import ETC_TYPES, etc_block_bytes, etc_out_bpp, oracle_decode_etc
from swiftshader_amd import abi, runtime

pytestmark = pytest.mark.gpu

SHAPES = [(4, 4), (13, 7), (1, 1), (30, 5), (257, 130), (1024, 768)]


def _gpu_decode(src, w, h, bpp, t, pitch, fill):
    s = torch.from_numpy(src).to("cuda:0")
    d = torch.full((pitch * h,), fill, dtype=torch.uint8, device="cuda:0")
    desc = abi.ETCDesc(s.data_ptr(), d.data_ptr(), w, h, pitch, bpp, t)
    rc = runtime.lib().swr_decode_etc(ctypes.byref(desc), None)
    torch.cuda.synchronize()
    return rc, d.cpu().numpy()


@pytest.mark.parametrize("name", list(ETC_TYPES))
def test_decode_matches_the_oracle(name):
    t = ETC_TYPES[name]
    rng = np.random.default_rng(300 + t)
    bpp = etc_out_bpp(t)
    for w, h in SHAPES:
        print(f"[gpu etc] {name} {w}x{h}", flush=True)
        nb = ((w + 3) // 4) * ((h + 3) // 4)
        src = rng.integers(0, 256, nb * etc_block_bytes(t), dtype=np.uint8)
        pitch = w * bpp + 12
        want = np.full(pitch * h, 0xA5, np.uint8)
        assert oracle_decode_etc(src, want.reshape(h, pitch), w, h, pitch, bpp, t) == 1
        rc, got = _gpu_decode(src, w, h, bpp, t, pitch, 0xA5)
        assert rc == 0 and np.array_equal(got, want), (name, w, h)


def test_refused_cases():
    src = np.zeros(64, np.uint8)
    rc, _ = _gpu_decode(src, 4, 4, 4, 7, 16, 0)  # no input type 7
    assert rc == abi.ERROR_INVALID_ARGUMENT
    rc, _ = _gpu_decode(src, 4, 4, 2, abi.ETC_RGB, 16, 0)  # RGB8 writes 4-byte pixels
    assert rc == abi.ERROR_INVALID_ARGUMENT
    rc, _ = _gpu_decode(src, 4, 4, 4, abi.ETC_R_UNSIGNED, 16, 0)  # R11 writes 2-byte pixels
    assert rc == abi.ERROR_INVALID_ARGUMENT
    rc, _ = _gpu_decode(src, 4, 4, 4, abi.ETC_RGBA, 18, 0)  # pitch not a multiple of 4
    assert rc == abi.ERROR_INVALID_ARGUMENT
