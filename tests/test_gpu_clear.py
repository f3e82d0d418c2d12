"""swr_clear on the GPU against the oracle, bit for bit: every format, 1 and
4 samples, whole-image and clipped areas (pixels outside stay untouched)."""
import numpy as np
import pytest
import torch

from oracle_runner import oracle_clear
from swiftshader_amd import abi, runtime

pytestmark = pytest.mark.gpu

CASES = [
    (abi.FORMAT_R8G8B8A8_UNORM, (4,), np.uint8, [0.25, 1.5, float("nan"), 0.5019608]),
    (abi.FORMAT_R32G32B32A32_SFLOAT, (4,), np.float32, [0.1, -2.0, 1e30, float("nan")]),
    (abi.FORMAT_D32_SFLOAT, (), np.float32, 1.25),
    (abi.FORMAT_S8_UINT, (), np.uint8, 0x1A5),
    (abi.FORMAT_B8G8R8A8_UNORM, (4,), np.uint8, [0.25, 1.5, float("nan"), 0.5019608]),
    (abi.FORMAT_D16_UNORM, (), np.uint16, 0.3337),
    (abi.FORMAT_D16_UNORM, (), np.uint16, float("nan")),
    (abi.FORMAT_R8G8B8A8_SRGB, (4,), np.uint8, [0.0021, 0.5, float("nan"), 0.25]),
    (abi.FORMAT_B8G8R8A8_SRGB, (4,), np.uint8, [0.8, -0.1, 0.04, 1.2]),
]
AREAS = [None, (3, 2, 20, 9), (-5, -4, 12, 40), (30, 0, 10, 10), (4, 4, 0, 5)]


@pytest.mark.parametrize("samples", [1, 4])
@pytest.mark.parametrize("case", range(len(CASES)))
def test_clear_matches_the_oracle(case, samples):
    fmt, tail, dtype, value = CASES[case]
    H, W = 23, 37
    rng = np.random.default_rng(case * 10 + samples)
    shape = (samples, H, W) + tail
    if dtype in (np.uint8, np.uint16):
        host = rng.integers(0, np.iinfo(dtype).max + 1, shape, dtype=dtype)
    else:
        host = rng.standard_normal(shape).astype(np.float32)
    dev = torch.from_numpy(host.copy().view(np.int16) if dtype == np.uint16 else host.copy()).to("cuda:0")
    want = host.copy()
    for area in AREAS:
        print(f"[gpu clear] format {fmt} samples {samples} area {area}", flush=True)
        runtime.clear(dev, fmt, value, area)
        oracle_clear(want, fmt, value, area)
        torch.cuda.synchronize()
        got = dev.cpu().numpy()
        assert np.array_equal(got.view(np.uint8), want.view(np.uint8)), area


@pytest.mark.parametrize("fmt", [abi.FORMAT_R8G8B8A8_SRGB, abi.FORMAT_B8G8R8A8_SRGB])
def test_srgb_resolve_matches_the_oracle(fmt):
    """sRGB attachments resolve in linear space (the reference's blit routine)."""
    from oracle_runner import oracle_resolve
    import ctypes
    rng = np.random.default_rng(fmt)
    src = rng.integers(0, 256, (4, 21, 35, 4), dtype=np.uint8)
    src[:, :5] = src[:1, :5]  # equal samples: the resolve returns them
    s = torch.from_numpy(src).to("cuda:0")
    d = torch.zeros((1, 21, 35, 4), dtype=torch.uint8, device="cuda:0")
    r = abi.ResolveDesc(35, 21, fmt, s.data_ptr(), 35 * 4, 21 * 35 * 4, 4, d.data_ptr(), 35 * 4)
    assert runtime.lib().swr_resolve(ctypes.byref(r), None) == 0
    torch.cuda.synchronize()
    want = oracle_resolve(src, fmt)
    got = d.cpu().numpy()
    assert np.array_equal(got, want)
    assert np.array_equal(got[0, :5], src[0, :5])


def test_clear_batch_matches_the_oracle():
    """swr_clear_batch: every format, 1 and 4 samples, whole and clipped areas,
    more clears than one launch takes, several areas of one target -- bit for
    bit the oracle's result of the same clears one by one."""
    rng = np.random.default_rng(77)
    targets, wants, items = [], [], []
    for k in range(60):
        fmt, tail, dtype, value = CASES[k % len(CASES)]
        samples = 4 if k % 3 == 0 else 1
        H, W = 9 + k % 7, 13 + (5 * k) % 11
        shape = (samples, H, W) + tail
        if dtype in (np.uint8, np.uint16):
            host = rng.integers(0, np.iinfo(dtype).max + 1, shape, dtype=dtype)
        else:
            host = rng.standard_normal(shape).astype(np.float32)
        dev = torch.from_numpy(host.copy().view(np.int16) if dtype == np.uint16 else host.copy()).to("cuda:0")
        want = host.copy()
        # Two areas of one target whose byte spans are disjoint: the top rows
        # and the bottom rows (clipped on the right).
        areas = [None] if k % 4 else [(0, 0, W, H // 2), (1, H // 2, W, H)]
        for area in areas:
            items.append((dev, fmt, value, area))
            oracle_clear(want, fmt, value, area)
        targets.append(dev)
        wants.append(want)
    print(f"[gpu clear] batch of {len(items)} clears", flush=True)
    runtime.clear_batch(items)
    torch.cuda.synchronize()
    for dev, want in zip(targets, wants):
        got = dev.cpu().numpy()
        assert np.array_equal(got.view(np.uint8), want.view(np.uint8))


def test_clear_batch_refuses_overlapping_targets():
    """Clears of one batch run concurrently: areas whose byte spans overlap
    (here two column ranges of the same rows) are refused before anything is
    launched, and the targets stay untouched."""
    dev = torch.zeros((1, 8, 8, 4), dtype=torch.uint8, device="cuda:0")
    items = [(dev, abi.FORMAT_R8G8B8A8_UNORM, [1, 0, 0, 1], (0, 0, 5, 8)),
             (dev, abi.FORMAT_R8G8B8A8_UNORM, [0, 1, 0, 1], (4, 0, 4, 8))]
    with pytest.raises(runtime.SwrError, match="overlap"):
        runtime.clear_batch(items)
    torch.cuda.synchronize()
    assert int(dev.sum().item()) == 0
