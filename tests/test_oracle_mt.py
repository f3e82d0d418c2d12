"""The multithreaded oracle (swo_draw_mt) renders the same bytes as swo_draw.

swo_draw_mt splits row pairs over host threads as the reference splits them
over its clusters (QuadRasterizer.cpp:46-49), with primitives set up in
batches by all threads; every pixel is still visited by one thread in
primitive order, so colour, depth, stencil, the primitive_id record and the
occlusion count must not change with the thread count.  Each parity case
runs at 1, 3 and 8 threads (3 leaves row pairs unevenly split; 8 exceeds
the row pairs of the small targets' tiles).
"""
import numpy as np
import pytest

from cases import parity_cases
from oracle_runner import oracle_render, oracle_render_ids
from swiftshader_amd import abi

CASES = {name: d for name, d in parity_cases()}
NAMES = ["textured_small", "clip_heavy", "stencil_msaa4", "blend_variety", "scissor_cull_viewport",
         "lines_clipped_last_msaa4_stencil", "points_sizes_msaa4_clipped_stencil", "d16_msaa4_bias_stencil",
         "rgba8_msaa4_aniso", "depth_bounds_gouraud"]


def _bytes(a):
    return None if a is None else a.view(np.uint8)


@pytest.mark.parametrize("name", NAMES)
def test_mt_equals_single_thread(name):
    d = CASES[name]
    ref = oracle_render(d, threads=1)
    for n in (3, 8):
        got = oracle_render(d, threads=n)
        for a, b in zip(ref[:3], got[:3]):
            assert (a is None) == (b is None)
            if a is not None:
                assert np.array_equal(_bytes(a), _bytes(b)), (name, n)
        assert ref[3] == got[3], (name, n)


def test_mt_primitive_ids():
    d = CASES["float_msaa4_aniso"]
    assert d.color_format == abi.FORMAT_R32G32B32A32_SFLOAT
    ref = oracle_render_ids(d, threads=1)
    got = oracle_render_ids(d, threads=5)
    for a, b in zip(ref[:3], got[:3]):
        assert np.array_equal(_bytes(a), _bytes(b))
    assert ref[3] == got[3]
