#!/usr/bin/env python3
"""Writes golden vectors from the reference's own compiled sources.

oracle/_ref/libswr_ref.so is SwiftShader's src/Device/Clipper.cpp and
src/Pipeline/Constants.cpp compiled unmodified (oracle/Makefile, `make ref`).
This script runs them on seeded inputs and stores inputs and outputs:

  clipper_ref.npz    3000 random clip-space triangles (x, y, z scaled by |w|,
                     w in [-1, 3]) with their clip flags; for each the
                     reference's visible flag and output polygon
                     (sw::Clipper::Clip, Clipper.cpp:22-299)
  constants_ref.npz  the 4x sample offsets, MSAA y-range offsets and the
                     anisotropic weight tables (Constants.cpp:87-297)

The tests compare the oracle against these files, so the pin holds where the
reference sources (and the compiled reference) are absent.  Run from the
repository root after `make -C oracle ref`.
"""
import ctypes
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "tests"))
sys.path.insert(0, ROOT)

from oracle_runner import reference_lib  # noqa: E402
from test_oracle_clipper import CLIP_FINITE, clip_flags, random_triangles, run  # noqa: E402


def main():
    ref = reference_lib()
    if ref is None:
        sys.exit("oracle/_ref/libswr_ref.so is missing: run `make -C oracle ref` first")
    tris, flags, vis, counts, polys = [], [], [], [], []
    for tri in random_triangles(1, 3000):
        f = clip_flags(tri[0]) | clip_flags(tri[1]) | clip_flags(tri[2])
        if f == CLIP_FINITE:
            continue
        r, p = run(ref.ref_clip_triangle, tri, f)
        out = np.zeros((16, 4), np.float32)
        out[:len(p)] = p
        tris.append(np.stack(tri))
        flags.append(f)
        vis.append(r)
        counts.append(len(p))
        polys.append(out)
    np.savez_compressed(os.path.join(HERE, "clipper_ref.npz"), triangles=np.array(tris, np.float32),
                        flags=np.array(flags, np.int32), visible=np.array(vis, np.int32),
                        count=np.array(counts, np.int32), polygon=np.array(polys, np.float32))
    xf = (ctypes.c_int * 4)()
    yf = (ctypes.c_int * 4)()
    ms = (ctypes.c_int * 2)()
    cw = (ctypes.c_uint16 * 17)()
    uw = (ctypes.c_float * 17)()
    us = (ctypes.c_float * 17)()
    ref.ref_constants(xf, yf, ms, cw, uw, us)
    np.savez_compressed(os.path.join(HERE, "constants_ref.npz"), xf=np.array(xf, np.int32), yf=np.array(yf, np.int32),
                        ms_offsets=np.array(ms, np.int32), c_weight=np.array(cw, np.uint16),
                        uv_weight=np.array(uw, np.float32), uv_start=np.array(us, np.float32))
    print(f"wrote {len(tris)} clipper vectors and the constant tables")


if __name__ == "__main__":
    main()
