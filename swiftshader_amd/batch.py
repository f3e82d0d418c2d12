"""Batch plans: many draws prepared once, enqueued many times.

A thin owner of a ``swr_batch_plan`` (include/swr.h): create it from draw
descriptors, run it (no allocation, no host wait: capturable into a HIP
graph), autotune it, attach gradient targets and enqueue the batched
backward.  The plan reads the memory the descriptors name at run time.
"""
import ctypes
from typing import Optional, Sequence

from . import abi, runtime


class BatchPlan:
    def __init__(self, descs: Sequence[abi.DrawDesc], headroom_percent: int = 10, stream=None):
        self._descs = (abi.DrawDesc * len(descs))(*descs)
        self._grads = None
        self.count = len(descs)
        self._plan = ctypes.c_void_p()
        runtime.check(runtime.lib().swr_batch_plan_create(self._descs, self.count, headroom_percent,
                                                          ctypes.byref(self._plan), runtime._stream_handle(stream)),
                      "swr_batch_plan_create")

    def run(self, stream=None):
        runtime.check(runtime.lib().swr_batch_plan_run(self._plan, runtime._stream_handle(stream)),
                      "swr_batch_plan_run")

    def autotune(self, stream=None) -> abi.Tuning:
        best = abi.Tuning()
        runtime.check(runtime.lib().swr_batch_plan_autotune(self._plan, ctypes.byref(best),
                                                            runtime._stream_handle(stream)), "swr_batch_plan_autotune")
        return best

    def pairs_overflow(self) -> int:
        need = ctypes.c_uint32(0)
        runtime.check(runtime.lib().swr_batch_plan_check(self._plan, ctypes.byref(need)), "swr_batch_plan_check")
        return int(need.value)

    def set_grads(self, grads: Sequence[abi.GradDesc], stream=None):
        self._grads = (abi.GradDesc * len(grads))(*grads)
        runtime.check(runtime.lib().swr_batch_plan_set_grads(self._plan, self._grads, len(grads),
                                                             runtime._stream_handle(stream)), "swr_batch_plan_set_grads")

    def backward(self, stream=None):
        runtime.check(runtime.lib().swr_batch_plan_backward(self._plan, runtime._stream_handle(stream)),
                      "swr_batch_plan_backward")

    def close(self):
        if self._plan:
            runtime.lib().swr_batch_plan_destroy(self._plan)
            self._plan = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def grad_desc(d_color, d_positions=None, d_colors=None, d_texcoords=None, d_texture: Optional[Sequence] = None,
              sample_state=None) -> abi.GradDesc:
    """A swr_grad_desc over torch tensors: d_color (S, H, W, 4) float32, rows
    and pixels contiguous (the sample planes may lie apart)."""
    S, H, W, _ = d_color.shape
    if d_color.stride()[1:] != (W * 4, 4, 1):
        raise ValueError("d_color rows must be contiguous (S, H, W, 4) planes")
    g = abi.GradDesc()
    g.d_color = d_color.data_ptr()
    g.d_color_row_pitch_bytes = W * 16
    g.d_color_sample_pitch_bytes = d_color.stride(0) * 4 if S > 1 else H * W * 16
    g.d_positions = d_positions.data_ptr() if d_positions is not None else None
    g.d_colors = d_colors.data_ptr() if d_colors is not None else None
    g.d_texcoords = d_texcoords.data_ptr() if d_texcoords is not None else None
    for i, t in enumerate(d_texture or []):
        g.d_texture[i] = t.data_ptr()
    if sample_state is not None:
        g.sample_state = sample_state.data_ptr()
        g.sample_state_row_pitch_bytes = W * 16
        g.sample_state_sample_pitch_bytes = H * W * 16
    return g
