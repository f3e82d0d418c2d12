"""MI355X-native fixed-function triangle draw path (SwiftShader's sw::Renderer::draw).

Importing the package loads torch first, so that libswr.so binds to the
same HIP runtime instance torch uses.  The C ABI lives in include/swr.h;
``swiftshader_amd.runtime`` is the Python host side over it
(``from swiftshader_amd import runtime``).
"""
import torch  # noqa: F401  (must precede loading libswr.so)

from . import abi, scenes  # noqa: F401

__all__ = ["abi", "scenes"]
