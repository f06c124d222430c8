"""MI355X-native differentiable mesh rasterizer (host package).

The compute path is libgvv_amd.so (hand-written HIP kernels for gfx950 behind a C ABI);
this package owns handles, tensors and the autograd wiring.  ``CudaRenderer`` (next to this
package) is the drop-in for the reference's Python API.
"""
from ._lib import GvvError, load as load_library
from .renderer import (ALBEDO_MODES, SHADING_MODES, Rasterizer, RenderFunction, cuda_renderer_gpu,
                       cuda_renderer_grad_gpu, get_rasterizer)

__all__ = ["GvvError", "load_library", "ALBEDO_MODES", "SHADING_MODES", "Rasterizer", "RenderFunction",
           "cuda_renderer_gpu", "cuda_renderer_grad_gpu", "get_rasterizer"]
