"""MI355X-native path tracer behind the OptiXRender host interface.

The compute path is the HIP library ``liboxr.so`` (include/oxr.h); this
package holds its ctypes binding, the Python mirror of the reference's
Renderer/Scene interface and the synthetic scene generators.
"""
from . import abi
from .renderer import Context, Renderer, SceneData

__all__ = ["abi", "Context", "Renderer", "SceneData"]
