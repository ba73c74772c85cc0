"""spiral_pir — MI355X (gfx950) Spiral PIR server answer path.

Host-side Python face of libspiral_hip.so: the reference's operator surface
(``ops``), the resident database and per-query answer pipeline (``server``),
the client (``client``) and fixture helpers (``host``).  Device work always
runs in the HIP library; nothing here computes on the CPU except the client,
which is host code by nature.
"""
from . import params
from ._lib import SpiralError, version
from .params import Shape

__all__ = ["params", "Shape", "SpiralError", "version"]
