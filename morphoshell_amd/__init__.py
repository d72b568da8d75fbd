"""MI355X-native per-step force-and-dynamics path of MorphoShell.

Layers: ``meshing`` (synthetic meshes and programmed metrics), ``native``
(ctypes binding of the C-ABI library built from ``csrc/``), ``shell`` (host
mirror of the reference's per-step calls), ``build`` (hipcc recipe).
"""
from .meshing import Mesh, disc_mesh, square_mesh  # noqa: F401

__version__ = "0.1.0"
