"""Mesh bindings: the reference's ``stream.mesh`` (python/mesh/__init__.py:20-28)."""
from . import ashape2D, ashape3D, mesh
from .ashape2D import AlphaShape2D
from .ashape3D import AlphaShape3D
from .mesh import Mesh2D

__all__ = ["AlphaShape2D", "AlphaShape3D", "Mesh2D", "ashape2D", "ashape3D", "mesh"]
