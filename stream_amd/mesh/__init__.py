"""Mesh bindings: the reference's ``stream.mesh`` (python/mesh/__init__.py:20-28)."""
from . import ashape2D
from .ashape2D import AlphaShape2D

__all__ = ["AlphaShape2D", "ashape2D"]
