"""Numerics bindings: the reference's ``stream.numerics`` (python/numerics/__init__.py:20-30)."""
