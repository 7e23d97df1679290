"""Infinite-bound sentinel shared by every layer.

Mirrors ``src/sovopt/core/tolerances.py:16`` of the reference: any bound at or
beyond ``1e30`` in magnitude is treated as infinite, which is the MPS
convention. The device kernels test finiteness with the same threshold, so a
model crosses the host/device boundary without re-encoding its bounds.
"""

INF = 1.0e30

__all__ = ["INF"]
