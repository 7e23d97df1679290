"""Model files: MPS in and out, CPLEX LP in.  ``read_model`` is the entry the
reference offers at ``src/sovopt/io/__init__.py:12-25``: the reader follows the
format extension, found through the compression table of ``io/mps.py``.
"""

from .lp_format import LPFormatError, read_lp
from .mps import MPSError, _format_extension, read_mps, write_mps

__all__ = ["read_model", "read_mps", "write_mps", "read_lp", "MPSError", "LPFormatError"]

_READERS = {".lp": read_lp}


def read_model(path, name=None):
    """Read a model file: CPLEX LP for ``.lp``, MPS otherwise; either may be
    compressed (``.gz``, ``.bz2``, ``.xz``, ``.lzma``)."""
    return _READERS.get(_format_extension(path), read_mps)(path, name=name)
