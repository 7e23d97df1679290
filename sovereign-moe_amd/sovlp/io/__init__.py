"""Model files (reference ``src/sovopt/io/__init__.py:12-25``): MPS in and out,
CPLEX LP in.  ``read_model`` picks the reader by extension, looking through
compression suffixes: ``.lp`` to ``read_lp``, everything else to ``read_mps``.
"""

from .lp_format import LPFormatError, read_lp
from .mps import MPSError, read_mps, write_mps

__all__ = ["read_model", "read_mps", "write_mps", "read_lp", "MPSError", "LPFormatError"]


def read_model(path, name=None):
    """Read a model file (``.mps`` or ``.lp``, optionally ``.gz``, ``.bz2``, ``.xz``)."""
    base = str(path).lower()
    for suf in (".gz", ".bz2", ".xz", ".lzma"):
        if base.endswith(suf):
            base = base[: -len(suf)]
            break
    if base.endswith(".lp"):
        return read_lp(path, name=name)
    return read_mps(path, name=name)
