"""Model files (reference ``src/sovopt/io/__init__.py:12-25``): MPS in and out.

``read_model`` picks the reader by extension, looking through compression
suffixes.  The CPLEX LP format (``.lp``) is not provided: it raises
``NotImplementedError`` rather than being misread as MPS.
"""

from .mps import MPSError, read_mps, write_mps

__all__ = ["read_model", "read_mps", "write_mps", "MPSError"]


def read_model(path, name=None):
    """Read a model file; ``.mps`` (optionally ``.gz``, ``.bz2``, ``.xz``) only."""
    base = str(path).lower()
    for suf in (".gz", ".bz2", ".xz", ".lzma"):
        if base.endswith(suf):
            base = base[: -len(suf)]
            break
    if base.endswith(".lp"):
        raise NotImplementedError("the CPLEX LP format is not provided; convert the model to MPS")
    return read_mps(path, name=name)
