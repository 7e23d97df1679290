"""sovlp -- MI355X-native PDLP engine, a drop-in for SOVOPT's ``--method pdlp --device gpu``.

Python host code over ``libsovlp.so`` (hand-written gfx950 HIP kernels behind a C ABI,
``include/sovlp.h``), with PyTorch-ROCm tensors as device buffers.
"""

__version__ = "0.1.0"
