"""Branch-and-bound support on the GPU: batched node relaxation (``bnr``)."""
