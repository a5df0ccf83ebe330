"""culsh_amd — MI355X-native LSH (p-stable, multiprobe) kNN index/query path.

A drop-in for the cuLSH HashTables/SearchTables path: hand-written gfx950 HIP
kernels in ``lib/libculsh.so`` behind the C ABI of ``include/culsh_c.h``,
with this package as the Python mirror of the reference classes.
"""
from .lsh import (FORM_PLAIN, FORM_PRECISION, FORM_PRECISION_GROUPED,  # noqa: F401
                  Batch, HashTables, SearchTables, Server, layout_requests)

__all__ = ["HashTables", "SearchTables", "Batch", "Server", "layout_requests", "FORM_PLAIN",
           "FORM_PRECISION", "FORM_PRECISION_GROUPED"]
