"""dtvproj -- MI355X-native cone-beam projector pair (A, A^T).

Python host layer over the gfx950 HIP library ``libdtvproj.so``:

* ``geometry``  -- Standard / Beijing / physical circular geometries reduced
  to a canonical FP64 view table.
* ``ops``       -- tensor-level projector calls (forward A, matched adjoint
  A^T, reference backprojectors) through the C ABI.
* ``autograd``  -- differentiable matched pair (backward of A is A^T and
  vice versa).
* ``tv``, ``unroll``, ``batched``, ``autotune``, ``phantom`` -- the
  reconstruction building blocks around the projector.
"""
from . import geometry  # noqa: F401

__version__ = "0.1.0"
