"""CPU oracle for the cone-beam projector pair -- TEST INFRASTRUCTURE ONLY.

This package is a CPU restatement (PyTorch FP64 by default) of the reference
sampling schemes of ILoveU3D/DTV-net ``model/ConeBeamLayers``.  It exists to
check the HIP path and to serve as the timed CPU baseline; nothing in the
product imports it.  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may use it, and only as the checker.

How each function is pinned (none of it is pinned by reference golden
vectors: the reference ships no tests, fixtures or data):

* ``cone.std_forward``/``bj_forward``/``circ_forward`` -- against the literal
  scalar loops in ``brute.py`` (tiny inputs), the closed-form perspective
  position of a point object, chord lengths of a constant volume, exactness
  of trilinear sampling on affine fields, rotation equivariance.
* ``cone.*_adjoint`` (matched transpose) -- against torch autograd VJP of the
  FP64 forward and the dense brute-force matrix transpose.
* ``cone.std_bp_ref``/``bj_bp_ref`` (reference backprojectors) -- against the
  scalar loops in ``brute.py``, the closed-form constant-sinogram response
  and the line-intersection closed form.
* ``tv.*`` -- against the dense sparse-matrix definition of the reference's
  difference operators on tiny volumes.

Parity to the reference CUDA binaries themselves is unpinned: they need an
NVIDIA GPU and texture hardware (8-bit fixed-point filter weights).
"""
