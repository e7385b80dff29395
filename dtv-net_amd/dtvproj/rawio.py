"""Raw float32 volumes and sinograms, the reference's on-disk format.

The reference stores every array as headerless little-endian float32 in C
order and reads it with ``np.fromfile(path, "float32")`` followed by a
reshape (data/simulateLoader.py:16-17, data/realLoader.py:26-34,
iterations/iteration.py:22-23): volumes ``[1, Z, Y, X]`` (``[1, 64, 256,
256]``), Beijing sinograms ``[1, angles*planes, V, U]`` (``[1, 3024, 128,
80]``).  These helpers add the size check the reference leaves to
``reshape`` and an optional memory map for large files.
"""
from __future__ import annotations

import os

import numpy as np

VOLUME_SCALE = (1, 64, 256, 256)
SINO_SCALE = (1, 144 * 21, 128, 80)


def read_raw(path: str, shape, scale: float = 1.0, mmap: bool = False) -> np.ndarray:
    """float32 array of ``shape`` from a headerless raw file, times ``scale``.

    Raises ValueError naming the file when its size does not match."""
    shape = tuple(int(s) for s in shape)
    n = int(np.prod(shape))
    size = os.path.getsize(path)
    if size != 4 * n:
        raise ValueError(f"{path}: {size} bytes, expected {4 * n} for float32 {list(shape)}")
    if mmap:
        a = np.memmap(path, dtype="<f4", mode="r", shape=shape)
        return np.asarray(a, dtype=np.float32) * np.float32(scale) if scale != 1.0 else a
    a = np.fromfile(path, dtype="<f4").reshape(shape)
    return a * np.float32(scale) if scale != 1.0 else a


def write_raw(path: str, array) -> None:
    """Write ``array`` as headerless little-endian float32 (C order)."""
    try:
        import torch
        if isinstance(array, torch.Tensor):
            array = array.detach().cpu().numpy()
    except ImportError:
        pass
    np.ascontiguousarray(array, dtype="<f4").tofile(path)
