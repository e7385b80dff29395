"""Drop-in for ``data/realLoader.py``: measured scans stored as raw float32
label / input volumes and sinograms plus a per-scan ``.mat`` holding the
Beijing ``projection_matrix`` (realLoader.py:9-49).

For each item the volumes and sinogram are scaled by ``amp`` = 6e4, the
volumes clipped at zero, and the scan's projection vectors installed as the
Beijing module's ``parameters`` (the reference's way of switching geometry
per scan).  Items are ``(input, proj, img)`` on the CPU.
"""
import os

import numpy as np
import torch
from torch.utils.data import Dataset

from dtvproj.rawio import read_raw
import model.ConeBeamLayers.Beijing.BeijingGeometry as Geometry


def load_projection_matrix(path):
    """float32 [views, 12] ``projection_matrix`` from a MATLAB file."""
    import scipy.io as sco
    m = np.array(sco.loadmat(path)["projection_matrix"]).astype(np.float32)
    if m.ndim != 2 or m.shape[1] != 12:
        raise ValueError(f"{path}: projection_matrix has shape {m.shape}, expected [views, 12]")
    return torch.from_numpy(m).contiguous()


class Real256Input(Dataset):
    def __init__(self, dirs, device, scale=(1, 64, 256, 256), pScale=(1, 144 * 21, 128, 80)):
        self.dir = []
        for dir in dirs:
            for item in os.listdir(os.path.join(dir, "label")):
                self.dir.append({
                    "label": os.path.join(dir, "label", item),
                    "input": os.path.join(dir, "input", item),
                    "matlab": os.path.join(dir, "matlab", item.replace(".raw", ".mat")),
                    "sino": os.path.join(dir, "sino", item),
                })
        self.scale = scale
        self.pScale = pScale
        self.device = device
        self.amp = 6e4

    def __getitem__(self, index):
        check = self.dir[index]
        img = torch.relu(torch.from_numpy(read_raw(check["label"], self.scale, self.amp)))
        proj = torch.from_numpy(read_raw(check["sino"], self.pScale, self.amp))
        inp = torch.relu(torch.from_numpy(read_raw(check["input"], self.scale, self.amp)))
        Geometry.parameters = load_projection_matrix(check["matlab"])
        return inp, proj, img

    def __len__(self):
        return len(self.dir)
