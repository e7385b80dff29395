"""Drop-in for ``data/simulateLoader.py``: datasets that read raw float32
volumes and project them with the Beijing geometry in ``__getitem__``
(simulateLoader.py:7-83), on the gfx950 projector.

Each item is ``(input, proj, img)`` with the batch dimension squeezed and
every tensor on the CPU, as the reference returns them; ``proj`` is
``ForwardProjection.apply`` of the volume (sampleInterval included, planes
interleaved in the reference's public layout).
"""
import os

import torch
from torch.utils.data import Dataset

from dtvproj.rawio import read_raw
from model.ConeBeamLayers.Beijing.BeijingGeometry import ForwardProjection


def _list(dir):
    return [os.path.join(dir, item) for item in os.listdir(dir)]


def _project(img, device):
    """(input placeholder, proj, img) for one volume [1, Z, Y, X]."""
    x = torch.unsqueeze(torch.from_numpy(img), 0).to(device)
    proj = ForwardProjection.apply(x)
    return x, proj


class Stimulated256(Dataset):
    """Label volume -> (zeros, A(label), label) (simulateLoader.py:7-27)."""

    def __init__(self, dir, device, scale=(1, 64, 256, 256)):
        self.dir = _list(dir)
        self.scale = scale
        self.device = device

    def __getitem__(self, index):
        img = read_raw(self.dir[index], self.scale)
        x, proj = _project(img, self.device)
        inp = torch.zeros_like(x)
        return torch.squeeze(inp, 0).cpu(), torch.squeeze(proj, 0).cpu(), torch.squeeze(x, 0).cpu()

    def __len__(self):
        return len(self.dir)


class Stimulated256Input(Dataset):
    """Label and a same-named initial volume under ``inputDir`` ->
    (input, A(label), label) (simulateLoader.py:29-55)."""

    def __init__(self, dir, inputDir, device, scale=(1, 64, 256, 256)):
        self.dir = []
        self.input = []
        for item in os.listdir(dir):
            self.dir.append(os.path.join(dir, item))
            self.input.append(os.path.join(inputDir, item))
        self.scale = scale
        self.device = device

    def __getitem__(self, index):
        img = read_raw(self.dir[index], self.scale)
        x, proj = _project(img, self.device)
        inp = torch.unsqueeze(torch.from_numpy(read_raw(self.input[index], self.scale)), 0)
        return torch.squeeze(inp, 0).cpu(), torch.squeeze(proj, 0).cpu(), torch.squeeze(x, 0).cpu()

    def __len__(self):
        return len(self.dir)


class Stimulated256InputShepp(Dataset):
    """As :class:`Stimulated256Input` with both volumes scaled by 1000 and the
    projection run on the current device (simulateLoader.py:57-83)."""

    def __init__(self, dir, inputDir, scale=(1, 64, 256, 256)):
        self.dir = []
        self.input = []
        for item in os.listdir(dir):
            self.dir.append(os.path.join(dir, item))
            self.input.append(os.path.join(inputDir, item))
        self.scale = scale

    def __getitem__(self, index):
        device = torch.device("cuda", torch.cuda.current_device())
        img = read_raw(self.dir[index], self.scale, scale=1000.0)
        x, proj = _project(img, device)
        inp = torch.from_numpy(read_raw(self.input[index], self.scale, scale=1000.0))
        return inp, torch.squeeze(proj, 0).cpu(), torch.squeeze(x, 0).cpu()

    def __len__(self):
        return len(self.dir)
