"""Helpers shared by the GPU parity tests."""
import sys

import numpy as np
import torch

from golikov_2d_amd import api
from oracle import oracle as O


def say(msg):
    print(f"[gpu-case] {msg}", flush=True)
    sys.stdout.flush()


def dev():
    return torch.device("cuda:0")


def to_dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).to(dev())


def to_host(t):
    torch.cuda.synchronize()
    return t.cpu().numpy()


def rel_l1(a, b):
    """per-field relative L1: sum|a-b| / sum|b| over the leading axis"""
    a = a.reshape(a.shape[0], -1)
    b = b.reshape(b.shape[0], -1)
    return np.abs(a - b).sum(axis=1) / np.maximum(np.abs(b).sum(axis=1), 1e-300)


def grids(nx, ny, bc, gamma=1.4, **kw):
    """(product grid, oracle grid) with identical fields"""
    return api.make_grid(nx, ny, gamma=gamma, bc=bc, **kw), O.make_grid(nx, ny, gamma=gamma, bc=bc, **kw)
