"""Device-memory helpers for the GPU tests (torch allocates, libmpkx computes)."""
import sys

import numpy as np


def torch_cuda():
    import torch
    if not torch.cuda.is_available():
        raise RuntimeError("GPU test run without a GPU")
    return torch


def to_dev(a):
    torch = torch_cuda()
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda")


def zeros_dev(n):
    torch = torch_cuda()
    return torch.zeros(n, dtype=torch.float64, device="cuda")


def to_host(t):
    torch = torch_cuda()
    torch.cuda.synchronize()
    return t.cpu().numpy()


def say(*args):
    print(*args, flush=True)
    sys.stdout.flush()
