"""Host<->device helpers for the GPU tests (u32 data carried in int32 tensors)."""
import numpy as np


def to_dev(a, dev):
    import torch
    a = np.ascontiguousarray(a)
    if a.dtype == np.uint32:
        a = a.view(np.int32)
    return torch.from_numpy(a).to(dev)


def from_dev(t):
    a = t.cpu().numpy()
    return a.view(np.uint32) if a.dtype == np.int32 else a


def empty_u32(n, dev):
    import torch
    return torch.empty(max(n, 1), dtype=torch.int32, device=dev)


def filled_u32(n, dev, value=0x5A5A5A5A):
    """Output buffer pre-filled with a marker so that unwritten slots show."""
    import torch
    return torch.full((max(n, 1),), value - (1 << 32) if value >= (1 << 31) else value,
                      dtype=torch.int32, device=dev)


def banner(msg):
    print(f"[case] {msg}", flush=True)
