"""Helpers for the GPU parity tests: upload numpy arrays, download tensors."""
import numpy as np


def dev(a, device="cuda"):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).to(device)


def host(t):
    return t.detach().cpu().numpy()


def rel(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def banner(msg):
    print(f"[gpu-test] {msg}", flush=True)
