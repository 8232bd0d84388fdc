"""Helpers shared by the GPU parity tests (host <-> device copies, errors)."""
import numpy as np
import torch


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda")


def host(t):
    torch.cuda.synchronize()
    return t.detach().cpu().numpy()


def rel_l2(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    den = np.linalg.norm(b.ravel())
    num = np.linalg.norm((a - b).ravel())
    return num / den if den > 0 else num


def rel_l2_vec(A, B):
    num = sum(np.linalg.norm((np.asarray(a) - np.asarray(b)).ravel()) ** 2 for a, b in zip(A, B))
    den = sum(np.linalg.norm(np.asarray(b).ravel()) ** 2 for b in B)
    return np.sqrt(num / den) if den > 0 else np.sqrt(num)


def log(msg):
    print(msg, flush=True)
