"""Host <-> device helpers for the GPU tests (torch provides device memory)."""
import numpy as np


def to_device(a):
    import torch
    a = np.ascontiguousarray(a)
    if a.dtype == np.uint16:
        t = torch.from_numpy(a.view(np.int16))
    else:
        t = torch.from_numpy(a)
    return t.cuda()


def empty_like_np(a):
    import torch
    dt = {np.dtype(np.uint8): torch.uint8, np.dtype(np.uint16): torch.int16,
          np.dtype(np.float32): torch.float32}[np.dtype(a.dtype)]
    return torch.empty(a.shape, dtype=dt, device="cuda")


def to_host(t, dtype):
    a = t.cpu().numpy()
    if np.dtype(dtype) == np.uint16:
        return a.view(np.uint16)
    return a.astype(dtype, copy=False)


def compare(gpu, ref, what=""):
    """Return (max_abs_diff, fraction of differing samples)."""
    g = gpu.astype(np.float64)
    r = ref.astype(np.float64)
    d = np.abs(g - r)
    return float(d.max()), float(np.count_nonzero(d) / d.size)


def assert_parity(gpu, ref, floating=False, what=""):
    """Integer planes: within 1 LSB on at most 0.1% of samples.
    Float planes: within 1e-5 relative (floor 1e-5 absolute)."""
    if floating:
        g = gpu.astype(np.float64)
        r = ref.astype(np.float64)
        err = np.abs(g - r) / np.maximum(np.abs(r), 1.0)
        assert float(err.max()) <= 1e-5, f"{what}: max rel err {err.max():.3g}"
    else:
        mx, frac = compare(gpu, ref)
        assert mx <= 1 and frac <= 1e-3, f"{what}: max diff {mx}, differing fraction {frac:.3g}"
