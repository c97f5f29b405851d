"""Seeded synthetic clips of the benchmark configurations (numpy, host).

Each generator is deterministic per (seed, frame, plane)."""
import numpy as np


def gray8_textured(width=640, height=360, frames=24, seed=1):
    """Smooth 2D gradient + random-orientation sinusoid texture (amplitude 40)
    + Gaussian noise sigma=5, clipped to [0,255] and rounded."""
    rng = np.random.default_rng(seed)
    yy, xx = np.mgrid[0:height, 0:width].astype(np.float64)
    out = []
    for n in range(frames):
        r = np.random.default_rng((seed, n))
        gx, gy, base = r.uniform(-0.1, 0.1), r.uniform(-0.1, 0.1), r.uniform(80, 170)
        img = base + gx * (xx - width / 2) + gy * (yy - height / 2)
        theta = r.uniform(0, np.pi)
        period = r.uniform(6, 40)
        phase = r.uniform(0, 2 * np.pi)
        img += 40 * np.sin(2 * np.pi * (xx * np.cos(theta) + yy * np.sin(theta)) / period + phase)
        img += r.normal(0, 5, size=img.shape)
        out.append(np.clip(np.rint(img), 0, 255).astype(np.uint8))
    del rng
    return out


def yuv420_sinusoids(width=1920, height=1080, frames=8, bits=16, seed=2, noise_frac=0.02, first_frame=0):
    """Per plane: 8 seeded random-orientation sinusoids (periods 4-64 px)
    drifting 0-2 px/frame + Gaussian noise at noise_frac of full scale.
    Returns frames as [Y, U, V] lists of uint8/uint16 arrays."""
    peak = (1 << bits) - 1
    dt = np.uint8 if bits == 8 else np.uint16
    dims = [(height, width), (height // 2, width // 2), (height // 2, width // 2)]
    params = []
    for p in range(3):
        r = np.random.default_rng((seed, p))
        params.append(dict(theta=r.uniform(0, np.pi, 8), period=r.uniform(4, 64, 8), phase=r.uniform(0, 2 * np.pi, 8),
                           amp=r.uniform(0.2, 1.0, 8), drift=r.uniform(0, 2, 8)))
    grids = [np.mgrid[0:h, 0:w].astype(np.float32) for h, w in dims]
    clip = []
    for n in range(first_frame, first_frame + frames):
        planes = []
        for p in range(3):
            yy, xx = grids[p]
            q = params[p]
            acc = np.zeros(dims[p], np.float32)
            for k in range(8):
                c, s = np.cos(q["theta"][k]), np.sin(q["theta"][k])
                arg = (xx * c + yy * s + q["drift"][k] * n) * (2 * np.pi / q["period"][k]) + q["phase"][k]
                acc += np.float32(q["amp"][k]) * np.sin(arg)
            acc = 0.5 + acc / np.float32(2 * 8 * 1.2)
            r = np.random.default_rng((seed, n, p, 7))
            img = acc * peak + r.normal(0, noise_frac * peak, size=dims[p]).astype(np.float32)
            planes.append(np.clip(np.rint(img), 0, peak).astype(dt))
        clip.append(planes)
    return clip


def yuv420_sinusoids_device(width, height, frames, bits=16, seed=2, noise_frac=0.02, device="cuda"):
    """The yuv420_sinusoids generator evaluated on the device with torch (for
    clips too large to generate on the host): the same seeded sinusoid
    parameters and formula; the Gaussian noise comes from a torch generator
    seeded with (seed), so its samples differ from the host generator's.
    Returns frames as [Y, U, V] lists of int16 tensors holding the uint16
    (bits > 8) or uint8 tensors holding the 8-bit samples."""
    import torch
    peak = (1 << bits) - 1
    dims = [(height, width), (height // 2, width // 2), (height // 2, width // 2)]
    params = []
    for p in range(3):
        r = np.random.default_rng((seed, p))
        params.append(dict(theta=r.uniform(0, np.pi, 8), period=r.uniform(4, 64, 8), phase=r.uniform(0, 2 * np.pi, 8),
                           amp=r.uniform(0.2, 1.0, 8), drift=r.uniform(0, 2, 8)))
    gen = torch.Generator(device=device).manual_seed(int(seed) & 0x7FFFFFFF)
    clip = []
    for n in range(frames):
        planes = []
        for p in range(3):
            h, w = dims[p]
            yy = torch.arange(h, device=device, dtype=torch.float32).view(h, 1)
            xx = torch.arange(w, device=device, dtype=torch.float32).view(1, w)
            q = params[p]
            acc = torch.zeros((h, w), device=device, dtype=torch.float32)
            for k in range(8):
                c, s = float(np.cos(q["theta"][k])), float(np.sin(q["theta"][k]))
                arg = (xx * c + yy * s + float(q["drift"][k]) * n) * float(2 * np.pi / q["period"][k]) + float(q["phase"][k])
                acc += float(q["amp"][k]) * torch.sin(arg)
            acc = 0.5 + acc / float(2 * 8 * 1.2)
            img = acc * peak + torch.randn((h, w), device=device, generator=gen) * (noise_frac * peak)
            v = torch.clamp(torch.round(img), 0, peak).to(torch.int32)
            planes.append(v.to(torch.uint8) if bits == 8 else v.to(torch.int16))
        clip.append(planes)
    return clip
