"""Pinning of the oracle's sampled noise models against the reference's own
known-answer tests, restated with independent binary64 arithmetic:
- DFTTest DFTNoise: tests/unit/dft_noise.cpp:10-37 (independent windows,
  direct 3D DFT, calibration E1/E2*alpha/n, tolerance 0.2% + 0.002);
- FFT3D pattern model: tests/unit/fft3d_models.cpp:108-126 (direct DFT,
  rectangular windows, linear pfactor, sampling weights r/(r+cutoff^2));
- FFT3D pshow preview: fft3d_models.cpp:143-147;
- FFT3D search: fft3d_models.cpp:152-165 (automatic search equals the
  explicit px/py choice; candidates exclude the two border rings).
The table plumbing is pinned bit-exactly: a sampled table equal to the
uniform threshold reproduces the uniform-sigma plane."""
import math

import numpy as np
import pytest


def _direct_dft3(v):
    T, S, _ = v.shape
    K = S // 2 + 1
    out = np.zeros((T, S, K), np.complex128)
    t, y, x = np.meshgrid(np.arange(T), np.arange(S), np.arange(S), indexing="ij")
    for kt in range(T):
        for ky in range(S):
            for kx in range(K):
                ph = np.exp(-2j * np.pi * (kt * t / T + ky * y / S + kx * x / S))
                out[kt, ky, kx] = (v * ph).sum()
    return out


def _raw(i, n, window):
    return 1.0 if window == 7 else 0.5 - 0.5 * math.cos(2 * math.pi * (i + 0.5) / n)


@pytest.mark.parametrize("T", [1, 3, 5])
@pytest.mark.parametrize("S", [3, 4])
@pytest.mark.parametrize("mean", [False, True])
@pytest.mark.parametrize("window", [0, 7])
@pytest.mark.parametrize("overlap_sel", [0, 1])
@pytest.mark.parametrize("alpha", [0.5, 2.0])
def test_dft_noise_model_against_direct_dft(oracle, T, S, mean, window, overlap_sel, alpha):
    overlap = 0 if overlap_sel == 0 else S - 1
    locations = [(0, 0, 0, 0), (2, 1, 1, 2), (0, 0, 0, 0)]
    # planes[f][p][y, x] = f*7 + p*13 + y + 3x reproduces the reference's fill
    frames = [[np.fromfunction(lambda y, x: f * 7 + p * 13 + y + 3 * x, (8, 12)).astype(np.uint8) for p in range(2)]
              for f in range(2 + T)]
    params = dict(tbsize=T, sbsize=S, sosize=overlap, smode=1, swin=window, twin=window, zmean=mean, ftype=0,
                  tmode=0 if T % 2 else 1)
    table = oracle.dft_noise_model(frames, locations, bits=8, alpha=alpha, **params)
    h = np.zeros((T, S, S))
    E1 = 0.0
    for z in range(T):
        for y in range(S):
            for x in range(S):
                def axis(j):
                    e = sum(_raw(i, S, window) ** 2 for i in range(j % (S - overlap), S, S - overlap))
                    return _raw(j, S, window) / math.sqrt(e)
                h[z, y, x] = np.float32(_raw(z, T, window) * _raw(y, S, window) * _raw(x, S, window) / math.sqrt(T * S * S))
                E1 += (_raw(z, T, window) * axis(y) * axis(x) / math.sqrt(T * S * S)) ** 2
    E2 = float((h.astype(np.float64) ** 2).sum())
    G = _direct_dft3(255 * h)
    expected = np.zeros(G.shape)
    for (f, p, ly, lx) in locations:
        v = np.stack([frames[f + z][p][ly:ly + S, lx:lx + S].astype(np.float64) for z in range(T)]) * h
        X = _direct_dft3(v)
        ratio = X[0, 0, 0].real / G[0, 0, 0].real if mean else 0.0
        expected += np.abs(X - ratio * G) ** 2 * E1 / E2 * alpha / len(locations)
    assert np.all(np.abs(table - expected.reshape(-1)) <= 0.002 + 0.002 * np.abs(expected.reshape(-1)))


def test_dft_noise_constant_sample_is_mean_free(oracle):
    frames = [[np.full((40, 40), 77, np.uint8)] for _ in range(3)]
    table = oracle.dft_noise_model(frames, [(0, 0, 5, 9), (0, 0, 20, 3)], bits=8, tbsize=3, sbsize=16, sosize=12,
                                   zmean=True)
    scale = oracle.dft_noise_model([[np.full((40, 40), 1, np.uint8)] for _ in range(3)], [(0, 0, 0, 0)], bits=8,
                                   tbsize=3, sbsize=16, sosize=12, zmean=False)
    assert np.abs(table).max() <= 1e-6 * np.abs(scale).max() * 77 * 77


def test_dfttest_sampled_table_plumbing_bit_exact(oracle):
    rng = np.random.default_rng(4)
    img = rng.integers(0, 256, (48, 64)).astype(np.uint8)
    params = dict(tbsize=1, sbsize=16, sosize=12, ftype=0, sigma=8.0)
    _, ws = oracle.dft_window_3d(1, 16, 12, 1, 0, 7)
    uniform = oracle.dfttest_plane([img], bits=8, **params)
    table = np.full(16 * 9, np.float32(8.0) / np.float32(ws), np.float32)
    sampled = oracle.dfttest_plane([img], bits=8, noise_table=table, **params)
    assert np.array_equal(uniform, sampled)
    zero = oracle.dfttest_plane([img], bits=8, noise_table=np.zeros_like(table), **params)
    identity = oracle.dfttest_plane([img], bits=8, **dict(params, sigma=0.0))
    assert np.array_equal(zero, identity)


@pytest.mark.parametrize("factor", [0.5, 1.0, 2.0])
def test_fft3d_pattern_model_against_direct_dft(oracle, factor):
    src = np.fromfunction(lambda y, x: ((x * 3 + y * 7) % 11) / 8, (16, 16)).astype(np.float32)
    (bx, by), power, model = oracle.fft3d_pattern(src, bits=32, floating=True, bw=4, bh=4, ow=0, oh=0,
                                                  pfactor=factor, px=1, py=1, degrid=0.0, bt=5)
    assert (bx, by) == (1, 1)
    X = np.fft.rfft2(src[:4, :4].astype(np.float64))
    for y in range(4):
        for x in range(3):
            r = (min(y, 4 - y) / 2) ** 2 + (x / 2) ** 2
            want = factor * abs(X[y, x]) ** 2 * r / (r + 0.01)
            assert abs(model[y * 3 + x] - want) <= 2e-4 * max(1.0, abs(want))


def test_fft3d_preview_places_selected_block(oracle):
    inp = np.full((16, 16), 128, np.uint8)
    out = oracle.fft3d_plane([inp, inp, inp], bits=8, bw=4, bh=4, ow=0, oh=0, pfactor=1.0, pshow=True)
    want = np.zeros((16, 16), np.uint8)
    want[4:8, 4:8] = 128
    assert np.array_equal(out, want)


def test_fft3d_search_first_minimum_and_border_exclusion(oracle):
    src = np.zeros((20, 20), np.float32)
    for y in range(4, 16):
        for x in range(4, 16):
            src[y, x] = (1e8 if (8 <= x < 12 and y >= 12) else 3e8) * (1 if (x + y) % 2 else -1)
    cfg = dict(bits=32, floating=True, bw=4, bh=4, ow=0, oh=0, pfactor=1.0, degrid=0.0)
    (ax, ay), pa, ma = oracle.fft3d_pattern(src, **cfg)
    (cx, cy), pc, mc = oracle.fft3d_pattern(src, **dict(cfg, px=3, py=4))
    assert (ax, ay) == (3, 4) and (cx, cy) == (3, 4)
    assert np.array_equal(ma, mc)
    assert ma[8] > 1e15
    # every candidate of a flat plane ties: the first (top-left interior) wins
    (fx, fy), _, _ = oracle.fft3d_pattern(np.ones((20, 20), np.float32), **cfg)
    assert (fx, fy) == (2, 2)
