"""DFTTest block sizes outside the 8/16/32-point codelets and the centre mode
(smode=0): the direct-DFT kernels against the oracle, which evaluates the
same direct DFTs (binary64 twiddles, index-order sums), so integer and float
outputs must match bit for bit. The oracle's direct DFT itself is pinned to a
binary64 DFT in test_oracle_kats.py; the reference's own PocketFFT bits for
these lengths are not reproduced by either (parity unpinned)."""
import numpy as np
import pytest

from gpu_util import empty_like_np, to_device, to_host

pytestmark = pytest.mark.gpu


def _plane(rng, h, w, bits, floating=False):
    if floating:
        yy, xx = np.mgrid[0:h, 0:w]
        img = 0.5 + 0.3 * np.sin(xx * 0.21 + yy * 0.13) + rng.normal(0, 0.05, (h, w))
        return img.astype(np.float32)
    peak = (1 << bits) - 1
    yy, xx = np.mgrid[0:h, 0:w]
    img = peak * (0.5 + 0.3 * np.sin(xx * 0.21 + yy * 0.13)) + rng.normal(0, peak * 0.04, (h, w))
    return np.clip(np.rint(img), 0, peak).astype(np.uint8 if bits == 8 else np.uint16)


def _run(nf, slots, bits, floating=False, targets=None, **params):
    h, w = slots[0].shape
    plan = nf.dfttest_plan(w, h, bits=bits, floating=floating, **params)
    dst = empty_like_np(slots[0])
    plan.process([to_device(s) for s in slots], dst, targets=targets)
    return to_host(dst, slots[0].dtype)


@pytest.mark.parametrize("B,O", [(12, 6), (24, 12), (6, 3), (20, 5), (10, 5)])
@pytest.mark.parametrize("T", [1, 3])
def test_generic_sizes_match_oracle(nf, oracle, B, O, T):
    rng = np.random.default_rng(B * 10 + T)
    slots = [_plane(rng, 70, 90, 16) for _ in range(T)]
    params = dict(tbsize=T, sbsize=B, sosize=O, sigma=8.0)
    print(f"generic B={B} O={O} T={T}", flush=True)
    got = _run(nf, slots, 16, **params)
    ref = oracle.dfttest_plane(slots, bits=16, **params)
    assert np.array_equal(got, ref)


@pytest.mark.parametrize("ftype,zmean", [(1, True), (2, False), (3, True), (4, False)])
def test_generic_gain_types(nf, oracle, ftype, zmean):
    rng = np.random.default_rng(ftype)
    slots = [_plane(rng, 50, 60, 32, floating=True) for _ in range(3)]
    params = dict(tbsize=3, sbsize=12, sosize=6, sigma=4.0, sigma2=2.0, pmin=0.5, pmax=400.0, ftype=ftype,
                  zmean=zmean)
    got = _run(nf, slots, 32, floating=True, **params)
    ref = oracle.dfttest_plane(slots, bits=32, floating=True, **params)
    assert np.array_equal(got, ref)


@pytest.mark.parametrize("B", [5, 7, 15])
@pytest.mark.parametrize("T", [1, 3])
def test_center_mode_matches_oracle(nf, oracle, B, T):
    rng = np.random.default_rng(B + 100 * T)
    slots = [_plane(rng, 40, 52, 8) for _ in range(T)]
    params = dict(tbsize=T, sbsize=B, smode=0, sigma=6.0)
    print(f"centre mode B={B} T={T}", flush=True)
    got = _run(nf, slots, 8, **params)
    ref = oracle.dfttest_plane(slots, bits=8, **params)
    assert np.array_equal(got, ref)


def test_center_mode_temporal_blocks(nf, oracle):
    """smode=0 with tmode=1: every covering temporal block adds its centre
    sample (plan.cpp:596-599)."""
    rng = np.random.default_rng(8)
    frames = [_plane(rng, 36, 44, 16) for _ in range(4)]
    T, tosize = 2, 1
    params = dict(tbsize=T, tmode=1, tosize=tosize, sbsize=7, smode=0, sigma=6.0)
    blocks = oracle.temporal_blocks(2, len(frames), T, tosize)
    slots = [frames[s] for _, _, sl in blocks for s in sl]
    targets = [t for _, t, _ in blocks]
    got = _run(nf, slots, 16, targets=targets, **params)
    ref = oracle.dfttest_plane(slots, bits=16, targets=targets, **params)
    assert np.array_equal(got, ref)


@pytest.mark.parametrize("B,smode", [(16, 1), (12, 1), (9, 0)])
def test_tmode1_tbsize1(nf, oracle, B, smode):
    """tmode=1 with tbsize=1 synthesises (U*V)*h (plan.cpp:594-595), not
    U*(h*V): both device paths follow it."""
    rng = np.random.default_rng(B)
    img = _plane(rng, 48, 64, 32, floating=True)
    params = dict(tbsize=1, tmode=1, tosize=0, sbsize=B, sosize=B // 2 if smode else 0, smode=smode, sigma=5.0)
    got = _run(nf, [img], 32, floating=True, targets=[0], **params)
    ref = oracle.dfttest_plane([img], bits=32, floating=True, targets=[0], **params)
    assert np.array_equal(got, ref)


def test_generic_dither(nf, oracle):
    rng = np.random.default_rng(2)
    img = _plane(rng, 40, 56, 8)
    params = dict(tbsize=1, sbsize=12, sosize=6, sigma=8.0, dither=2, dither_seed=3)
    h, w = img.shape
    plan = nf.dfttest_plan(w, h, bits=8, **params)
    dst = empty_like_np(img)
    plan.process([to_device(img)], dst, frame=4, plane=0)
    ref = oracle.dfttest_plane([img], bits=8, frame=4, plane=0, **params)
    assert np.array_equal(to_host(dst, np.uint8), ref)


def test_generic_batch_and_nonfinite(nf, oracle):
    import torch
    rng = np.random.default_rng(6)
    planes = [_plane(rng, 40, 48, 16), _plane(rng, 30, 34, 16)]
    params = dict(tbsize=1, sbsize=12, sosize=6, sigma=8.0)
    jobs = []
    for pl in planes:
        h, w = pl.shape
        jobs.append((nf.dfttest_plan(w, h, bits=16, **params), [to_device(pl)], empty_like_np(pl)))
    flags = torch.zeros(1, dtype=torch.int32, device="cuda")
    nf.process_batch(jobs, flags)
    torch.cuda.synchronize()
    assert int(flags.item()) == 0
    for (_, _, dst), pl in zip(jobs, planes):
        assert np.array_equal(to_host(dst, np.uint16), oracle.dfttest_plane([pl], bits=16, **params))
    img = _plane(rng, 30, 40, 32, floating=True)
    img[3, 4] = np.nan
    plan = nf.dfttest_plan(40, 30, bits=32, floating=True, tbsize=1, sbsize=9, smode=0)
    with pytest.raises(nf.NonFinite):
        plan.process([to_device(img)], empty_like_np(img))


def test_generic_rejections(nf):
    with pytest.raises(nf.InvalidArgument):
        nf.dfttest_plan(64, 48, bits=8, tbsize=1, sbsize=8, smode=0)  # centre mode needs an odd sbsize
    with pytest.raises(nf.Unsupported):
        nf.dfttest_plan(256, 200, bits=8, tbsize=15, sbsize=64, sosize=32)  # block volume beyond the LDS


# ------------------------------------------------------------------ FFT3D
@pytest.mark.parametrize("bw,bh", [(24, 24), (32, 16), (16, 32), (12, 20), (48, 48)])
@pytest.mark.parametrize("bt", [1, 3])
def test_fft3d_generic_blocks_match_oracle(nf, oracle, bw, bh, bt):
    rng = np.random.default_rng(bw * 100 + bh + bt)
    slots = [_plane(rng, 72, 96, 16) for _ in range(bt)]
    params = dict(bt=bt, bw=bw, bh=bh, sigma=3.0)
    print(f"fft3d generic bw={bw} bh={bh} bt={bt}", flush=True)
    plan = nf.fft3d_plan(96, 72, bits=16, **params)
    dst = empty_like_np(slots[0])
    plan.process([to_device(s) for s in slots], dst)
    ref = oracle.fft3d_plane(slots, bits=16, **params)
    assert np.array_equal(to_host(dst, np.uint16), ref)


@pytest.mark.parametrize("extra", [dict(bt=-1, sharpen=0.5, dehalo=0.3),
                                   dict(bt=1, sigma=3.0, sigma2=2.0, sigma3=1.0, sigma4=4.0, degrid=0.5),
                                   dict(bt=1, sigma=3.0, sharpen=0.4, beta=1.5),
                                   dict(bt=2, sigma=3.0, sharpen=0.3)])
def test_fft3d_generic_models(nf, oracle, extra):
    rng = np.random.default_rng(5)
    bt = max(1, extra["bt"])
    slots = [_plane(rng, 60, 80, 8) for _ in range(bt)]
    params = dict(bw=20, bh=12, **extra)
    plan = nf.fft3d_plan(80, 60, bits=8, chroma=True, **params)
    dst = empty_like_np(slots[0])
    plan.process([to_device(s) for s in slots], dst)
    ref = oracle.fft3d_plane(slots, bits=8, chroma=True, **params)
    assert np.array_equal(to_host(dst, np.uint8), ref)


def test_fft3d_generic_kalman(nf, oracle):
    rng = np.random.default_rng(9)
    params = dict(bt=0, sigma=2.0, bw=24, bh=24, kratio=2.0)
    frames = [_plane(rng, 72, 96, 10) & 1023 for _ in range(5)]
    plan = nf.fft3d_plan(96, 72, bits=10, **params)
    state = nf.KalmanState(plan)
    ref = oracle.KalmanState(params, 96, 72, bits=10)
    for n, fr in enumerate(frames):
        src = to_device(fr)
        state.advance(src)
        ref.advance(fr)
        dst = empty_like_np(fr)
        state.render(src, dst)
        assert np.array_equal(to_host(dst, np.uint16), ref.render()), n


def test_fft3d_generic_spectrum_cache(nf, oracle):
    import torch
    rng = np.random.default_rng(13)
    frames = [_plane(rng, 60, 72, 16) for _ in range(6)]
    params = dict(bt=3, bw=24, bh=12, sigma=3.0)
    plan = nf.fft3d_plan(72, 60, bits=16, **params)
    nf.set_spectrum_cache(plan, 4)
    dev = [to_device(f) for f in frames]
    flags = torch.zeros(1, dtype=torch.int32, device="cuda")
    for n in range(1, 5):
        dst = empty_like_np(frames[0])
        nf.process_batch([(plan, [dev[n - 1], dev[n], dev[n + 1]], dst)], flags, frames=[n])
        torch.cuda.synchronize()
        ref = oracle.fft3d_plane(frames[n - 1:n + 2], bits=16, **params)
        assert np.array_equal(to_host(dst, np.uint16), ref), n
    assert int(flags.item()) == 0


def test_fft3d_generic_rejects_oversized_blocks(nf):
    with pytest.raises(nf.Unsupported):
        nf.fft3d_plan(400, 300, bits=8, bt=5, bw=120, bh=120)


# ---------------------------------------------------- sampled noise models
@pytest.mark.parametrize("B,T,smode", [(12, 1, 1), (12, 3, 1), (20, 1, 1), (9, 1, 0), (9, 3, 0)])
def test_generic_dft_noise_model(nf, oracle, B, T, smode):
    rng = np.random.default_rng(B * 7 + T)
    h, w = 70, 90
    frames = [[rng.integers(0, 256, (h, w)).astype(np.uint8)] for _ in range(T + 3)]
    locations = [(0, 0, 3, 5), (2, 0, h - B, w - B), (1, 0, 20, 30)]
    params = dict(tbsize=T, tmode=0, sbsize=B, sosize=B * 3 // 4 if smode else 0, smode=smode, ftype=0, sigma=8.0,
                  zmean=True)
    plan = nf.dfttest_plan(w, h, bits=8, nlocation=[v for loc in locations for v in loc], **params)
    dev = [[to_device(p) for p in fr] for fr in frames]
    nf.dfttest_prepare_noise(plan, [dev[f + z][p] for (f, p, _, _) in locations for z in range(T)])
    ref_table = oracle.dft_noise_model(frames, locations, bits=8, **params)
    assert np.array_equal(nf.read_model(plan), ref_table)
    slots = [frames[z][0] for z in range(T)]
    dst = empty_like_np(slots[0])
    plan.process([to_device(s) for s in slots], dst)
    ref = oracle.dfttest_plane(slots, bits=8, noise_table=ref_table, **params)
    assert np.array_equal(to_host(dst, np.uint8), ref)


@pytest.mark.parametrize("bw,bh", [(24, 24), (32, 24), (20, 12)])
@pytest.mark.parametrize("bt", [1, 3])
def test_generic_fft3d_pattern(nf, oracle, bw, bh, bt):
    rng = np.random.default_rng(bw + bh + bt)
    h, w = 120, 150
    frames = [rng.integers(0, 256, (h, w)).astype(np.uint8) for _ in range(bt)]
    params = dict(bw=bw, bh=bh, bt=bt, pfactor=1.5, sharpen=0.3, sigma=3.0)
    plan = nf.fft3d_plan(w, h, bits=8, **params)
    sel = nf.fft3d_prepare_pattern(plan, to_device(frames[0]))
    (rx, ry), _, ref_model = oracle.fft3d_pattern(frames[0], bits=8, **params)
    assert sel == (rx, ry)
    assert np.array_equal(nf.read_model(plan), ref_model)
    dst = empty_like_np(frames[0])
    plan.process([to_device(f) for f in frames], dst)
    ref = oracle.fft3d_plane(frames, bits=8, pattern=ref_model, **params)
    assert np.array_equal(to_host(dst, np.uint8), ref)
    preview = nf.fft3d_plan(w, h, bits=8, pshow=True, **params)
    dst = empty_like_np(frames[0])
    preview.process([to_device(f) for f in frames], dst)
    ref = oracle.fft3d_plane(frames, bits=8, pshow=True, **params)
    assert np.array_equal(to_host(dst, np.uint8), ref)
