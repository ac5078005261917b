"""The benchmark workloads at their full sizes against the oracle.

* soup_1080p: 1920x1080, 2,000,000 faces (the headline frame): ids and depth
  bit for bit against the oracle's raster (run on every host core), colour
  within 1e-5 of the binary32 and 1e-4 of the binary64 shading oracle (the
  latter wherever the binary32 oracle itself meets it).
* batch_views: the 64-view batch (1k-500k faces, 256^2 to 1920x1080) rendered
  in one launch sequence; every scene against the oracle's raster and shade.
* ibl_1080p with the learned surrogate: 1920x1080, 1M faces, trained on 16M
  analytic pairs; >= 40 dB against the analytic frame, every fallback pixel
  bit-identical to the analytic frame, every other pixel the network's own.
"""
import math
import os

import numpy as np
import pytest

from oracle import orc

pytestmark = pytest.mark.gpu

COLOR_TOL_F32 = 1e-5   # abs, tone-mapped colour vs the binary32 oracle
COLOR_TOL_F64 = 1e-4   # abs, tone-mapped colour vs the binary64 oracle
NTHREADS = max(1, len(os.sched_getaffinity(0)))


def _oracle_frame(w, ids_gpu=None):
    """(ids, depth, img32, img64) of workload w from the oracle: parallel raster
    with the serial argmin rule (equal to the serial raster, which the CPU tests
    check), then the binary32 and binary64 deferred shade."""
    from taichi_three_amd import workloads as wl, _abi
    from taichi_three_amd.matr import lower_material
    W, H = w.res
    cam = orc.camera(w.W2V, w.V2W)
    keys = orc.raster_omp(W, H, cam, w.verts, flags=3, nthreads=NTHREADS)
    depth, ids = orc.unpack(keys)
    lo = lower_material(wl.make_material(w.material), dict(attrs=w.attrs))
    flags = _abi.VCOLOR | _abi.BG | _abi.TONEMAP | (_abi.SMOOTH if w.smoothing else 0)
    mesh = orc.HostMesh(w.verts, w.norms if w.smoothing else None, None, w.colors)
    lt = wl.make_lighting(w).to_struct()
    img32 = orc.shade(W, H, cam, mesh, lo.struct, lt, ids, flags, nthreads=NTHREADS)
    img64 = orc.shade(W, H, cam, mesh, lo.struct, lt, ids, flags, f64=True, nthreads=NTHREADS)
    return ids, depth, img32, img64


def _check_colour(img, o32, o64, tag):
    """Colour within 1e-5 of the binary32 oracle at every pixel, and within 1e-4
    of the binary64 oracle wherever the binary32 oracle itself is: at sharp
    highlights (roughness 0.05 makes D change by ~6e5 per unit of N.H) the
    reference's own binary32 arithmetic strays past 1e-4 of binary64 on a few
    pixels, where no result within 1e-5 of it can meet both bounds; those pixels
    stay within 1e-5 of the binary32 value and are counted."""
    e32 = float(np.abs(img - o32).max())
    print(f'{tag}: colour max |gpu - binary32| = {e32:.2e}', flush=True)
    assert e32 <= COLOR_TOL_F32, (tag, e32)
    d64 = np.abs(img - o64).max(axis=2)
    ref_ok = np.abs(o32 - o64).max(axis=2) <= COLOR_TOL_F64
    e64 = float(d64[ref_ok].max()) if ref_ok.any() else 0.0
    assert e64 <= COLOR_TOL_F64, (tag, e64)
    nbad = int((~ref_ok).sum())
    assert nbad <= 1e-4 * ref_ok.size, (tag, nbad)
    return e32, e64, nbad


def test_soup_1080p_2m_full_size(gpu):
    """The headline frame itself: 245 binning groups, 8,160 tiles."""
    import torch
    from taichi_three_amd import workloads as wl
    w = wl.soup_1080p()
    print(f'case: soup_1080p 1920x1080 2M seed 1 (oracle on {NTHREADS} threads)', flush=True)
    scene, mesh = wl.build_scene(w)
    scene.render()
    torch.cuda.synchronize()
    ids = scene.engine.ids.to_numpy().T
    depth = scene.engine.depth.to_numpy().T
    img = scene.img.to_numpy().transpose(1, 0, 2)
    del scene, mesh
    oids, odepth, o32, o64 = _oracle_frame(w)
    assert (oids >= 0).sum() > 500_000
    assert np.array_equal(ids, oids), f'{(ids != oids).sum()} id mismatches'
    assert np.array_equal(depth, odepth), f'{(depth != odepth).sum()} depth mismatches'
    e32, e64, nbad = _check_colour(img, o32, o64, 'soup_1080p')
    print(f'soup_1080p: {int((oids >= 0).sum())} covered, colour err {e32:.2e} (f32) {e64:.2e} (f64), '
          f'{nbad} pixels where the binary32 oracle itself is beyond 1e-4 of binary64', flush=True)


def test_batch_64_views_full_size(gpu):
    """All 64 scenes of the batch workload in one render: ids and depth bit for
    bit, colour within the tolerances, against the oracle scene by scene."""
    import torch
    from taichi_three_amd import workloads as wl
    from taichi_three_amd.batch import BatchRenderer
    ws = wl.batch_views()
    assert len(ws) == 64
    print(f'case: batch_views 64 scenes, {sum(len(w.verts) for w in ws)} faces', flush=True)
    plan = BatchRenderer(ws)
    plan.render()
    torch.cuda.synchronize()
    worst = (0.0, 0.0, 0)
    for i, w in enumerate(ws):
        bids, bdep = plan.resolve(i)
        ids, depth = bids.cpu().numpy(), bdep.cpu().numpy()
        img = plan.image(i).cpu().numpy()
        oids, odepth, o32, o64 = _oracle_frame(w)
        assert np.array_equal(ids, oids), (i, w.res, len(w.verts), int((ids != oids).sum()))
        assert np.array_equal(depth, odepth), i
        e = _check_colour(img, o32, o64, i)
        worst = (max(worst[0], e[0]), max(worst[1], e[1]), worst[2] + e[2])
    plan.close()
    print(f'batch: 64 scenes match, colour err {worst[0]:.2e} (f32) {worst[1]:.2e} (f64), '
          f'{worst[2]} pixels where the binary32 oracle itself is beyond 1e-4 of binary64', flush=True)


def _psnr(a, b):
    mse = float(np.mean((a.astype(np.float64) - b.astype(np.float64)) ** 2))
    return 10 * math.log10(1.0 / max(mse, 1e-20))


def test_surrogate_1080p_1m_16m_samples(gpu):
    """The learned-shading workload at its size: the network trained on 16M
    analytic pairs shades the 1080p/1M-face IBL frame to >= 40 dB of the analytic
    frame; the pixels it hands to the analytic fallback are bit-identical to the
    analytic frame and every other pixel is the network's own value (the same
    frame shaded with the fallback switched off)."""
    import torch
    from taichi_three_amd import workloads as wl
    from taichi_three_amd.neural import NeuralShading
    print('case: surrogate ibl_1080p 1M faces, 16M samples', flush=True)
    w = wl.ibl_1080p()
    scene, mesh = wl.build_scene(w)
    scene.render()
    torch.cuda.synchronize()
    ref = scene.img.to_numpy().copy()
    ns = NeuralShading(scene.lighting, seed=1)
    scene.neural = ns
    scene.render()
    untrained = scene.img.to_numpy().copy()
    ns.train(nsamples=1 << 24, steps=30000, batch=16384)
    scene.render()
    torch.cuda.synchronize()
    img = scene.img.to_numpy().copy()
    nfall = ns.fallback_count(scene.engine)
    thr = ns.threshold
    ns.threshold = float('inf')
    scene.render()
    torch.cuda.synchronize()
    pure = scene.img.to_numpy().copy()
    ns.threshold = thr
    covered = int((scene.engine.ids.to_numpy() >= 0).sum())
    p_tr, p_un = _psnr(img, ref), _psnr(untrained, ref)
    print(f'psnr trained {p_tr:.2f} dB, untrained {p_un:.2f} dB, fallback {nfall} of {covered}', flush=True)
    assert p_tr >= 40.0
    assert p_tr - p_un >= 15.0
    moved = np.any(img != pure, axis=2)
    assert np.array_equal(img[moved], ref[moved]), 'a fallback pixel differs from the analytic frame'
    # the pixels the fallback rewrote are the queued ones (minus any whose network
    # value happened to equal the analytic one bit for bit)
    assert int(moved.sum()) <= nfall
    assert int(moved.sum()) >= 0.99 * nfall
    assert nfall < covered
