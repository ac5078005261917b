"""A frame's raster captured in a graph (torch.cuda.CUDAGraph) replays as it
ran: the binned raster keeps no counter that one call leaves for the next to
reset, so every replay of the captured clear + raster (with the keys scribbled
over and other eager work between replays) gives the eager ids and depth."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize('n', [60_000, 1_200_000])
def test_captured_raster_replays_like_eager(gpu, n):
    import torch
    import taichi_three_amd as tina
    from taichi_three_amd import workloads as wl
    print(f'case: captured raster n={n}', flush=True)
    rng = np.random.default_rng(7)
    W, H = 640, 360
    verts = wl.triangle_soup(rng, n, np.log(0.02), 0.6)
    verts[:50] = wl.triangle_soup(rng, 50, np.log(1.0), 0.2)  # giant faces
    eng = tina.Engine(W, H, maxfaces=n)
    eng.set_camera(*wl.default_camera(W / H))
    r = tina.TriangleRaster(eng, maxfaces=n)
    m = tina.SimpleMesh(maxfaces=n)
    m.set_face_verts(verts)

    def frame():
        eng.clear_depth()
        r.set_object(m)
        r.render_occup()

    frame()
    ids0, d0 = eng.resolve()
    torch.cuda.synchronize()
    ids0, d0 = ids0.clone(), d0.clone()
    assert (ids0 >= 0).sum() > 0.1 * W * H
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        frame()
    torch.cuda.current_stream().wait_stream(side)
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        frame()
    keys = eng.keys()
    for k in range(4):
        keys.fill_(k)  # scribble: the captured clear is applied by the tiles themselves
        frame()        # and an eager frame between replays
        g.replay()
        ids, d = eng.resolve()
        torch.cuda.synchronize()
        assert torch.equal(ids, ids0), f'replay {k}: {(ids != ids0).sum().item()} ids differ'
        assert torch.equal(d, d0), f'replay {k}'


def test_captured_neural_frame_replays_like_eager(gpu):
    """The surrogate's pass (k_shade_nn + the analytic fallback queue) captured
    with the rest of Scene.render(): its queue counters are emptied by the pass
    itself, so each replay shades the same pixels the eager frame does (an
    untrained network, so that many pixels take the fallback)."""
    import torch
    from taichi_three_amd import workloads as wl
    from taichi_three_amd.neural import NeuralShading
    print('case: captured neural frame 320x180', flush=True)
    w = wl.ibl_1080p(seed=6, res=(320, 180), nfaces=50_000, sky=(128, 64), precision=4)
    scene, mesh = wl.build_scene(w)
    ns = NeuralShading(scene.lighting, seed=3)
    scene.neural = ns
    scene.render()
    torch.cuda.synchronize()
    ref = scene.img.to_numpy().copy()
    nref = ns.fallback_count(scene.engine)
    assert 0 < nref < 320 * 180
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        scene.render()
    torch.cuda.current_stream().wait_stream(side)
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        scene.render()
    for k in range(3):
        scene.image.storage.fill_(float(k + 1))
        scene.render()
        g.replay()
        torch.cuda.synchronize()
        img = scene.img.to_numpy()
        assert np.array_equal(img, ref), f'replay {k}: {int((img != ref).any(axis=-1).sum())} pixels differ'
        assert ns.fallback_count(scene.engine) == nref
