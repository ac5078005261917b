"""Batched multi-view rendering: every scene of a batch equals the same scene
rendered alone (ids and depth bit-exact, colour identical), for two different
tunings; graph replays equal the eager render; the autotuner returns the
fastest of its candidates."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _views(n=6):
    from taichi_three_amd import workloads as wl
    return wl.batch_views(n=n, seed=2, faces=(1000, 40_000))


def _single(w):
    import torch
    from taichi_three_amd import workloads as wl
    scene, mesh = wl.build_scene(w)
    scene.render()
    torch.cuda.synchronize()
    ids = scene.engine.ids.to_numpy().T.copy()
    depth = scene.engine.depth.to_numpy().T.copy()
    img = scene.img.to_numpy().transpose(1, 0, 2).copy()
    return ids, depth, img


@pytest.mark.parametrize('tuning', [(4096, 64, 0, 0, 0), (1024, 16, 256, 1024, 1024)])
def test_batch_equals_single_scene_renders(gpu, tuning):
    import torch
    from taichi_three_amd.batch import BatchRenderer
    ws = _views()
    assert {tuple(w['res']) for w in ws} >= {(1920, 1080)} or len({tuple(w['res']) for w in ws}) >= 2
    print(f'case: batch of {len(ws)} tuning {tuning}', flush=True)
    plan = BatchRenderer(ws, tuning)
    plan.render()
    torch.cuda.synchronize()
    for i, w in enumerate(ws):
        ids, depth, img = _single(w)
        bids, bdep = plan.resolve(i)
        bimg = plan.image(i).cpu().numpy()
        assert (ids >= 0).sum() > 0
        assert np.array_equal(bids.cpu().numpy(), ids), i
        assert np.array_equal(bdep.cpu().numpy(), depth), i
        err = float(np.abs(bimg - img).max())
        assert err == 0.0, (i, err)
    plan.close()


def test_graph_replay_matches_eager(gpu):
    """Three replays of the captured graph, with eager work and synchronisation
    between them, each equal to the eager render."""
    import torch
    from taichi_three_amd.batch import BatchRenderer
    ws = _views(4)
    print('case: batch graph replay', flush=True)
    plan = BatchRenderer(ws)
    plan.render()
    torch.cuda.synchronize()
    ref = [plan.image(i).cpu().numpy().copy() for i in range(len(ws))]
    ref_ids = [plan.resolve(i)[0].cpu().numpy().copy() for i in range(len(ws))]
    plan.capture()
    for rep in range(3):
        for img in plan.scenes.images:
            img.fill_(float(rep + 7))
        torch.cuda.synchronize()
        plan.replay()
        torch.cuda.synchronize()
        for i in range(len(ws)):
            assert np.array_equal(plan.image(i).cpu().numpy(), ref[i]), (rep, i)
            assert np.array_equal(plan.resolve(i)[0].cpu().numpy(), ref_ids[i]), (rep, i)
    plan.close()


def test_autotuner_returns_the_fastest_candidate(gpu):
    from taichi_three_amd import autotune
    from taichi_three_amd.batch import BatchScenes, Tuning
    print('case: autotune', flush=True)
    scenes = BatchScenes(_views(4))
    space = [Tuning(4096, 64), Tuning(2048, 16, 256, 1024, 1024)]
    best, results = autotune.tune(scenes, space=space, reps=3)
    assert len(results) == 2 and all(np.isfinite(ms) and ms > 0 for _, ms in results)
    assert best == min(results, key=lambda r: r[1])[0]
