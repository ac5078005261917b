"""Differentiable shading on the HIP path: the forward against the binary64
oracle, the adjoint kernel against central finite differences of the binary64
oracle (h = 1e-3) and against a pure-PyTorch autograd reference, on a 64x64
render of a 492-vertex icosphere; then an optimisation at 1024x1024 on the
100,002-vertex icosphere."""
import numpy as np
import pytest

from oracle import diffshade as ds
from test_diff_oracle import diff_case

pytestmark = pytest.mark.gpu

NAMES = ('colors', 'normals', 'rough', 'metal', 'intens')


def _setup(res=64, freq=7, seed=0):
    import torch
    import taichi_three_amd as tina
    from taichi_three_amd import workloads as wl
    from taichi_three_amd.diff import DiffShading
    W, H, W2V, V2W, verts, faces, p, lights, amb = diff_case(freq=freq, res=res, seed=seed)
    eng = tina.Engine(W, H, maxfaces=len(faces))
    view, proj = wl.default_camera(1.0)
    eng.set_camera(view, proj)
    d = DiffShading(eng, verts, faces, lights, ambient=amb)
    d.rasterize()
    ids = eng.ids.to_numpy().T.copy()       # [H][W]
    tp = {k: torch.from_numpy(v).cuda().contiguous() for k, v in p.items()}
    return d, ids, (W, H, W2V, V2W, verts, faces, lights, amb), p, tp


def _oracle_img(case, ids, p):
    W, H, W2V, V2W, verts, faces, lights, amb = case
    return ds.render(W, H, W2V, V2W, verts, faces, ids, p['normals'], p['colors'], p['rough'], p['metal'],
                     lights, p['intens'], amb)


def test_forward_matches_the_binary64_oracle(gpu):
    import torch
    print('case: dshade forward 64x64 icosphere(7)', flush=True)
    d, ids, case, p, tp = _setup()
    img = d.forward_raw(tp['colors'], tp['normals'], tp['rough'], tp['metal'], tp['intens'])
    torch.cuda.synchronize()
    ref = _oracle_img(case, ids, p)
    assert (ids >= 0).sum() > 500
    err = float(np.abs(img.cpu().numpy() - ref).max())
    assert err <= 1e-4 * max(1.0, float(np.abs(ref).max())), err


def _gpu_grads(d, tp, target):
    import torch
    leaves = {k: v.clone().requires_grad_(True) for k, v in tp.items()}
    img = d.render(leaves['colors'], leaves['normals'], leaves['rough'], leaves['metal'], leaves['intens'])
    loss = 0.5 * ((img - target) ** 2).sum()
    loss.backward()
    return {k: leaves[k].grad.detach().cpu().numpy().astype(np.float64) for k in NAMES}


def _target(case, ids, seed=1):
    W, H, W2V, V2W, verts, faces, lights, amb = case
    _, _, _, _, _, _, q, _, _ = diff_case(seed=seed)
    return ds.render(W, H, W2V, V2W, verts, faces, ids, q['normals'], q['colors'], q['rough'], q['metal'],
                     lights, q['intens'], amb)


def test_adjoint_matches_central_differences(gpu):
    """dLoss/dparam of L = 0.5 |img - target|^2 from the adjoint kernel vs
    (L(p + h) - L(p - h)) / 2h with h = 1e-3, L evaluated in binary64."""
    import torch
    print('case: dshade adjoint vs FD h=1e-3', flush=True)
    d, ids, case, p, tp = _setup()
    tgt = _target(case, ids)
    g = _gpu_grads(d, tp, torch.from_numpy(tgt.astype(np.float32)).cuda())
    faces = case[5]
    visible = np.unique(faces[np.unique(ids[ids >= 0])].reshape(-1))
    rng = np.random.default_rng(2)
    p64 = {k: v.astype(np.float64) for k, v in p.items()}

    def L(q):
        img = _oracle_img(case, ids, q)
        return 0.5 * float(((img - tgt) ** 2).sum())

    h = 1e-3
    for name in NAMES:
        arr = p64[name]
        if name == 'intens':
            idx = [(i,) for i in range(arr.shape[0])]
        else:
            vs = rng.choice(visible, 8, replace=False)
            idx = [(v,) if arr.ndim == 1 else (v, int(rng.integers(3))) for v in vs]
        scale = np.abs(g[name]).max()
        for ix in idx:
            qp = {k: v.copy() for k, v in p64.items()}
            qm = {k: v.copy() for k, v in p64.items()}
            qp[name][ix] += h
            qm[name][ix] -= h
            fd = (L(qp) - L(qm)) / (2 * h)
            assert abs(g[name][ix] - fd) <= 2e-3 * scale + 1e-3 * abs(fd), (name, ix, g[name][ix], fd, scale)


def _torch_reference_grads(case, ids, p, target):
    """The same shading written with plain PyTorch ops in float64 on the CPU,
    differentiated by autograd."""
    import torch
    W, H, W2V, V2W, verts, faces, lights, amb = case
    pix, f, w = ds.weights(W, H, W2V, verts, faces, ids)
    V = torch.from_numpy(ds.view_dirs(W, H, V2W, pix))
    w = torch.from_numpy(w)
    fv = torch.from_numpy(faces[f].astype(np.int64))
    leaves = {k: torch.from_numpy(v.astype(np.float64)).requires_grad_(True) for k, v in p.items()}

    def interp(A):
        return (w[..., None] * A[fv]).sum(1) if A.dim() == 2 else (w * A[fv]).sum(1)
    pos = interp(torch.from_numpy(verts.astype(np.float64)))
    nr = interp(leaves['normals'])
    n = nr / torch.sqrt((nr * nr).sum(1, keepdim=True))
    a, r, m = interp(leaves['colors']), interp(leaves['rough']), interp(leaves['metal'])
    f0 = m[:, None] * a + (1 - m[:, None]) * 0.04
    C = torch.tensor(amb, dtype=torch.float64)[None] * ((1 - f0) * a + f0)
    r2 = torch.clamp(r * r, min=1e-6)
    for l, (lp, lc) in enumerate(lights):
        Lv = torch.tensor(lp, dtype=torch.float64)[None] - pos
        dist = torch.sqrt((Lv * Lv).sum(1))
        Lv = Lv / dist[:, None]
        cos = (n * Lv).sum(1)
        hv = Lv + V
        hv = hv / torch.sqrt((hv * hv).sum(1, keepdim=True))
        NoH = torch.clamp((hv * n).sum(1), min=1e-10)
        NoL = torch.clamp((Lv * n).sum(1), min=1e-10)
        NoV = torch.clamp((V * n).sum(1), min=1e-10)
        VoH = torch.clamp((hv * V).sum(1), 1e-10, 1 - 1e-10)
        den = 1 - NoH * NoH * (1 - r2)
        ndf = r2 / (den * den)
        k = r2 / 2
        vdf = 1 / ((NoV * k + 1 - k) * (NoL * k + 1 - k)) / ((1 - r2) + 4 * np.pi * r2)
        F = f0 + (1 - f0) * ((1 - VoH) ** 5)[:, None]
        brdf = (1 - f0) * a / np.pi + f0 * (ndf * vdf)[:, None] * F
        lit = cos[:, None] * (leaves['intens'][l] * torch.tensor(lc, dtype=torch.float64)[None]
                              / (dist * dist)[:, None]) * brdf
        C = C + torch.where((cos > 0)[:, None], lit, torch.zeros_like(lit))
    tgt = torch.from_numpy(target.reshape(-1, 3)[pix])
    loss = 0.5 * ((C - tgt) ** 2).sum()
    loss.backward()
    return {k: leaves[k].grad.numpy() for k in NAMES}


def test_adjoint_matches_pytorch_autograd(gpu):
    import torch
    print('case: dshade adjoint vs pytorch autograd', flush=True)
    d, ids, case, p, tp = _setup()
    tgt = _target(case, ids)
    g = _gpu_grads(d, tp, torch.from_numpy(tgt.astype(np.float32)).cuda())
    ref = _torch_reference_grads(case, ids, p, tgt)
    for name in NAMES:
        scale = np.abs(ref[name]).max()
        err = float(np.abs(g[name] - ref[name]).max())
        assert err <= 1e-3 * scale, (name, err, scale)


def test_fit_at_1024_on_the_100k_vertex_icosphere(gpu):
    """L2 fit of colours, roughness/metallic, normals and the four intensities to a
    target rendered from hidden parameters: 40 Adam steps must cut the loss by
    more than half."""
    import torch
    print('case: dshade fit 1024x1024 icosphere(100)', flush=True)
    d, ids, case, p, tp = _setup(res=1024, freq=100, seed=3)
    assert d.nverts == 100_002 and d.nfaces == 200_000
    _, _, _, _, _, _, q, _, _ = diff_case(freq=100, res=1024, seed=4)
    hidden = {k: torch.from_numpy(v).cuda().contiguous() for k, v in q.items()}
    with torch.no_grad():
        target = d.forward_raw(hidden['colors'], hidden['normals'], hidden['rough'], hidden['metal'],
                               hidden['intens']).clone()
    leaves = {k: v.clone().requires_grad_(True) for k, v in tp.items()}
    opt = torch.optim.Adam(list(leaves.values()), lr=0.02)
    losses = []
    for it in range(40):
        opt.zero_grad()
        img = d.render(leaves['colors'], leaves['normals'], leaves['rough'], leaves['metal'], leaves['intens'])
        loss = 0.5 * ((img - target) ** 2).sum()
        loss.backward()
        opt.step()
        with torch.no_grad():
            leaves['rough'].clamp_(0.05, 1.0)
            leaves['metal'].clamp_(0.0, 1.0)
            leaves['colors'].clamp_(0.0, 1.0)
        losses.append(float(loss))
    print(f'loss {losses[0]:.4g} -> {losses[-1]:.4g}', flush=True)
    assert np.isfinite(losses).all()
    assert losses[-1] < 0.5 * losses[0]
