"""Benchmark of the rasterise-and-shade path on one MI355X.

Workload (one step = one frame): the 1920x1080 frame of a 2,000,000-face soup
(centres U[-1,1]^3, edge lognormal(ln 0.01, 0.7), random orientation), PBR with
per-face roughness/metallic and per-vertex albedo, four point lights, culling
and clipping on -- clear, coverage, deferred shade and ACES, with the geometry
already resident in HBM.  Prints one JSON line.

  python bench.py [--steps K] [--warmup W] [--gpus 1] [--no-cpu]
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = 'fps & Mtri/s at 1080p/2M tris; neural vs analytic shade ms @ PSNR≥40dB'
HBM_PEAK_GBS = 8000.0
BF16_PEAK_TFLOPS = 2500.0                         # dense bf16 MFMA peak of one MI355X
MLP_FLOP_PER_ROW = 2 * (12 * 64 + 64 * 64 * 2 + 64 * 4)   # the 12-64-64-64-4 network, one row


def cpu_info():
    model = 'unknown'
    try:
        for line in open('/proc/cpuinfo'):
            if line.startswith('model name'):
                model = line.split(':', 1)[1].strip()
                break
    except OSError:
        pass
    return model


def cpu_baseline(w, budget_s=12.0):
    """The oracle's OpenMP port of the reference algorithm (one task per face with a
    serial bbox scan, then a per-pixel deferred shade), built -O3 -march=native on
    this host, on every core of the process's affinity mask."""
    import numpy as np
    from oracle import orc
    from taichi_three_amd import workloads as wl, _abi
    from taichi_three_amd.matr import lower_material
    orc.use_native()
    nthreads = len(os.sched_getaffinity(0))
    W, H = w.res
    cam = orc.camera(w.W2V, w.V2W)
    lo = lower_material(wl.make_material(w.material), dict(attrs=w.attrs))
    lt = wl.make_lighting(w).to_struct()
    mesh = orc.HostMesh(w.verts, w.norms, None, w.colors)
    flags = _abi.VCOLOR | _abi.BG | _abi.TONEMAP
    times = []
    t_start = time.perf_counter()
    while True:
        t0 = time.perf_counter()
        keys = orc.raster_omp(W, H, cam, w.verts, flags=3, nthreads=nthreads)
        _, ids = orc.unpack(keys)
        orc.shade(W, H, cam, mesh, lo.struct, lt, ids, flags, nthreads=nthreads)
        times.append(time.perf_counter() - t0)
        if time.perf_counter() - t_start > budget_s or len(times) >= 50:
            break
    med = float(np.median(times))
    return {'value': round(1.0 / med, 3), 'unit': 'fps', 'cores': nthreads, 'kind': 'port',
            'sample': f'{len(times)} full frames of the same 1080p/2M workload (median), oracle built '
                      f'-O3 -march=native -ffp-contract=off, {nthreads} OpenMP threads, CPU: {cpu_info()}',
            'ms_per_frame': round(med * 1e3, 2)}


def stage_times(scene, reps):
    """Mean device ms per render phase, from events recorded on the stream the
    kernels are launched on (torch's current stream)."""
    import numpy as np
    import torch
    stages, evs = {}, []

    def on_phase(name):
        e = torch.cuda.Event(enable_timing=True)
        e.record()
        evs.append((name, e))

    for _ in range(reps):
        evs.clear()
        scene.render(on_phase=on_phase)
        torch.cuda.synchronize()
        for (n0, e0), (_, e1) in zip(evs[:-1], evs[1:]):
            stages.setdefault(n0, []).append(e0.elapsed_time(e1))
    return {k: float(np.median(v)) for k, v in stages.items()}


def traffic_for(stage):
    """HBM bytes per launch of a stage's kernels, from the committed counter
    summary profiles/traffic_soup_1080p.json (rocprofv3 --pmc FETCH_SIZE and
    WRITE_SIZE, scaled by what the counters read for a copy of known size), or
    None when that summary is absent."""
    try:
        with open(os.path.join(ROOT, 'profiles', 'traffic_soup_1080p.json')) as f:
            return json.load(f)['stages'][stage]['hbm_bytes']
    except (OSError, KeyError, ValueError):
        return None


def psnr(a, b):
    import numpy as np
    mse = float(np.mean((a.astype(np.float64) - b.astype(np.float64)) ** 2))
    return 10.0 * np.log10(1.0 / max(mse, 1e-30))


def eval_only(ns, n, reps):
    """The two evaluations alone, without the per-pixel fetch of normals,
    material and view that the shade pass shares: the network and the analytic
    Cook-Torrance + IBL shader on the same n feature rows (drawn like the
    training pairs, seed 7), device ms per call and the PSNR between them."""
    import numpy as np
    import torch
    X, _ = ns.samples(n, seed=7)

    def timed(fn, calls=10):  # median over reps of the mean of `calls` back-to-back calls
        fn(X)
        torch.cuda.synchronize()
        ts = []
        for _ in range(reps):
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            for _ in range(calls):
                fn(X)
            e1.record()
            torch.cuda.synchronize()
            ts.append(e0.elapsed_time(e1) / calls)
        return float(np.median(ts))

    t_nn, t_an = timed(ns.infer), timed(ns.analytic)
    o_nn = ns.infer(X)[:, :3].cpu().numpy()
    o_an = ns.analytic(X)[:, :3].cpu().numpy()
    flop = n * MLP_FLOP_PER_ROW
    # the floor of one launch in this timing (back-to-back calls of a one-element fill)
    z = torch.zeros(1, device='cuda')
    t_launch = timed(lambda _x: z.fill_(1.0))
    # four times the rows: the per-row rates once the fixed cost of a launch is amortised
    X4 = torch.cat([X] * 4)
    t_nn4 = timed(lambda _x: ns.infer(X4))
    t_an4 = timed(lambda _x: ns.analytic(X4))
    del X4
    return {'rows': n, 'analytic_ms': round(t_an, 4), 'mlp_ms': round(t_nn, 4),
            'speedup': round(t_an / t_nn, 3), 'psnr_db': round(psnr(o_nn, o_an), 2),
            'mlp_tflops': round(flop / (t_nn * 1e-3) / 1e12, 2),
            'roofline': {'bound': 'mfma', 'kernel': 'k_mlp_rows', 'achieved': round(flop / (t_nn * 1e-3) / 1e12, 1),
                         'peak': BF16_PEAK_TFLOPS, 'unit': 'TFLOP/s',
                         'frac': round(flop / (t_nn * 1e-3) / 1e12 / BF16_PEAK_TFLOPS, 4), 'traffic': None,
                         'flop_per_row': MLP_FLOP_PER_ROW,
                         'mfma_floor_ms': round(flop / (BF16_PEAK_TFLOPS * 1e12) * 1e3, 4),
                         'launch_floor_ms': round(t_launch, 4)},
            'rows_x4': {'rows': 4 * n, 'analytic_ms': round(t_an4, 4), 'mlp_ms': round(t_nn4, 4),
                        'speedup': round(t_an4 / t_nn4, 3),
                        'mlp_tflops': round(4 * flop / (t_nn4 * 1e-3) / 1e12, 1)}}


def neural_vs_analytic(reps, train_steps):
    """The learned surrogate against the analytic Cook-Torrance + IBL shade at
    1920x1080 with 1M faces: train on 16M analytic pairs, then time the shade
    phase both ways and compare the frames."""
    import torch
    from taichi_three_amd import workloads as wl
    from taichi_three_amd.neural import NeuralShading
    w = wl.ibl_1080p()
    scene, mesh = wl.build_scene(w)
    scene.render()
    torch.cuda.synchronize()
    ref = scene.img.to_numpy().copy()
    analytic = stage_times(scene, reps)['shade']
    ns = NeuralShading(scene.lighting, seed=1)
    t0 = time.perf_counter()
    ns.train(nsamples=1 << 24, steps=train_steps, batch=16384)
    torch.cuda.synchronize()
    t_train = time.perf_counter() - t0
    scene.neural = ns
    scene.render()
    torch.cuda.synchronize()
    img = scene.img.to_numpy()
    nfall = ns.fallback_count(scene.engine)
    neural = stage_times(scene, reps)['shade']
    ids = scene.engine.ids.to_numpy()
    covered = int((ids >= 0).sum())
    # the same trained network at other thresholds: fewer analytic pixels, lower PSNR
    # (inf: no fallback at all)
    sweep = {}
    for th in (0.03, float('inf')):
        ns.threshold = th
        scene.render()
        torch.cuda.synchronize()
        sweep[str(th)] = {'fallback_frac': round(ns.fallback_count(scene.engine) / max(1, covered), 4),
                          'psnr_db': round(psnr(scene.img.to_numpy(), ref), 2),
                          'neural_shade_ms': round(stage_times(scene, reps)['shade'], 4)}
    ns.threshold = 1e-2
    rows = eval_only(ns, covered, reps)
    return {'workload': 'ibl_1080p: 1920x1080, 1M-face soup, PBR under a 512x256 procedural HDR sky + 1 sun',
            'analytic_shade_ms': round(analytic, 4), 'neural_shade_ms': round(neural, 4),
            'speedup': round(analytic / neural, 3), 'psnr_db': round(psnr(img, ref), 2),
            'eval_only': rows,
            'fallback_pixels': nfall, 'covered_pixels': covered,
            'fallback_frac': round(nfall / max(1, covered), 4), 'threshold': ns.threshold,
            'other_thresholds': sweep,
            'train': {'samples': 1 << 24, 'steps': train_steps, 'batch': 16384, 'seconds': round(t_train, 2)},
            'mlp_flop_per_px': MLP_FLOP_PER_ROW}


def post_passes(reps):
    """Blooming and FXAA on the 1920x1080 image (device ms per pass, median)
    with the algorithmic bytes each moves: bloom reads the image and adds to it
    (3 x 12 B/px) plus its half-resolution passes (2 x 2 x 12 B per half pixel);
    FXAA reads the image, writes luminance and a copy, reads them back and
    writes the image (12 + 4 + 12 + 4 + 12 + 12 B/px)."""
    import numpy as np
    import torch
    import taichi_three_amd as tina
    from taichi_three_amd.field import Field
    W, H = 1920, 1080
    img = Field((torch.rand((H, W, 3), device='cuda') * 3).contiguous())
    out = {}
    for name, p, nbytes in (('bloom', tina.Blooming((W, H)), 36 * W * H + 48 * (W // 2) * (H // 2)),
                            ('fxaa', tina.FXAA((W, H)), 56 * W * H)):
        p.apply(img)
        torch.cuda.synchronize()
        ts = []
        for _ in range(reps):
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            p.apply(img)
            e1.record()
            torch.cuda.synchronize()
            ts.append(e0.elapsed_time(e1))
        ms = float(np.median(ts))
        out[name] = {'ms': round(ms, 4), 'algorithmic_GBs': round(nbytes / (ms * 1e-3) / 1e9, 1)}
    out['bloom']['radius'] = min(W, H) // 16
    return out


def other_rasters(reps):
    """Particle and wire passes beside the triangle frame (device ms, median):
    1 M particles of radius 0.003-0.008 at 1920x1080 and 64 k of radius
    0.01-0.02 at 512x512 (the particle example's size), uniform in [-1, 1]^3,
    Diffuse, camera at distance 3.2 with a 45 degree field of view; 200 k
    random wires in [-1.2, 1.2]^3 at 1920x1080.  'raster' is the coverage
    pass (with the key clear), 'shade' the colour pass."""
    import numpy as np
    import torch
    import taichi_three_amd as tina
    from taichi_three_amd.util.matrix import lookat, perspective
    out = {}

    def timed(scene, tag):
        ev = []

        def mark(name):
            e = torch.cuda.Event(enable_timing=True)
            e.record()
            ev.append((name, e))
        for _ in range(3):
            scene.render()
        torch.cuda.synchronize()
        per = {}
        for _ in range(reps):
            ev.clear()
            scene.render(on_phase=mark)
            torch.cuda.synchronize()
            for (a, ea), (_, eb) in zip(ev, ev[1:]):
                per.setdefault(a, []).append(ea.elapsed_time(eb))
        out[tag] = {k: round(float(np.median(v)), 4) for k, v in per.items() if k in ('raster', 'shade')}

    rng = np.random.default_rng(0)
    for tag, (W, H, n, rlo, rhi) in (('particles_1080p_1M', (1920, 1080, 1 << 20, 0.003, 0.008)),
                                     ('particles_512_64k', (512, 512, 65536, 0.01, 0.02))):
        scene = tina.Scene(W, H, maxpars=n)
        pars = tina.SimpleParticles(n)
        pars.set_particles(rng.uniform(-1, 1, (n, 3)).astype(np.float32))
        pars.set_particle_radii(rng.uniform(rlo, rhi, n).astype(np.float32))
        scene.add_object(pars, tina.Diffuse())
        scene.set_camera(lookat([0, 0, 0], [0.4, 0.3, 3.2]), perspective(45, W / H, 0.1, 50))
        timed(scene, tag)
        out[tag]['particles_per_s'] = round(n / (out[tag]['raster'] + out[tag]['shade']) * 1e3, 0)
    W, H, n = 1920, 1080, 200_000
    scene = tina.Scene(W, H)
    mesh = tina.SimpleMesh(n, npolygon=2)
    mesh.set_face_verts(rng.uniform(-1.2, 1.2, (n, 2, 3)).astype(np.float32))
    scene.add_object(mesh)
    scene.set_camera(lookat([0, 0, 0], [0.4, 0.3, 3.2]), perspective(45, W / H, 0.1, 50))
    timed(scene, 'wires_1080p_200k')
    return out


def batched_views(reps, tune_reps=3):
    """64 independent scenes (1k-500k faces log-uniform, resolutions from
    {256^2, 512^2, 1024^2, 1920x1080}, cameras on the radius-3 sphere, seed 2;
    the 512x512 workload's triangle and Blinn-Phong material generators)
    rendered as one batch: autotuned bin sizes and grids, then timed graph
    replays; beside it the same scenes rendered one Scene.render() at a time."""
    import torch
    from taichi_three_amd import autotune, workloads as wl
    from taichi_three_amd.batch import BatchRenderer, BatchScenes
    ws = wl.batch_views()
    scenes = BatchScenes(ws)
    t0 = time.perf_counter()
    best, results = autotune.tune(scenes, reps=tune_reps)
    t_tune = time.perf_counter() - t0
    plan = BatchRenderer(scenes, best)
    ms = autotune.time_plan(plan, reps=reps)
    # algorithmic bytes of one batch: per scene the positions of every face, the
    # key clear, the shade pass's key read and the RGB write, and per visible face
    # its positions, normals and albedo (36 + 36 + 36 B)
    nbytes = 0
    for i, (W, H) in enumerate(scenes.res):
        ids, _ = plan.resolve(i)
        V = int(torch.unique(ids[ids >= 0]).numel())
        nbytes += 36 * scenes.nfaces[i] + (8 + 8 + 12) * W * H + 108 * V
    plan.close()
    # the same scenes, one Scene at a time (each its own engine), device time per frame
    serial = 0.0
    for w in ws:
        scene, mesh = wl.build_scene(w)
        for _ in range(2):
            scene.render()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(3):
            scene.render()
        e1.record()
        torch.cuda.synchronize()
        serial += e0.elapsed_time(e1) / 3
        del scene, mesh
    torch.cuda.empty_cache()
    return {'workload': 'batch_views: 64 scenes, 1k-500k faces (log-uniform), {256^2,512^2,1024^2,1920x1080}, '
                        'cameras on the radius-3 sphere, Blinn-Phong over per-vertex albedo, seed 2',
            'scenes': scenes.n, 'faces': scenes.total_faces, 'pixels': scenes.total_pixels,
            'batch_ms': round(ms, 4), 'serial_scene_render_ms': round(serial, 4),
            'speedup_vs_serial': round(serial / ms, 3),
            'scenes_per_s': round(scenes.n / (ms * 1e-3), 1),
            'mtri_per_s': round(scenes.total_faces / (ms * 1e-3) / 1e6, 1),
            'mpix_per_s': round(scenes.total_pixels / (ms * 1e-3) / 1e6, 1),
            'roofline': {'bound': 'hbm', 'achieved': round(nbytes / (ms * 1e-3) / 1e9, 1), 'peak': HBM_PEAK_GBS,
                         'unit': 'GB/s', 'frac': round(nbytes / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                         'bytes': nbytes},
            'tuning': best.as_dict(), 'tuned_candidates': len(results), 'tune_seconds': round(t_tune, 1)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=50)
    ap.add_argument('--warmup', type=int, default=10)
    ap.add_argument('--no-cpu', action='store_true')
    ap.add_argument('--no-neural', action='store_true')
    ap.add_argument('--no-batch', action='store_true')
    ap.add_argument("--train-steps", type=int, default=60000)
    ap.add_argument('--nfaces', type=int, default=2_000_000)
    ap.add_argument('--copy-mb', type=int, default=0,
                    help='first run three kernels of known traffic on this many MiB (a device copy, '
                         'a sum that only reads, a gather of 64-byte rows in random order) to '
                         'calibrate the HBM byte counters')
    ap.add_argument('--frame-only', action='store_true',
                    help='time the 1080p frame alone (no surrogate, batch, side passes, SSAO or CPU '
                         'baseline): the short run the counter profiles use')
    args = ap.parse_args()
    if args.gpus != 1:
        print(f'bench.py measures one GPU only (asked for --gpus {args.gpus}): the path runs on a '
              'single GPU; sharding over several GPUs is a design, not something this script runs.',
              file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch
    from taichi_three_amd import workloads as wl

    if args.copy_mb > 0:
        n = args.copy_mb << 18
        src = torch.ones(n, dtype=torch.float32, device='cuda')
        dst = torch.empty_like(src)
        dst.copy_(src)                       # reads and writes copy_mb MiB
        tot = src.sum()                      # reads copy_mb MiB
        rows = src.view(-1, 16)              # 64-byte rows
        perm = torch.randperm(rows.shape[0], device='cuda')
        torch.cuda.synchronize()
        dst.view(-1, 16).copy_(rows.index_select(0, perm))   # a gather of every row in random order
        torch.cuda.synchronize()
        del src, dst, tot, rows, perm

    w = wl.soup_1080p(nfaces=args.nfaces)
    W, H = w.res
    T = len(w.verts)
    scene, mesh = wl.build_scene(w)
    torch.cuda.synchronize()

    for _ in range(args.warmup):
        scene.render()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        scene.render()
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    ms = dt / args.steps * 1e3
    fps = args.steps / dt

    # per-stage device time with events on the stream the kernels run on
    reps = max(10, args.steps)
    stage_ms = stage_times(scene, reps)

    ids = scene.engine.ids.to_numpy()
    V = int(len(np.unique(ids[ids >= 0])))
    covered = int((ids >= 0).sum())
    npix = W * H
    # algorithmic bytes per launch: compulsory reads and writes of each stage
    # (the 'clear' phase only records the clear: k_tile applies it, writing every
    # key once, so the key write belongs to the raster stage)
    bytes_stage = {
        'clear': 0,
        'raster': 36 * T + 8 * npix,                    # read every face's positions, write the packed keys
        'shade': 8 * npix + 80 * V + 12 * npix,         # keys, visible faces' pos+albedo+rough/metal, RGB
    }
    dom = max(stage_ms, key=stage_ms.get)
    achieved = bytes_stage[dom] / (stage_ms[dom] * 1e-3) / 1e9
    frame_bytes = sum(bytes_stage.values())
    out = {
        'metric': METRIC,
        'value': round(fps, 3),
        'unit': 'fps',
        'n_gpus': 1,
        'steps': args.steps,
        'warmup': args.warmup,
        'ms_per_step': round(ms, 4),
        'higher_is_better': True,
        'scaling': 'weak',
        'vs_baseline': None,
        'dtype': 'f32',
        'data': 'synthetic (seeded generator, geometry resident in HBM)',
        'config': {'workload': 'soup_1080p: 1920x1080, 2M-face soup, PBR per-face roughness/metallic, 4 point lights',
                   'width': W, 'height': H, 'faces': T, 'culling': True, 'clipping': True},
        'mtri_per_s': round(T * fps / 1e6, 2),
        'mpix_per_s': round(npix * fps / 1e6, 2),
        'stage_ms': {k: round(v, 4) for k, v in stage_ms.items()},
        'visible_faces': V,
        'covered_pixels': covered,
        'frame_roofline': {'bytes': frame_bytes, 'achieved_GBs': round(frame_bytes / (ms * 1e-3) / 1e9, 1),
                           'frac': round(frame_bytes / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)},
        'roofline': {'bound': 'hbm', 'kernel': {'clear': '(none)', 'raster': 'k_bin+k_tile',
                                                'shade': 'k_shade'}[dom],
                     'achieved': round(achieved, 1), 'peak': HBM_PEAK_GBS, 'unit': 'GB/s',
                     'frac': round(achieved / HBM_PEAK_GBS, 4), 'traffic': traffic_for(dom)},
    }
    del scene, mesh
    torch.cuda.empty_cache()
    if args.frame_only:
        print(json.dumps(out))
        return
    if not args.no_neural:
        out['neural'] = neural_vs_analytic(max(10, args.steps // 2), args.train_steps)
    if not args.no_batch:
        out['batch'] = batched_views(max(5, args.steps // 4))
    out['postp'] = post_passes(max(10, args.steps // 2))
    out['other_rasters'] = other_rasters(max(10, args.steps // 2))
    # SSAO on the same frame: the shade stage then also writes the normal buffer,
    # runs the 64-sample occlusion, the 4x4 box and ACES as a separate pass
    scene_ao, _ = wl.build_scene(w, ssao=True)
    ao_ms = stage_times(scene_ao, max(10, args.steps // 2))
    out['ssao'] = {'shade_ms': round(ao_ms['shade'], 4), 'extra_ms': round(ao_ms['shade'] - stage_ms['shade'], 4),
                   'nsamples': scene_ao.ssao.nsamples}
    del scene_ao
    if not args.no_cpu:
        out['cpu_baseline'] = cpu_baseline(w)
        out['speedup_vs_cpu'] = round(fps / out['cpu_baseline']['value'], 1)
    print(json.dumps(out))


if __name__ == '__main__':
    main()
