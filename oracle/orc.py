"""ctypes wrapper of the CPU oracle (liborc.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, always as the reference of a comparison.  The product package never
imports it.  The restatement it wraps is described in tina_oracle.h; what pins
each function is listed in oracle/README.md.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, 'liborc.so')

_lib = None


def build(force=False):
    if force or not os.path.exists(LIB):
        subprocess.check_call(['make', '-s', '-C', HERE, 'liborc.so'] + (['-B'] if force else []))


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        _lib = ctypes.CDLL(LIB)
    return _lib


def use_native():
    """Rebuilds the oracle -O3 -march=native for this host (liborc_native.so)
    and routes every later call through it: the CPU baseline's build."""
    global _lib
    subprocess.check_call(['make', '-s', '-B', '-C', HERE, 'liborc_native.so'])
    _lib = ctypes.CDLL(os.path.join(HERE, 'liborc_native.so'))
    return _lib


def _p(a):
    return ctypes.c_void_p(0 if a is None else a.ctypes.data)


def camera(W2V, V2W, bias=(0.5, 0.5)):
    from taichi_three_amd import _abi
    cam = _abi.Camera()
    cam.W2V[:] = [float(x) for x in np.asarray(W2V, dtype=np.float32).reshape(-1)]
    cam.V2W[:] = [float(x) for x in np.asarray(V2W, dtype=np.float32).reshape(-1)]
    cam.bias[:] = [float(bias[0]), float(bias[1])]
    return cam


def raster(W, H, cam, verts, id_base=0, flags=3, depth=None, occup=None):
    """Serial coverage (orc_raster).  Returns (depth, occup) [H, W] i32."""
    verts = np.ascontiguousarray(verts, dtype=np.float32)
    if depth is None:
        depth = np.empty((H, W), dtype=np.int32)
        occup = np.empty((H, W), dtype=np.int32)
        lib().orc_clear(W, H, _p(depth), _p(occup))
    lib().orc_raster(W, H, ctypes.byref(cam), _p(verts), len(verts), ctypes.c_uint32(id_base),
                     ctypes.c_uint32(flags), _p(depth), _p(occup))
    return depth, occup


def raster_particles(W, H, cam, pos, radius, id_base=0, flags=2, depth=None, occup=None):
    """Serial particle coverage (orc_raster_particles).  Returns (depth, occup) [H, W] i32."""
    pos = np.ascontiguousarray(pos, dtype=np.float32).reshape(-1, 3)
    radius = np.ascontiguousarray(np.broadcast_to(np.asarray(radius, dtype=np.float32), (len(pos),)))
    if depth is None:
        depth = np.empty((H, W), dtype=np.int32)
        occup = np.empty((H, W), dtype=np.int32)
        lib().orc_clear(W, H, _p(depth), _p(occup))
    lib().orc_raster_particles(W, H, ctypes.byref(cam), _p(pos), _p(radius), len(pos), ctypes.c_uint32(id_base),
                               ctypes.c_uint32(flags), _p(depth), _p(occup))
    return depth, occup


def raster_wires(W, H, cam, verts, id_base=0, flags=0, depth=None, occup=None):
    """Serial wire depth pass (orc_raster_wires); verts [n, 2, 3].  Returns (depth, occup)."""
    verts = np.ascontiguousarray(verts, dtype=np.float32).reshape(-1, 2, 3)
    if depth is None:
        depth = np.empty((H, W), dtype=np.int32)
        occup = np.empty((H, W), dtype=np.int32)
        lib().orc_clear(W, H, _p(depth), _p(occup))
    lib().orc_raster_wires(W, H, ctypes.byref(cam), _p(verts), len(verts), ctypes.c_uint32(id_base),
                           ctypes.c_uint32(flags), _p(depth), _p(occup))
    return depth, occup


def particle_mesh(pos, radius, colors=None, id_base=0):
    """HostMesh of a particle set for shade(..., flags | PARTICLES): centres in
    verts, radii in norms, colours in colors."""
    pos = np.ascontiguousarray(pos, dtype=np.float32).reshape(-1, 3)
    radius = np.broadcast_to(np.asarray(radius, dtype=np.float32), (len(pos),))
    return HostMesh(pos, norms=radius, colors=colors, id_base=id_base)


def raster_omp(W, H, cam, verts, id_base=0, flags=3, keys=None, nthreads=1):
    verts = np.ascontiguousarray(verts, dtype=np.float32)
    if keys is None:
        keys = np.empty((H, W), dtype=np.uint64)
        lib().orc_clear_keys(W, H, _p(keys), nthreads)
    lib().orc_raster_omp(W, H, ctypes.byref(cam), _p(verts), len(verts), ctypes.c_uint32(id_base),
                         ctypes.c_uint32(flags), _p(keys), nthreads)
    return keys


def unpack(keys):
    H, W = keys.shape
    depth = np.empty((H, W), dtype=np.int32)
    occup = np.empty((H, W), dtype=np.int32)
    lib().orc_unpack_keys(W, H, _p(keys), _p(depth), _p(occup))
    return depth, occup


class HostMesh:
    """tina_mesh over host arrays (kept alive by this object)."""

    def __init__(self, verts, norms=None, coors=None, colors=None, id_base=0):
        from taichi_three_amd import _abi
        self.arrays = [None if a is None else np.ascontiguousarray(a, dtype=np.float32)
                       for a in (verts, norms, coors, colors)]
        m = _abi.Mesh()
        m.verts, m.norms, m.coors, m.colors = [0 if a is None else a.ctypes.data for a in self.arrays]
        m.nfaces = len(verts)
        m.id_base = id_base
        self.struct = m


def shade(W, H, cam, mesh, material, lights, occup, flags, bg=(0, 0, 0), ibl=None, f64=False, nthreads=8):
    """Deferred shading of the pixels of `mesh` (orc_shade_f32 / f64).  Returns
    the image [H, W, 3] (pixels outside the mesh's id range are 0 unless BG)."""
    occup = np.ascontiguousarray(occup, dtype=np.int32)
    if f64:
        img = np.zeros((H, W, 3), dtype=np.float64)
        bga = (ctypes.c_double * 3)(*[float(x) for x in np.broadcast_to(bg, 3)])
        fn = lib().orc_shade_f64
    else:
        img = np.zeros((H, W, 3), dtype=np.float32)
        bga = (ctypes.c_float * 3)(*[float(x) for x in np.broadcast_to(bg, 3)])
        fn = lib().orc_shade_f32
    fn(W, H, ctypes.byref(cam), ctypes.byref(mesh.struct), ctypes.byref(material), ctypes.byref(lights),
       ctypes.byref(ibl) if ibl is not None else None, _p(occup), _p(img), bga, ctypes.c_uint32(flags), nthreads)
    return img


def shade_point(material, lights, pos, nrm, viewdir, flags=0, ibl=None, slots=None, f64=True):
    """Shade one point (no tone mapping): Lighting.shade_color(material, pos, nrm, viewdir)."""
    if f64:
        T, fn, ct = np.float64, lib().orc_shade_point_f64, ctypes.c_double
    else:
        T, fn, ct = np.float32, lib().orc_shade_point_f32, ctypes.c_float
    a = [np.ascontiguousarray(x, dtype=T) for x in (pos, nrm, viewdir)]
    s = None if slots is None else np.ascontiguousarray(slots, dtype=T)
    out = np.zeros(3, dtype=T)
    fn(ctypes.byref(material), ctypes.byref(lights), ctypes.byref(ibl) if ibl is not None else None,
       ctypes.c_uint32(flags), _p(a[0]), _p(a[1]), _p(a[2]), _p(s), _p(out))
    return out


def aces(img, f64=False):
    img = np.array(img, dtype=np.float64 if f64 else np.float32, order='C')
    n = img.size // 3
    (lib().orc_aces_f64 if f64 else lib().orc_aces_f32)(_p(img), n)
    return img


def skybox_sample(img, d):
    img = np.ascontiguousarray(img, dtype=np.float32)
    out = np.zeros(3)
    lib().orc_skybox_sample_f64(_p(img), img.shape[0], img.shape[1], _p(np.asarray(d, dtype=np.float64)), _p(out))
    return out


def bake_lambert(env, ow, oh, nsamples, nthreads=8):
    env = np.ascontiguousarray(env, dtype=np.float32)
    out = np.zeros((ow, oh, 3), dtype=np.float32)
    lib().orc_bake_lambert(_p(env), env.shape[0], env.shape[1], _p(out), ow, oh, nsamples, nthreads)
    return out


def bake_ggx(env, ow, oh, roughness, nsamples, nthreads=8):
    env = np.ascontiguousarray(env, dtype=np.float32)
    out = np.zeros((ow, oh, 3), dtype=np.float32)
    lib().orc_bake_ggx(_p(env), env.shape[0], env.shape[1], _p(out), ow, oh, ctypes.c_float(roughness),
                       nsamples, nthreads)
    return out


def mesh_transform(trans, trans_normal, verts, norms=None):
    """MeshTransform on points [..., 3] (and normals [..., 3]): (verts', norms')."""
    t = np.ascontiguousarray(trans, dtype=np.float32).reshape(16)
    tn = np.ascontiguousarray(trans_normal, dtype=np.float32).reshape(9)
    v = np.ascontiguousarray(verts, dtype=np.float32)
    ov = np.empty_like(v)
    n = None if norms is None else np.ascontiguousarray(norms, dtype=np.float32)
    on = None if n is None else np.empty_like(n)
    lib().orc_mesh_transform(_p(t), _p(tn), _p(v), _p(ov), ctypes.c_long(v.size // 3), _p(n), _p(on),
                             ctypes.c_long(0 if n is None else n.size // 3))
    return ov, on


def face_normals(soup):
    """normalized((b - a) x (c - a)) of every face of soup [nf, 3, 3]: [nf, 3]."""
    v = np.ascontiguousarray(soup, dtype=np.float32)
    out = np.empty((len(v), 3), dtype=np.float32)
    lib().orc_face_normals(_p(v), ctypes.c_long(len(v)), _p(out))
    return out


def smooth_normals(soup, faces, nverts):
    """Vertex normals of MeshSmoothNormal: faces [nf, 3] vertex indices, positions
    from the soup [nf, 3, 3] (faces visited in ascending order): [nverts, 3]."""
    v = np.ascontiguousarray(soup, dtype=np.float32)
    f = np.ascontiguousarray(faces, dtype=np.int32)
    out = np.empty((nverts, 3), dtype=np.float32)
    lib().orc_smooth_normals(_p(v), _p(f), ctypes.c_long(len(f)), int(nverts), _p(out))
    return out


def ssr(W, H, cam, mesh, material, depth, occup, norm, image, coor=None, nsamples=32, nsteps=32, stepsize=2.0,
        tolerance=15.0, blur=4, out=None):
    """SSR.render_at (orc_ssr) for the pixels of `mesh` (a HostMesh): returns /
    fills out [H, W, 4] (rgb sum, hits) / nsamples; other pixels left as they are."""
    depth = np.ascontiguousarray(depth, dtype=np.int32)
    occup = np.ascontiguousarray(occup, dtype=np.int32)
    norm = np.ascontiguousarray(norm, dtype=np.float32)
    image = np.ascontiguousarray(image, dtype=np.float32)
    coor = None if coor is None else np.ascontiguousarray(coor, dtype=np.float32)
    if out is None:
        out = np.zeros((H, W, 4), dtype=np.float32)
    lib().orc_ssr(W, H, ctypes.byref(cam), ctypes.byref(mesh.struct), ctypes.byref(material), _p(depth), _p(occup),
                  _p(norm), _p(coor), _p(image), int(nsamples), int(nsteps), ctypes.c_float(stepsize),
                  ctypes.c_float(tolerance), int(blur), _p(out))
    return out


def ssr_apply(image, buf, blur=4):
    """SSR.apply (orc_ssr_apply): image * (1 - a) + rgb of the blur x blur box; returns a new image."""
    img = np.array(image, dtype=np.float32, order='C')
    buf = np.ascontiguousarray(buf, dtype=np.float32)
    H, W = img.shape[:2]
    lib().orc_ssr_apply(W, H, _p(img), _p(buf), int(blur))
    return img
