"""Mesh nodes: the input side of TriangleRaster.set_object (tina/mesh/*.py).

The reference's node protocol is a set of ti.func getters that set_object
inlines and copies into the raster's own [maxfaces, 3] fields every frame
(tina/core/triangle.py:71-86).  Here a mesh keeps its face data resident in HBM
as [T, 3, 3] f32 tensors (the same AoS layout) and `face_buffers()` hands the
raster those tensors directly, so no per-frame copy happens.

Extensions beyond the reference (needed by the benchmark workloads):
  set_face_colors(c[T,3,3])   per-vertex albedo, bound to Input('color')
  set_face_attr(name, a[T] or a[T,3])   per-face arrays read by tina.FaceAttr(name)
"""
import ctypes

import numpy as np

from . import _abi
from .engine import MAX

__all__ = ['SimpleMesh', 'PrimitiveMesh', 'MeshFlatNormal', 'MeshSmoothNormal', 'ConnectiveMesh',
           'MeshTransform', 'MeshEditBase', 'MeshModel', 'MeshToWire', 'MeshFlipCulling', 'MeshNoCulling',
           'MeshFlipNormal', 'MeshGrid', 'export_simple_mesh', 'simple_mesh_to_connective', 'export_connective_mesh']


def _torch():
    import torch
    return torch


def _as_device(arr, shape_tail, device=None):
    torch = _torch()
    dev = device or torch.device('cuda')
    if isinstance(arr, torch.Tensor):
        t = arr.to(device=dev, dtype=torch.float32)
    else:
        t = torch.as_tensor(np.ascontiguousarray(np.asarray(arr, dtype=np.float32)), device=dev)
    if tuple(t.shape[1:]) != tuple(shape_tail):
        raise ValueError(f'expected an array of shape [n, {", ".join(map(str, shape_tail))}], got {tuple(t.shape)}')
    return t.contiguous()


class SimpleMesh:
    """Face-soup mesh (tina/mesh/simple.py:4-115)."""

    def __init__(self, maxfaces=MAX, npolygon=3):
        if npolygon not in (2, 3):
            raise NotImplementedError('meshes of triangles (npolygon=3) or wires (npolygon=2) are rasterised here')
        self.maxfaces = int(maxfaces)
        self.npolygon = npolygon
        self.nfaces = 0
        self.verts = None
        self.norms = None
        self.coors = None
        self.colors = None
        self.mtlids = None
        self.attrs = {}
        self.version = 0

    def get_npolygon(self):
        return self.npolygon

    def get_nfaces(self):
        return min(self.nfaces, self.maxfaces)

    def pre_compute(self):
        pass

    def _n(self, arr):
        return min(int(arr.shape[0]), self.maxfaces)

    def set_face_verts(self, verts):
        """verts: [nfaces, 3, 3]; the face count is verts.shape[0] (simple.py:53-67)."""
        n = self._n(verts)
        self.verts = _as_device(verts[:n], (self.npolygon, 3))
        self.nfaces = n
        self.version += 1

    def set_face_norms(self, norms):
        self.norms = _as_device(norms[:self.nfaces], (self.npolygon, 3))
        self.version += 1

    def set_face_coors(self, coors):
        self.coors = _as_device(coors[:self.nfaces], (self.npolygon, 2))
        self.version += 1

    def set_face_mtlids(self, mtlids):
        self.mtlids = np.asarray(mtlids[:self.nfaces], dtype=np.int32)

    def set_material_id(self, mtlid):
        self.mtlids = np.full(self.nfaces, mtlid, dtype=np.int32)

    def set_face_colors(self, colors):
        """Per-vertex colours [nfaces, 3, 3] bound to Input('color') (extension)."""
        self.colors = _as_device(colors[:self.nfaces], (self.npolygon, 3))
        self.version += 1

    def set_face_attr(self, name, values):
        """Per-face scalar [nfaces] or vec3 [nfaces, 3] array read by tina.FaceAttr(name)."""
        torch = _torch()
        v = values if isinstance(values, torch.Tensor) else torch.as_tensor(np.asarray(values, dtype=np.float32))
        v = v.to(device='cuda', dtype=torch.float32)[:self.nfaces].contiguous()
        if v.dim() == 2 and v.shape[1] not in (1, 3):
            raise ValueError('face attributes must be [n] or [n, 3]')
        self.attrs[name] = v
        self.version += 1

    def face_buffers(self):
        return dict(verts=self.verts, norms=self.norms, coors=self.coors, colors=self.colors,
                    attrs=self.attrs, nfaces=self.get_nfaces() if self.verts is not None else 0)


class PrimitiveMesh(SimpleMesh):
    """Procedural meshes (tina/mesh/prim.py:6-111); faces[f][k] = (vert, norm, coor)."""

    def __init__(self, faces):
        faces = np.asarray(faces, dtype=np.float32)
        super().__init__(maxfaces=len(faces), npolygon=faces.shape[1])
        self.set_face_verts(faces[:, :, 0])
        self.set_face_norms(faces[:, :, 1])
        self.set_face_coors(faces[:, :, 2, :2])

    @staticmethod
    def sphere_faces(lons=32, lats=24, rad=1):
        """UV sphere face list (prim.py:21-47), as float64 (F, 3, 3, 3)."""
        def at(lat, lon):
            lat /= lats
            lon /= lons
            coor = lon, lat, 0
            lat = (lat - 0.5) * np.pi
            lon = (lon * 2 - 1) * np.pi
            x = np.cos(lat) * np.cos(lon)
            y = np.cos(lat) * np.sin(lon)
            z = np.sin(lat)
            return (x * rad, y * rad, z * rad), (x, y, z), coor

        faces = []
        for lat in range(lats):
            for lon in range(lons):
                v1 = at(lat, lon)
                v2 = at(lat, lon + 1)
                v3 = at(lat + 1, lon + 1)
                v4 = at(lat + 1, lon)
                if lat != 0:
                    faces.append([v1, v2, v3])
                if lat != lats - 1:
                    faces.append([v3, v4, v1])
        return np.array(faces, dtype=np.float64)

    @classmethod
    def sphere(cls, lons=32, lats=24, rad=1):
        return cls(cls.sphere_faces(lons, lats, rad))

    @staticmethod
    def icosphere(freq=1, rad=1.0):
        """Geodesic icosphere of frequency `freq` (extension: the reference has
        only UV spheres): every icosahedron face split into freq^2 triangles,
        vertices pushed to the sphere and shared.  Returns (verts [10 freq^2 + 2][3],
        faces [20 freq^2][3] int32, counter-clockwise seen from outside), float64."""
        t = (1.0 + 5.0 ** 0.5) / 2.0
        ico_v = np.array([(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t),
                          (0, -1, -t), (0, 1, -t), (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)], np.float64)
        ico_f = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4), (11, 10, 2),
                 (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9), (4, 9, 5),
                 (2, 4, 11), (6, 2, 10), (8, 6, 7), (9, 8, 1)]
        n = int(freq)
        index, verts, faces = {}, [], []

        def vid(p):
            key = tuple(np.round(p / np.linalg.norm(p), 9))
            if key not in index:
                index[key] = len(verts)
                verts.append(p / np.linalg.norm(p))
            return index[key]

        for a, b, c in ico_f:
            A, B, C = ico_v[a], ico_v[b], ico_v[c]
            grid = {}
            for i in range(n + 1):
                for j in range(n + 1 - i):
                    grid[i, j] = vid(A + (B - A) * (i / n) + (C - A) * (j / n))
            for i in range(n):
                for j in range(n - i):
                    faces.append((grid[i, j], grid[i + 1, j], grid[i, j + 1]))
                    if i + j + 1 < n:
                        faces.append((grid[i + 1, j], grid[i + 1, j + 1], grid[i, j + 1]))
        V = np.array(verts) * rad
        F = np.array(faces, dtype=np.int32)
        # orient every face outward
        P = V[F]
        nrm = np.cross(P[:, 1] - P[:, 0], P[:, 2] - P[:, 0])
        flip = (nrm * P.sum(axis=1)).sum(axis=1) < 0
        F[flip] = F[flip][:, ::-1]
        return V, F

    @classmethod
    def cylinder(cls, lons=32, lats=4, rad=1, hei=2):
        """Capped cylinder (prim.py:50-94)."""
        def at(lat, lon, rad=rad):
            lat /= lats
            lon /= lons
            coor = lon, lat, 0
            lat = lat - 0.5
            lon = (lon * 2 - 1) * np.pi
            x = np.cos(lon)
            y = np.sin(lon)
            return (x * rad, y * rad, lat * hei), (x, y, 0), coor

        faces = []
        for lat in range(lats):
            for lon in range(lons):
                v1, v2, v3, v4 = at(lat, lon), at(lat, lon + 1), at(lat + 1, lon + 1), at(lat + 1, lon)
                faces.append([v1, v2, v3])
                faces.append([v3, v4, v1])
        for lon in range(lons):
            norm, coor = (0, 0, -1), (0, 0, -1)
            faces.append([(at(0, lon + 1)[0], norm, coor), (at(0, lon)[0], norm, coor),
                          ((0, 0, -hei / 2), norm, coor)])
        for lon in range(lons):
            norm, coor = (0, 0, 1), (0, 0, 1)
            faces.append([((0, 0, hei / 2), norm, coor), (at(lats, lon)[0], norm, coor),
                          (at(lats, lon + 1)[0], norm, coor)])
        return cls(np.array(faces, dtype=np.float64))


class MeshEditBase:
    """Forwards everything to the wrapped mesh (tina/mesh/base.py)."""

    def __init__(self, mesh):
        self.mesh = mesh

    def __getattr__(self, attr):
        return getattr(self.mesh, attr)

    def face_buffers(self):
        return self.mesh.face_buffers()


class MeshFlatNormal(MeshEditBase):
    """Per-face normals normalized((b - a) x (c - a)) on every corner (tina/mesh/norm.py:5-14),
    computed on the device (tina_face_normals) from the wrapped mesh's positions."""

    def __init__(self, mesh):
        super().__init__(mesh)
        self._out = None

    def face_buffers(self):
        fb = dict(self.mesh.face_buffers())
        v = fb['verts']
        if v is None:
            return fb
        torch = _torch()
        nf = int(fb['nfaces'])
        if self._out is None or self._out.shape[0] < nf or self._out.device != v.device:
            self._out = torch.empty((max(nf, 1), 3, 3), dtype=torch.float32, device=v.device)
        out = self._out[:nf]
        _abi.check(_abi.load().tina_face_normals(_abi.ptr(v), nf, 3, _abi.ptr(out), _abi.stream_handle()))
        fb['norms'] = out
        return fb


class ConnectiveMesh:
    """Indexed mesh: vertex arrays plus faces[nf, 3] indices (tina/mesh/conn.py:4-95,
    tina/mesh/model.py:4-73).  Faces are expanded into a device soup for the raster."""

    def __init__(self, verts=None, faces=None, norms=None, coors=None, colors=None):
        self.npolygon = 3
        self.verts_v = self.faces = self.norms_v = self.coors_v = self.colors_v = None
        self.attrs = {}
        self.version = 0
        if verts is not None:
            self.set_mesh(verts, faces, norms, coors, colors)

    def get_npolygon(self):
        return 3

    def get_nfaces(self):
        return 0 if self.faces is None else int(self.faces.shape[0])

    def get_max_vert_nindex(self):
        return 0 if self.verts_v is None else int(self.verts_v.shape[0])

    def set_mesh(self, verts, faces, norms=None, coors=None, colors=None):
        torch = _torch()
        self.verts_v = _as_device(verts, (3,))
        self.faces = torch.as_tensor(np.asarray(faces, dtype=np.int64), device='cuda')
        self.norms_v = None if norms is None else _as_device(norms, (3,))
        self.coors_v = None if coors is None else _as_device(coors, (2,))
        self.colors_v = None if colors is None else _as_device(colors, (3,))
        self.version += 1

    def set_vertex_colors(self, colors):
        self.colors_v = _as_device(colors, (3,))
        self.version += 1

    def set_face_attr(self, name, values):
        SimpleMesh.set_face_attr(self, name, values)

    @property
    def nfaces(self):
        return self.get_nfaces()

    def face_buffers(self):
        f = self.faces
        g = (lambda a: None if a is None else a[f].contiguous())
        return dict(verts=g(self.verts_v), norms=g(self.norms_v), coors=g(self.coors_v),
                    colors=g(self.colors_v), attrs=self.attrs, nfaces=self.get_nfaces())


class MeshModel:
    """A mesh read from an OBJ file (tina/mesh/model.py:4-73): vertex, texcoord
    and normal arrays with faces[nf, 3, 3] (vertex, texcoord, normal) index
    triples, as tina.readobj returns them.  `obj` is that dictionary or a path
    (further arguments go to readobj).  The faces are expanded once into a
    resident [nf, 3, ...] face soup for the raster, on the device."""

    def __init__(self, obj, *args, **kwargs):
        if isinstance(obj, str):
            from .assimp import readobj
            obj = readobj(obj, *args, **kwargs)
        torch = _torch()
        f = np.asarray(obj['f'])
        if f.ndim == 2:  # one index per corner for every stream
            f = np.stack([f, f, f], axis=2)
        self.faces = torch.as_tensor(np.ascontiguousarray(f, dtype=np.int64), device='cuda')
        self.verts_v = _as_device(obj['v'], (3,))
        self.coors_v = _as_device(obj['vt'], (2,))
        self.norms_v = _as_device(obj['vn'], (3,))
        self.maxfaces = int(f.shape[0])
        self.maxverts = int(self.verts_v.shape[0])
        self.maxcoors = int(self.coors_v.shape[0])
        self.maxnorms = int(self.norms_v.shape[0])
        self.attrs = {}
        self.version = 0
        self._soup = None

    def get_npolygon(self):
        return 3

    def pre_compute(self):
        pass

    def get_max_vert_nindex(self):
        return self.maxverts

    def get_nfaces(self):
        return self.maxfaces

    @property
    def nfaces(self):
        return self.maxfaces

    def get_face_vert_indices(self, n):
        """(i, j, k): the vertex indices of face n."""
        return tuple(int(x) for x in self.faces[n, :, 0].tolist())

    def set_face_attr(self, name, values):
        SimpleMesh.set_face_attr(self, name, values)

    def face_buffers(self):
        if self._soup is None or self._soup[0] != self.version:
            f = self.faces
            self._soup = (self.version, dict(verts=self.verts_v[f[:, :, 0]].contiguous(),
                                             coors=self.coors_v[f[:, :, 1]].contiguous(),
                                             norms=self.norms_v[f[:, :, 2]].contiguous()))
        soup = self._soup[1]
        return dict(verts=soup['verts'], norms=soup['norms'], coors=soup['coors'], colors=None,
                    attrs=self.attrs, nfaces=self.maxfaces)


def _vertex_indices(mesh):
    """The [nf, 3] vertex indices of the indexed mesh under a chain of edit nodes
    (get_face_vert_indices: ConnectiveMesh faces, MeshModel's vertex index column)."""
    m = mesh
    while not hasattr(m, 'faces') and isinstance(m, MeshEditBase):
        m = m.mesh
    f = getattr(m, 'faces', None)
    if f is None or not hasattr(m, 'get_max_vert_nindex'):
        raise NotImplementedError('MeshSmoothNormal needs an indexed mesh (ConnectiveMesh or MeshModel)')
    return m, (f[:, :, 0] if f.dim() == 3 else f)


class MeshSmoothNormal(MeshEditBase):
    """Area-blind vertex normals (tina/mesh/norm.py:17-55): per vertex the sum of
    the normalized face normals of its faces, added in ascending face order (the
    reference's loop run serially), then normalized.  On the device: the
    vertex -> (face, corner) lists are built once per topology
    (tina_vertex_incidence, a stable sort keeps each list in face order), the
    face normals come from the wrapped mesh's current positions
    (tina_face_normals) and each vertex sums its list in order
    (tina_smooth_normals).  cached=True computes them once, as the reference's
    materialize callback does (update_normal() recomputes); cached=False every
    frame."""

    def __init__(self, mesh, cached=True):
        super().__init__(mesh)
        self._base, _ = _vertex_indices(mesh)
        self.cached = cached
        self._norm = None
        self._version = None
        self._csr = None

    def _incidence(self):
        torch = _torch()
        base, idx = _vertex_indices(self.mesh)
        key = (id(base), getattr(base, 'version', 0))
        if self._csr is not None and self._csr[0] == key:
            return self._csr[1:]
        lib = _abi.load()
        faces = idx.to(torch.int32).contiguous()
        nf, nv = int(faces.shape[0]), int(base.get_max_vert_nindex())
        offsets = torch.empty(nv + 1, dtype=torch.int32, device=faces.device)
        lst = torch.empty(max(3 * nf, 1), dtype=torch.int32, device=faces.device)
        nbytes = ctypes.c_size_t(0)
        _abi.check(lib.tina_vertex_incidence(_abi.ptr(faces), nf, nv, _abi.ptr(offsets), _abi.ptr(lst), None,
                                             ctypes.byref(nbytes), _abi.stream_handle()))
        temp = torch.empty(int(nbytes.value), dtype=torch.uint8, device=faces.device)
        _abi.check(lib.tina_vertex_incidence(_abi.ptr(faces), nf, nv, _abi.ptr(offsets), _abi.ptr(lst),
                                             _abi.ptr(temp), ctypes.byref(nbytes), _abi.stream_handle()))
        self._csr = (key, faces.long(), offsets, lst, nv)
        return self._csr[1:]

    def update_normal(self):
        torch = _torch()
        faces, offsets, lst, nv = self._incidence()
        fb = self.mesh.face_buffers()
        v = fb['verts']
        nf = int(faces.shape[0])
        lib = _abi.load()
        fn = torch.empty((max(nf, 1), 3), dtype=torch.float32, device=v.device)
        _abi.check(lib.tina_face_normals(_abi.ptr(v), nf, 1, _abi.ptr(fn), _abi.stream_handle()))
        out = torch.empty((max(nv, 1), 3), dtype=torch.float32, device=v.device)
        _abi.check(lib.tina_smooth_normals(_abi.ptr(fn), _abi.ptr(offsets), _abi.ptr(lst), nv, _abi.ptr(out),
                                           _abi.stream_handle()))
        self._norm = out[:nv]
        self._version = getattr(self._base, 'version', 0)

    @property
    def norm(self):
        """The vertex normals [nverts, 3] (MeshSmoothNormal.norm)."""
        if self._norm is None:
            self.update_normal()
        return self._norm

    def face_buffers(self):
        if self._norm is None or not self.cached or self._version != getattr(self._base, 'version', 0):
            self.update_normal()
        faces, _, _, _ = self._incidence()
        fb = dict(self.mesh.face_buffers())
        fb['norms'] = self._norm[faces].contiguous()
        return fb


class MeshTransform(MeshEditBase):
    """Applies a 4x4 transform to positions and its inverse transpose to normals
    (tina/mesh/trans.py:5-40): every call of face_buffers (each set_object, as the
    reference's get_face_verts runs inside it) transforms the wrapped mesh's faces
    on the device (tina_mesh_transform) with the reference's mapply_pos order.
    Both matrices are held in binary32 as the reference's fields hold them;
    trans_normal is the upper-left 3x3 of transpose(inv(trans)) taken in float64."""

    def __init__(self, mesh, trans=None):
        super().__init__(mesh)
        self._bufs = {}
        self.set_transform(np.eye(4) if trans is None else trans)

    def set_transform(self, trans):
        trans = np.asarray(trans, dtype=np.float64)
        self.trans = trans
        self._t32 = (ctypes.c_float * 16)(*[float(x) for x in trans.astype(np.float32).reshape(-1)])
        tn = np.transpose(np.linalg.inv(trans))[:3, :3]
        self._tn32 = (ctypes.c_float * 9)(*[float(x) for x in tn.astype(np.float32).reshape(-1)])

    def _buf(self, name, like):
        b = self._bufs.get(name)
        if b is None or b.shape != like.shape or b.device != like.device:
            b = self._bufs[name] = _torch().empty_like(like)
        return b

    def face_buffers(self):
        fb = dict(self.mesh.face_buffers())
        v, n = fb['verts'], fb['norms']
        if v is None:
            return fb
        ov = self._buf('verts', v)
        on = self._buf('norms', n) if n is not None else None
        _abi.check(_abi.load().tina_mesh_transform(
            self._t32, self._tn32, _abi.ptr(v), _abi.ptr(ov), v.numel() // 3,
            _abi.ptr(n) if n is not None else None, _abi.ptr(on) if on is not None else None,
            0 if n is None else n.numel() // 3, _abi.stream_handle()))
        fb['verts'] = ov
        if on is not None:
            fb['norms'] = on
        return fb


class MeshToWire(MeshEditBase):
    """The edges of every polygon as wires (tina/mesh/wire.py:5-39): wire
    n = f * npolygon + e joins corner e of face f to corner (e + 1) % npolygon."""

    def get_npolygon(self):
        return 2

    def _src_npolygon(self):
        return self.mesh.get_npolygon() if hasattr(self.mesh, 'get_npolygon') else 3

    def get_nfaces(self):
        return self.mesh.get_nfaces() * self._src_npolygon()

    def face_buffers(self):
        fb = dict(self.mesh.face_buffers())
        k = self._src_npolygon()
        torch = _torch()
        nxt = torch.tensor([(e + 1) % k for e in range(k)], device=fb['verts'].device) if fb['verts'] is not None \
            else None

        def edges(a):
            if a is None:
                return None
            return torch.stack([a, a[:, nxt]], dim=2).reshape(-1, 2, a.shape[-1]).contiguous()

        out = dict(fb)
        for key in ('verts', 'norms', 'coors', 'colors'):
            out[key] = edges(fb.get(key))
        out['nfaces'] = int(fb['nfaces']) * k
        out['attrs'] = {}
        return out


def _flip_corners(a):
    return None if a is None else a.flip(1).contiguous()


class MeshFlipCulling(MeshEditBase):
    """Every face's corners in reverse order, so culling keeps the other side
    (tina/mesh/cull.py:5-26)."""

    def face_buffers(self):
        fb = dict(self.mesh.face_buffers())
        for key in ('verts', 'norms', 'coors', 'colors'):
            fb[key] = _flip_corners(fb.get(key))
        return fb


class MeshNoCulling(MeshEditBase):
    """Each face twice: face 2n as it is, face 2n + 1 with its corners reversed
    and its normals negated, so both sides survive culling (cull.py:29-57)."""

    def get_nfaces(self):
        return self.mesh.get_nfaces() * 2

    def face_buffers(self):
        torch = _torch()
        fb = dict(self.mesh.face_buffers())

        def pair(a, negate=False):
            if a is None:
                return None
            b = a.flip(1)
            if negate:
                b = -b
            return torch.stack([a, b], dim=1).reshape(-1, *a.shape[1:]).contiguous()
        for key in ('verts', 'coors', 'colors'):
            fb[key] = pair(fb.get(key))
        fb['norms'] = pair(fb.get('norms'), negate=True)
        fb['attrs'] = {k: v.repeat_interleave(2, dim=0).contiguous() for k, v in fb.get('attrs', {}).items()}
        fb['nfaces'] = int(fb['nfaces']) * 2
        return fb


class MeshFlipNormal(MeshEditBase):
    """Normals negated (cull.py:60-66)."""

    def face_buffers(self):
        fb = dict(self.mesh.face_buffers())
        if fb.get('norms') is not None:
            fb['norms'] = (-fb['norms']).contiguous()
        return fb


class MeshGrid:
    """A res.x x res.y grid of vertices (tina/mesh/grid.py:5-74): pos, nrm and
    tex are Fields indexed [i, j]; pos starts as the plane (u * 2 - 1, v * 2 - 1, 0)
    with (u, v) = (i, j) / (res - 1), tex = (u, v).  pre_compute() sets each
    normal to normalize(dx x dy) of the central differences (clamped at the
    border).  Cell (i, j) gives the triangles (a, b, c) and (a, c, d) of its
    corners a = [i, j], b = [i + 1, j], c = [i + 1, j + 1], d = [i, j + 1], or the
    quad (a, b, c, d) with as_quad (for MeshToWire)."""

    def __init__(self, res, as_quad=False, device=None):
        from .field import Field
        torch = _torch()
        if isinstance(res, int):
            res = (res, res)
        self.res = (int(res[0]), int(res[1]))
        self.as_quad = as_quad
        dev = device or torch.device('cuda')
        rx, ry = self.res
        i = torch.arange(rx, dtype=torch.float32, device=dev)[:, None].expand(rx, ry)
        j = torch.arange(ry, dtype=torch.float32, device=dev)[None, :].expand(rx, ry)
        u, v = i / float(rx - 1), j / float(ry - 1)
        self.tex = Field(torch.stack([u, v], dim=2).contiguous(), screen=False)
        self.pos = Field(torch.stack([u * 2 - 1, v * 2 - 1, torch.zeros_like(u)], dim=2).contiguous(), screen=False)
        self.nrm = Field(torch.zeros((rx, ry, 3), dtype=torch.float32, device=dev), screen=False)
        self.version = 0

    def get_npolygon(self):
        return 4 if self.as_quad else 3

    def get_nfaces(self):
        n = (self.res[0] - 1) * (self.res[1] - 1)
        return n if self.as_quad else n * 2

    def pre_compute(self):
        """nrm[i, j] = normalized(dx.cross(dy)) (grid.py:26-35)."""
        torch = _torch()
        p = self.pos.storage
        rx, ry = self.res
        ii = torch.arange(rx, device=p.device)
        jj = torch.arange(ry, device=p.device)
        i1, i2 = (ii + 1).clamp(max=rx - 1), (ii - 1).clamp(min=0)
        j1, j2 = (jj + 1).clamp(max=ry - 1), (jj - 1).clamp(min=0)
        dy = p[:, j1] - p[:, j2]
        dx = p[i1, :] - p[i2, :]
        n = torch.stack([dx[..., 1] * dy[..., 2] - dx[..., 2] * dy[..., 1],
                         dx[..., 2] * dy[..., 0] - dx[..., 0] * dy[..., 2],
                         dx[..., 0] * dy[..., 1] - dx[..., 1] * dy[..., 0]], dim=2)
        inv = 1.0 / torch.sqrt((n[..., 0] * n[..., 0] + n[..., 1] * n[..., 1]) + n[..., 2] * n[..., 2])
        self.nrm.storage.copy_(n * inv[..., None])

    def _faces(self, prop):
        a, b = prop[:-1, :-1], prop[1:, :-1]
        c, d = prop[1:, 1:], prop[:-1, 1:]
        if self.as_quad:
            f = [a, b, c, d]
            return _torch().stack(f, dim=2).reshape(-1, 4, prop.shape[-1]).contiguous()
        t0 = _torch().stack([a, b, c], dim=2)
        t1 = _torch().stack([a, c, d], dim=2)
        return _torch().stack([t0, t1], dim=2).reshape(-1, 3, prop.shape[-1]).contiguous()

    def face_buffers(self):
        # cell m = i * (res.y - 1) + j, its two triangles 2m and 2m + 1 (grid.py:46-60)
        return dict(verts=self._faces(self.pos.storage), norms=self._faces(self.nrm.storage),
                    coors=self._faces(self.tex.storage), colors=None, attrs={}, nfaces=self.get_nfaces())


def export_simple_mesh(mesh):
    """{'fv': [n, npolygon, 3], 'fn': ..., 'ft': [n, npolygon, 2]} float32 (tina/mesh/export.py:66-85)."""
    if hasattr(mesh, 'pre_compute'):
        mesh.pre_compute()
    fb = mesh.face_buffers()
    out = {'fv': fb['verts'].detach().cpu().numpy().astype(np.float32)}
    if fb.get('norms') is not None:
        out['fn'] = fb['norms'].detach().cpu().numpy().astype(np.float32)
    if fb.get('coors') is not None:
        out['ft'] = fb['coors'].detach().cpu().numpy().astype(np.float32)
    return out


def simple_mesh_to_connective(obj):
    """One vertex per face corner: 'f' [n, npolygon] = arange, 'v' [n * npolygon, 3]
    (export.py:88-98; normals and texcoords under 'fn' / 'ft' become 'vn' / 'vt')."""
    fv = np.asarray(obj['fv'])
    nf, k = fv.shape[:2]
    ret = {'f': np.arange(nf * k).reshape(nf, k), 'v': fv.reshape(nf * k, 3)}
    if 'fn' in obj:
        ret['vn'] = np.asarray(obj['fn']).reshape(nf * k, 3)
    if 'ft' in obj:
        ret['vt'] = np.asarray(obj['ft']).reshape(nf * k, -1)
    return ret


def export_connective_mesh(mesh):
    """An indexed mesh's own arrays (ConnectiveMesh, MeshModel), else the face soup
    with one vertex per corner (export.py:101-126)."""
    inner = mesh
    while isinstance(inner, MeshEditBase) and not hasattr(inner, 'faces'):
        inner = inner.mesh
    if hasattr(inner, 'faces') and hasattr(inner, 'verts_v') and inner is mesh:
        f = inner.faces.cpu().numpy()
        ret = {'f': f[..., 0] if f.ndim == 3 else f, 'v': inner.verts_v.cpu().numpy()}
        for src, dst in (('norms_v', 'vn'), ('coors_v', 'vt')):
            if getattr(inner, src, None) is not None:
                ret[dst] = getattr(inner, src).cpu().numpy()
        return ret
    return simple_mesh_to_connective(export_simple_mesh(mesh))
