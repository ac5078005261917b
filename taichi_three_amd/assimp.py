"""Wavefront OBJ input and output (tina/assimp/obj.py).

`readobj` returns the reference's dictionary: 'v' [nv, 3], 'vt' [nt, 2], 'vn'
[nn, 3] float32 and 'f' [nf, 3, 3] int32 index triples (vertex, texcoord,
normal; 0-based), plus 'usemtl' ([first face, name] pairs) and 'mtllib'.
Polygons are split into a fan around their first corner (quads as (0, 1, 2),
(2, 3, 0)); a missing index ('v//vn', 'v/vt', 'v') reads as 0, so every face
holds three triples (the reference keeps only the indices a file gives, which
its MeshModel cannot take when normals or texcoords are absent); empty arrays
read as one zero row.  `MeshModel` (tina/mesh/model.py) turns the dictionary into a mesh the
raster takes; everything here is host-side input handling, run once per asset.
"""
import io

import numpy as np

__all__ = ['readobj', 'writeobj', 'objorient', 'objautoscale', 'objverts', 'objnorms', 'objcoors', 'objmtlids',
           'objunpackmtls', 'objmknorm']


def _fan(corners):
    """Triangles of one polygon face (obj.py:4-14)."""
    n = len(corners)
    if n < 3:
        raise ValueError(f'an OBJ face needs at least 3 corners, got {n}')
    if n == 4:
        return [[corners[0], corners[1], corners[2]], [corners[2], corners[3], corners[0]]]
    return [[corners[0], corners[k], corners[k + 1]] for k in range(1, n - 1)]


def _numbers(fields):
    try:
        return [float(x) for x in fields]
    except ValueError:
        return None


def readobj(path, orient='xyz', scale=None, simple=False, usemtl=True, quadok=False):
    """Reads an OBJ file (a path or a binary file object) (obj.py:17-100).

    orient: axis permutation and signs applied to v and vn (see objorient);
    scale: a factor, or 'auto' to centre on the mean vertex and fit [-1, 1];
    simple: return (v, f[:, :, 0]) only; quadok: keep polygons unsplit."""
    if callable(getattr(path, 'read', None)):
        data = path.read()
    else:
        with open(path, 'rb') as fh:
            data = fh.read()
    if isinstance(data, str):
        data = data.encode()
    v, vt, vn, faces, mtls = [], [], [], [], []
    mtllib = None
    records = []
    for raw in data.splitlines():
        parts = raw.strip().split()
        if len(parts) < 2:
            continue
        key, fields = parts[0], parts[1:]
        if key in (b'v', b'vt', b'vn'):
            nums = _numbers(fields)
            if nums is None:
                continue
            {b'v': v, b'vt': vt, b'vn': vn}[key].append(nums)
        else:
            records.append((key, fields))
    # faces after every vertex record, as the reference reads the file twice
    for key, fields in records:
        if key == b'mtllib':
            mtllib = fields[0]
        elif key == b'usemtl':
            mtls.append([len(faces), fields[0]])
        elif key == b'f':
            corners = [([int(x) - 1 if x else 0 for x in field.split(b'/')] + [0, 0])[:3] for field in fields]
            if quadok:
                faces.append(corners)
            else:
                faces.extend(_fan(corners))
    out = {
        'v': np.array(v, dtype=np.float32) if v else np.zeros((1, 3), np.float32),
        'vt': np.array(vt, dtype=np.float32) if vt else np.zeros((1, 2), np.float32),
        'vn': np.array(vn, dtype=np.float32) if vn else np.zeros((1, 3), np.float32),
        'f': np.array(faces, dtype=np.int32) if faces else np.zeros((1, 3, 3), np.int32),
    }
    if usemtl:
        out['usemtl'] = mtls
        out['mtllib'] = mtllib
    if orient is not None:
        objorient(out, orient)
    if scale is not None:
        if scale == 'auto':
            objautoscale(out)
        else:
            out['v'] *= scale
    if simple:
        return out['v'], out['f'][:, :, 0]
    return out


def writeobj(path, obj):
    """Writes 'v', 'vt', 'vn' and 'f' (1-based v/vt/vn triples; a [nf, 3] index
    array writes i/i/i) (obj.py:103-125).  path: a file name or a text file object."""
    lines = ['# OBJ file written by taichi_three_amd.writeobj']
    for key in ('v', 'vt', 'vn'):
        if key in obj:
            lines += [f'{key} ' + ' '.join(repr(float(x)) for x in row) for row in np.asarray(obj[key])]
    if 'f' in obj:
        f = np.asarray(obj['f'])
        if f.ndim >= 3:
            lines += ['f ' + ' '.join('/'.join(str(int(i) + 1) for i in corner) for corner in face) for face in f]
        else:
            lines += ['f ' + ' '.join('/'.join([str(int(i) + 1)] * 3) for i in face) for face in f]
    text = '\n'.join(lines) + '\n'
    if callable(getattr(path, 'write', None)):
        path.write(text)
    else:
        with open(path, 'w') as fh:
            fh.write(text)


def objorient(obj, orient):
    """orient: three letters naming the source axis of x, y and z ('xzy' swaps y
    and z); an upper-case letter negates that axis; a leading '-' reverses every
    face's corner order (obj.py:184-205).  Applies to v and vn in place."""
    reverse = orient.startswith('-')
    if reverse:
        orient = orient[1:]
    src = ['xyz'.index(c.lower()) for c in orient]
    if src != [0, 1, 2]:
        obj['v'][:, :] = obj['v'][:, src]
        obj['vn'][:, :] = obj['vn'][:, src]
    for axis, c in enumerate(orient):
        if c.isupper():
            obj['v'][:, axis] *= -1
            obj['vn'][:, axis] *= -1
    if reverse:
        obj['f'][:] = obj['f'][:, ::-1, :].copy()


def objautoscale(obj):
    """Centres the vertices on their mean and scales the largest coordinate to 1."""
    obj['v'] -= np.average(obj['v'], axis=0)
    obj['v'] /= np.max(np.abs(obj['v']))


def objverts(obj):
    return obj['v'][obj['f'][:, :, 0]]


def objcoors(obj):
    return obj['vt'][obj['f'][:, :, 1]]


def objnorms(obj):
    return obj['vn'][obj['f'][:, :, 2]]


def objmtlids(obj):
    """Per-face material ids: 1 + the order in which each usemtl name first
    appears, 0 before the first usemtl (obj.py:147-163)."""
    f = obj['f']
    ids = np.zeros(len(f), dtype=np.int32)
    names = []
    marks = obj.get('usemtl') or []
    for k, (begin, name) in enumerate(marks):
        end = marks[k + 1][0] if k + 1 < len(marks) else len(f)
        if name not in names:
            names.append(name)
        ids[begin:end] = names.index(name) + 1
    return ids


def objunpackmtls(obj):
    """{material name: an obj holding only that material's faces} (obj.py:128-144)."""
    marks = obj.get('usemtl') or []
    parts = {}
    for k, (begin, name) in enumerate(marks):
        end = marks[k + 1][0] if k + 1 < len(marks) else len(obj['f'])
        chunk = obj['f'][begin:end]
        parts[name] = chunk if name not in parts else np.concatenate([parts[name], chunk], axis=0)
    return {name: {'f': f, 'v': obj['v'], 'vn': obj['vn'], 'vt': obj['vt']} for name, f in parts.items()}


def objmknorm(obj):
    """Replaces the normals by one per face, normalize((c - a) x (b - a)) as the
    reference computes it (obj.py:208-219), and points every corner at it."""
    fv, ft = obj['f'][:, :, 0], obj['f'][:, :, 1]
    p = obj['v'][fv]
    n = np.cross(p[:, 2] - p[:, 0], p[:, 1] - p[:, 0])
    n /= np.linalg.norm(n, axis=1, keepdims=True)
    fn = np.repeat(np.arange(len(fv), dtype=np.int32)[:, None], 3, axis=1)
    obj['vn'] = n.astype(np.float32)
    obj['f'] = np.stack([fv, ft, fn], axis=2).astype(np.int32)


def readobj_text(text, **kwargs):
    """readobj on OBJ source text."""
    return readobj(io.BytesIO(text.encode() if isinstance(text, str) else text), **kwargs)
