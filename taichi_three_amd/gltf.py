"""glTF 2.0 scenes (tina/assimp/gltf.py:6-247).

readgltf(path) parses the JSON (a path, a dict or a readable stream), its
buffers (base64 data URIs or files next to the .gltf) and images, and returns a
GltfScene whose extract(scene) adds one MeshModel + MeshTransform object per
primitive of each node of the default scene, with a tina.PBR material built
from pbrMetallicRoughness (tina.Lambert() when a material has none), as the
reference does:

  - node transform = translate(t) @ quaternion(r) @ scale(s) of the node's
    translation / rotation / scale (gltf.py:78-89); the nodes listed by the
    scene are taken, as in the reference (their children are not walked)
  - baseColorFactor -> basecolor (rgb), baseColorTexture -> Texture(image),
    metallicFactor / roughnessFactor -> metallic / roughness, their textures
    and metallicRoughnessTexture (blue = metallic, green = roughness) ->
    Texture of that channel (gltf.py:115-143)
  - images are decoded to [x, y, c] (rows swapped with columns, no flip, as
    the reference's np.swapaxes; uint8 scaled to [0, 1] by Texture)

Differences, all for files the reference would misread: an accessor's own
byteOffset is honoured (the reference reads from the start of the view), a
texture index goes through textures[i].source (the reference indexes images
directly; the two agree when the lists line up), and a primitive without
normals or texcoords gets zero ones so MeshModel can index them.
"""
import base64
import json
import os

import numpy as np

__all__ = ['readgltf', 'GltfScene', 'GltfNode', 'GltfPrimitive', 'GltfMaterial']

_COMPONENT = {5120: np.int8, 5121: np.uint8, 5122: np.int16, 5123: np.uint16, 5124: np.int32, 5125: np.uint32,
              5126: np.float32}
_WIDTH = {'SCALAR': 1, 'VEC2': 2, 'VEC3': 3, 'VEC4': 4, 'MAT2': 4, 'MAT3': 9, 'MAT4': 16}


class GltfMaterial:
    def __init__(self, name, pbr=None, images=()):
        self.name = name
        self.pbr = pbr
        self._images = images

    def extract(self):
        import taichi_three_amd as tina
        if self.pbr is None:
            return tina.Lambert()
        kw = {}
        for key, value in self.pbr.items():
            if key == 'baseColorFactor':
                kw['basecolor'] = tuple(float(x) for x in value[:3])
            elif key == 'baseColorTexture':
                kw['basecolor'] = tina.Texture(self._images(value['index']))
            elif key == 'metallicFactor':
                kw['metallic'] = float(value)
            elif key == 'metallicTexture':
                kw['metallic'] = tina.Texture(self._images(value['index']))
            elif key == 'roughnessFactor':
                kw['roughness'] = float(value)
            elif key == 'roughnessTexture':
                kw['roughness'] = tina.Texture(self._images(value['index']))
            elif key == 'metallicRoughnessTexture':
                img = self._images(value['index'])
                kw['metallic'] = tina.Texture(img[..., 2])
                kw['roughness'] = tina.Texture(img[..., 1])
        return tina.PBR(**kw)

    def __repr__(self):
        return f'GltfMaterial({self.name!r}, {self.pbr!r})'


class GltfPrimitive:
    def __init__(self):
        self.attrs = {}
        self.indices = None
        self.material = None

    def to_obj_mesh(self):
        """The readobj-style dictionary of this primitive (gltf.py:42-53)."""
        obj = {}
        for name, arr in self.attrs.items():
            if name == 'POSITION':
                obj['v'] = np.asarray(arr, dtype=np.float32)
            elif name == 'NORMAL':
                obj['vn'] = np.asarray(arr, dtype=np.float32)
            elif name.startswith('TEXCOORD') and 'vt' not in obj:
                obj['vt'] = np.asarray(arr, dtype=np.float32)
        nv = len(obj['v'])
        obj.setdefault('vn', np.zeros((nv, 3), np.float32))
        obj.setdefault('vt', np.zeros((nv, 2), np.float32))
        idx = np.arange(nv) if self.indices is None else np.asarray(self.indices).reshape(-1)
        obj['f'] = idx.astype(np.int32).reshape(len(idx) // 3, 3)
        return obj

    def extract(self, scene, trans):
        import taichi_three_amd as tina
        mesh = tina.MeshTransform(tina.MeshModel(self.to_obj_mesh()), trans)
        material = self.material.extract() if self.material is not None else None
        scene.add_object(mesh, material)
        return mesh

    def __repr__(self):
        return f'GltfPrimitive({sorted(self.attrs)}, indices={None if self.indices is None else len(self.indices)})'


class GltfNode:
    def __init__(self, name, translation=None, rotation=None, scale=None, primitives=None):
        self.name = name
        self.trans = (translation, rotation, scale)
        self.primitives = primitives

    def matrix(self):
        from .util.matrix import identity, quaternion, scale, translate
        offset, rotation, factor = self.trans
        m = identity()
        if factor is not None:
            m = scale(factor) @ m
        if rotation is not None:
            m = quaternion(rotation) @ m
        if offset is not None:
            m = translate(offset) @ m
        return m

    def extract(self, scene):
        m = self.matrix()
        return [p.extract(scene, m) for p in (self.primitives or [])]

    def __repr__(self):
        return f'GltfNode({self.name!r}, {self.primitives!r}, {self.trans!r})'


class GltfScene:
    def __init__(self, name, nodes=()):
        self.name = name
        self.nodes = list(nodes)

    def extract(self, scene):
        """Adds every primitive of every node to a tina.Scene; returns the scene (gltf.py:104-107)."""
        for node in self.nodes:
            node.extract(scene)
        return scene

    def __repr__(self):
        return f'GltfScene({self.name!r}, {self.nodes!r})'


def _decode_image(data, mime):
    try:
        from PIL import Image
        import io
        with io.BytesIO(data) as fh:
            return np.array(Image.open(fh))
    except ImportError:
        if data[:8] != b'\x89PNG\r\n\x1a\n':
            raise NotImplementedError(f'glTF image of type {mime or "unknown"}: only PNG is decoded without PIL')
        from .pngread import read_png
        return read_png(data)


def readgltf(path):
    """Parses a glTF 2.0 file (gltf.py:6-247); see the module docstring."""
    base = '.'
    if isinstance(path, str):
        base = os.path.dirname(os.path.abspath(path))
        with open(path, 'rb') as fh:
            root = json.load(fh)
    elif isinstance(path, dict):
        root = path
    else:
        root = json.load(path)

    def load_uri(uri):
        if uri.startswith('data:'):
            return base64.b64decode(uri[uri.index('base64,') + len('base64,'):].encode('ascii'))
        with open(os.path.join(base, uri), 'rb') as fh:
            return fh.read()

    buffers = [load_uri(b['uri']) for b in root.get('buffers', [])]

    def view_bytes(i):
        view = root['bufferViews'][i]
        off = view.get('byteOffset', 0)
        return buffers[view['buffer']][off:off + view['byteLength']], view.get('byteStride')

    def accessor(i):
        acc = root['accessors'][i]
        data, stride = view_bytes(acc['bufferView'])
        dt = np.dtype(_COMPONENT[acc['componentType']])
        width = _WIDTH[acc['type']]
        count = acc['count']
        off = acc.get('byteOffset', 0)
        if stride and stride != width * dt.itemsize:
            rows = np.frombuffer(data, dtype=np.uint8, count=stride * (count - 1) + width * dt.itemsize, offset=off)
            rows = np.lib.stride_tricks.as_strided(rows, shape=(count, width * dt.itemsize), strides=(stride, 1))
            arr = np.ascontiguousarray(rows).view(dt).reshape(count, width)
        else:
            arr = np.frombuffer(data, dtype=dt, count=count * width, offset=off).reshape(count, width)
        return arr[:, 0] if width == 1 else arr

    cache = {}

    def image(tex_index):
        textures = root.get('textures')
        src = textures[tex_index].get('source', tex_index) if textures else tex_index
        if src not in cache:
            img = root['images'][src]
            if 'bufferView' in img:
                data = view_bytes(img['bufferView'])[0]
            else:
                data = load_uri(img['uri'])
            cache[src] = np.swapaxes(_decode_image(bytes(data), img.get('mimeType')), 0, 1)
        return cache[src]

    materials = [GltfMaterial(m.get('name', 'Untitled'), m.get('pbrMetallicRoughness'), image)
                 for m in root.get('materials', [])]
    sc = root['scenes'][root.get('scene', 0)]
    nodes = []
    for node_id in sc.get('nodes', []):
        node = root['nodes'][node_id]
        prims = None
        if 'mesh' in node:
            prims = []
            for p in root['meshes'][node['mesh']]['primitives']:
                if p.get('mode', 4) != 4:
                    raise NotImplementedError('glTF primitives other than triangle lists (mode 4)')
                gp = GltfPrimitive()
                if 'indices' in p:
                    gp.indices = accessor(p['indices'])
                for name, acc_id in p['attributes'].items():
                    gp.attrs[name] = accessor(acc_id)
                if 'material' in p:
                    gp.material = materials[p['material']]
                prims.append(gp)
        nodes.append(GltfNode(node.get('name', 'Untitled'), node.get('translation'), node.get('rotation'),
                              node.get('scale'), prims))
    return GltfScene(sc.get('name', 'Untitled'), nodes)
