"""ctypes view of the C ABI in include/tina_hip.h and the loader of libtina_hip.so.

The library is the product: every render call goes through it.  There is no
CPU fallback -- if the shared object is missing or no GPU is visible, the calls
raise instead of computing anything elsewhere.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, 'libtina_hip.so')

TINA_OK = 0
TINA_ERR_INVALID = 1
TINA_ERR_CAPACITY = 2
TINA_ERR_HIP = 3
TINA_ERR_ALLOC = 4

CULL = 1
CLIP = 2

SMOOTH = 1
TEX = 2
VCOLOR = 4
TONEMAP = 8
BG = 16
IBL = 32
NOLIGHT = 64
PARTICLES = 128

MAX_LIGHTS = 16
MAX_PARAMS = 12
MAX_OPS = 16
MAX_STACK = 4
IBL_LEVELS = 7

P_CONST, P_PARAM, P_COLOR, P_POS, P_NORMAL, P_TEXCOORD, P_TEXTURE, P_FACE, P_FRESNEL = range(9)
OP_LAMBERT, OP_PHONG, OP_BLINNPHONG, OP_COOKTORRANCE, OP_EMISSION, OP_SCALE, OP_MIX, OP_ADD = range(8)

MAXDEPTH = 2 ** 30


class Camera(ctypes.Structure):
    _fields_ = [('W2V', ctypes.c_float * 16), ('V2W', ctypes.c_float * 16),
                ('bias', ctypes.c_float * 2), ('pad', ctypes.c_float * 2)]


class Lights(ctypes.Structure):
    _fields_ = [('n', ctypes.c_int32), ('pad', ctypes.c_int32 * 3),
                ('dirs', (ctypes.c_float * 4) * MAX_LIGHTS),
                ('colors', (ctypes.c_float * 4) * MAX_LIGHTS),
                ('ambient', ctypes.c_float * 4)]


class Param(ctypes.Structure):
    _fields_ = [('kind', ctypes.c_int32), ('dim', ctypes.c_int32),
                ('a', ctypes.c_int32), ('b', ctypes.c_int32), ('c', ctypes.c_int32),
                ('tex_w', ctypes.c_int32), ('tex_h', ctypes.c_int32), ('tex_c', ctypes.c_int32),
                ('value', ctypes.c_float * 4), ('ptr', ctypes.c_uint64)]


class Op(ctypes.Structure):
    _fields_ = [('op', ctypes.c_int32), ('a', ctypes.c_int32), ('b', ctypes.c_int32),
                ('pad', ctypes.c_int32)]


class Material(ctypes.Structure):
    _fields_ = [('nparams', ctypes.c_int32), ('nops', ctypes.c_int32), ('pad', ctypes.c_int32 * 2),
                ('params', Param * MAX_PARAMS), ('ops', Op * MAX_OPS)]


class Mesh(ctypes.Structure):
    _fields_ = [('verts', ctypes.c_uint64), ('norms', ctypes.c_uint64),
                ('coors', ctypes.c_uint64), ('colors', ctypes.c_uint64),
                ('nfaces', ctypes.c_int32), ('id_base', ctypes.c_uint32)]


class IBLTables(ctypes.Structure):
    _fields_ = [('diff', ctypes.c_uint64), ('diff_w', ctypes.c_int32), ('diff_h', ctypes.c_int32),
                ('spec', ctypes.c_uint64 * IBL_LEVELS),
                ('spec_w', ctypes.c_int32 * IBL_LEVELS), ('spec_h', ctypes.c_int32 * IBL_LEVELS),
                ('lut', ((ctypes.c_float * 2) * 3) * 3),
                ('rough_lo', ctypes.c_float * 6), ('rough_hi', ctypes.c_float * 6),
                ('rough_step', ctypes.c_float * 6)]


class DMesh(ctypes.Structure):
    _fields_ = [('verts', ctypes.c_uint64), ('faces', ctypes.c_uint64), ('normals', ctypes.c_uint64),
                ('colors', ctypes.c_uint64), ('rough', ctypes.c_uint64), ('metal', ctypes.c_uint64),
                ('intensity', ctypes.c_uint64), ('nfaces', ctypes.c_int32), ('nverts', ctypes.c_int32),
                ('id_base', ctypes.c_uint32), ('ambient', ctypes.c_float * 3), ('specular', ctypes.c_float)]


class SceneDesc(ctypes.Structure):
    _fields_ = [('width', ctypes.c_int32), ('height', ctypes.c_int32), ('nfaces', ctypes.c_int32),
                ('raster_flags', ctypes.c_uint32), ('shade_flags', ctypes.c_uint32), ('pad0', ctypes.c_int32),
                ('img', ctypes.c_uint64), ('cam', Camera), ('mesh', Mesh), ('mat', Material), ('lights', Lights),
                ('bg', ctypes.c_float * 3), ('pad1', ctypes.c_float)]


class BatchTuning(ctypes.Structure):
    _fields_ = [('bin_faces', ctypes.c_int32), ('max_tiles', ctypes.c_int32), ('grid_bin', ctypes.c_int32),
                ('grid_tile', ctypes.c_int32), ('grid_shade', ctypes.c_int32)]


class DGrads(ctypes.Structure):
    _fields_ = [('colors', ctypes.c_uint64), ('normals', ctypes.c_uint64), ('rough', ctypes.c_uint64),
                ('metal', ctypes.c_uint64), ('intensity', ctypes.c_uint64)]


# (name, restype, argtypes) of every exported entry point declared in include/tina_hip.h
_vp = ctypes.c_void_p
_u32 = ctypes.c_uint32
_i32 = ctypes.c_int
_i64 = ctypes.c_int64
EXPORTS = [
    ('tina_version', _i32, []),
    ('tina_last_error', ctypes.c_char_p, [_vp]),
    ('tina_ctx_create', _i32, [_i32, _i32, _i32, ctypes.POINTER(_vp)]),
    ('tina_ctx_destroy', _i32, [_vp]),
    ('tina_ctx_reserve', _i32, [_vp, _i32]),
    ('tina_set_camera', _i32, [_vp, ctypes.POINTER(Camera)]),
    ('tina_clear', _i32, [_vp, _vp]),
    ('tina_raster', _i32, [_vp, _vp, _i32, _u32, _u32, _vp]),
    ('tina_raster_particles', _i32, [_vp, _vp, _vp, _i32, _u32, _u32, _vp]),
    ('tina_raster_wires', _i32, [_vp, _vp, _i32, _u32, _u32, _vp]),
    ('tina_paint_ids', _i32, [_vp, _u32, _i32, _vp, _vp, _u32, _vp, _vp]),
    ('tina_normals', _i32, [_vp, ctypes.POINTER(Mesh), _u32, _vp, _vp]),
    ('tina_gbuffer', _i32, [_vp, ctypes.POINTER(Mesh), _u32, _i32, _i32, _vp, ctypes.c_float, _vp, _vp]),
    ('tina_ssao', _i32, [_vp, _vp, _vp, _i32, _vp, _i32, ctypes.c_float, ctypes.c_float, ctypes.c_float, _vp, _vp]),
    ('tina_ssao_apply', _i32, [_vp, _i32, _i32, _vp, _i32, _i32, _vp]),
    ('tina_ssr', _i32, [_vp, ctypes.POINTER(Mesh), ctypes.POINTER(Material), _vp, _vp, _vp, _i32, _i32,
                        ctypes.c_float, ctypes.c_float, _i32, _i32, _u32, _vp, _vp]),
    ('tina_ssr_apply', _i32, [_vp, _i32, _i32, _vp, _i32, _i32, _vp]),
    ('tina_shade', _i32, [_vp, ctypes.POINTER(Mesh), ctypes.POINTER(Material), ctypes.POINTER(Lights),
                          ctypes.POINTER(IBLTables), _vp, ctypes.POINTER(ctypes.c_float * 3), _u32, _vp]),
    ('tina_fill', _i32, [_vp, _vp, ctypes.POINTER(ctypes.c_float * 3), _u32, _vp]),
    ('tina_tonemap_aces', _i32, [_vp, _i32, _vp]),
    ('tina_resolve', _i32, [_vp, _vp, _vp, _vp]),
    ('tina_keys', _i32, [_vp, ctypes.POINTER(_vp)]),
    ('tina_bake_ibl', _i32, [_vp, _i32, _i32, _vp, _i32, _i32, ctypes.c_float, _i32, _vp]),
    ('tina_nn_samples', _i32, [ctypes.POINTER(Lights), ctypes.POINTER(IBLTables), _u32, _i64, ctypes.c_float,
                               ctypes.c_float, ctypes.c_float, ctypes.c_float, _vp, _vp, _vp]),
    ('tina_mlp_train_step', _i32, [_vp, _vp, _i64, _i64, _i32, _vp, _vp, _vp, _vp, ctypes.c_float, _i32, _vp,
                                   _vp]),
    ('tina_mlp_train_step_static', _i32, [_vp, _vp, _i64, _i64, _i32, _vp, _vp, _vp, ctypes.c_float, _i32, _vp,
                                          _vp]),
    ('tina_mlp_pack', _i32, [_vp, _vp, _vp]),
    ('tina_bloom', _i32, [_vp, _i32, _i32, ctypes.c_float, ctypes.c_float, ctypes.c_float, _vp, _i32, _vp, _vp,
                          _vp]),
    ('tina_fxaa', _i32, [_vp, _i32, _i32, ctypes.c_float, ctypes.c_float, ctypes.c_float, _vp, _vp, _vp]),
    ('tina_accumulate', _i32, [_vp, _vp, _i64, _i32, _vp]),
    ('tina_nn_analytic_rows', _i32, [ctypes.POINTER(Lights), ctypes.POINTER(IBLTables), _vp, _i64,
                                     ctypes.c_float, _vp, _vp]),
    ('tina_mlp_inference', _i32, [_vp, _vp, _i64, _vp, _vp]),
    ('tina_shade_nn', _i32, [_vp, ctypes.POINTER(Mesh), ctypes.POINTER(Material), ctypes.POINTER(Lights),
                             ctypes.POINTER(IBLTables), _vp, _vp, ctypes.POINTER(ctypes.c_float * 3), _u32,
                             ctypes.c_float, _vp]),
    ('tina_nn_fallback_count', _i32, [_vp, _vp, _vp]),
    ('tina_dshade_forward', _i32, [_vp, ctypes.POINTER(DMesh), ctypes.POINTER(Lights), _vp, _vp]),
    ('tina_dshade_backward', _i32, [_vp, ctypes.POINTER(DMesh), ctypes.POINTER(Lights), _vp,
                                    ctypes.POINTER(DGrads), _vp]),
    ('tina_batch_create', _i32, [_i32, ctypes.POINTER(SceneDesc), ctypes.POINTER(BatchTuning),
                                 ctypes.POINTER(_vp)]),
    ('tina_batch_destroy', _i32, [_vp]),
    ('tina_batch_render', _i32, [_vp, _vp]),
    ('tina_batch_capture', _i32, [_vp, _vp]),
    ('tina_batch_launch', _i32, [_vp, _vp]),
    ('tina_batch_keys', _i32, [_vp, _i32, ctypes.POINTER(_vp)]),
    ('tina_resolve_keys', _i32, [_vp, _i64, _vp, _vp, _vp]),
    ('tina_selftest_div', _i32, [_i32, ctypes.c_uint64, _u32, _vp, _vp]),
    ('tina_mesh_transform', _i32, [ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_float), _vp, _vp, _i64,
                                   _vp, _vp, _i64, _vp]),
    ('tina_face_normals', _i32, [_vp, _i64, _i32, _vp, _vp]),
    ('tina_vertex_incidence', _i32, [_vp, _i64, _i32, _vp, _vp, _vp, ctypes.POINTER(ctypes.c_size_t), _vp]),
    ('tina_smooth_normals', _i32, [_vp, _vp, _vp, _i32, _vp, _vp]),
]

_lib = None


def load():
    """Loads libtina_hip.so (built in-tree by __graft_entry__.build()).  Raises if absent."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            f'{LIB_PATH} is missing: build the HIP library first '
            '(python -c "import __graft_entry__ as g; g.build()")')
    lib = ctypes.CDLL(LIB_PATH)
    for name, res, args in EXPORTS:
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


class TinaError(RuntimeError):
    pass


def check(rc, ctx=None):
    if rc != TINA_OK:
        msg = load().tina_last_error(ctx)
        msg = msg.decode() if msg else 'unknown error'
        if rc in (TINA_ERR_INVALID, TINA_ERR_CAPACITY):
            raise ValueError(f'tina: {msg}')
        raise TinaError(f'tina (code {rc}): {msg}')


def stream_handle():
    import torch
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def ptr(t):
    """Address of a buffer: a torch tensor's data_ptr (device memory on the render
    path), a NumPy array's host address, or 0 for None."""
    if t is None:
        return 0
    if hasattr(t, 'data_ptr'):
        return int(t.data_ptr())
    return int(t.ctypes.data)
