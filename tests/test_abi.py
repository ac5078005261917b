"""The C-ABI library loads and exports every entry point include/*.h declares
(no compute calls: these run without a GPU)."""
import ctypes
import glob
import os
import re

from conftest import ROOT


def declared_symbols():
    names = set()
    for h in glob.glob(os.path.join(ROOT, 'include', '*.h')):
        src = open(h).read()
        src = re.sub(r'/\*.*?\*/', '', src, flags=re.S)
        for m in re.finditer(r'^\s*(?:const\s+)?[A-Za-z_][A-Za-z0-9_]*\s*\**\s*(tina_[a-z0-9_]+)\s*\(', src, re.M):
            names.add(m.group(1))
    return names


def test_library_exports_every_declared_symbol():
    from taichi_three_amd import _abi
    lib = _abi.load()
    names = declared_symbols()
    assert 'tina_raster' in names and 'tina_shade' in names
    missing = [n for n in sorted(names) if not hasattr(lib, n)]
    assert not missing, missing
    bound = {n for n, _, _ in _abi.EXPORTS}
    assert names <= bound, sorted(names - bound)


def test_struct_layouts_match_header(tmp_path):
    """The ctypes mirrors have the C compiler's sizes and field offsets."""
    import subprocess
    from taichi_three_amd import _abi
    checks = {
        'tina_camera': (_abi.Camera, ['W2V', 'V2W', 'bias']),
        'tina_lights': (_abi.Lights, ['n', 'dirs', 'colors', 'ambient']),
        'tina_param': (_abi.Param, ['kind', 'dim', 'a', 'tex_w', 'value', 'ptr']),
        'tina_op': (_abi.Op, ['op', 'a', 'b']),
        'tina_material': (_abi.Material, ['nparams', 'nops', 'params', 'ops']),
        'tina_mesh': (_abi.Mesh, ['verts', 'norms', 'coors', 'colors', 'nfaces', 'id_base']),
        'tina_ibl': (_abi.IBLTables, ['diff', 'diff_w', 'spec', 'spec_w', 'spec_h', 'lut', 'rough_lo',
                                      'rough_hi', 'rough_step']),
        'tina_dmesh': (_abi.DMesh, ['verts', 'faces', 'normals', 'colors', 'rough', 'metal', 'intensity',
                                    'nfaces', 'nverts', 'id_base', 'ambient', 'specular']),
        'tina_dgrads': (_abi.DGrads, ['colors', 'normals', 'rough', 'metal', 'intensity']),
        'tina_scene_desc': (_abi.SceneDesc, ['width', 'nfaces', 'shade_flags', 'img', 'cam', 'mesh', 'mat', 'lights',
                                             'bg']),
        'tina_batch_tuning': (_abi.BatchTuning, ['bin_faces', 'max_tiles', 'grid_bin', 'grid_tile', 'grid_shade']),
    }
    lines = ['#include <stdio.h>', '#include <stddef.h>', f'#include "{ROOT}/include/tina_hip.h"',
             'int main(void) {']
    for cname, (_, fields) in checks.items():
        lines.append(f'printf("%zu\\n", sizeof({cname}));')
        for f in fields:
            lines.append(f'printf("%zu\\n", offsetof({cname}, {f}));')
    lines.append('return 0; }')
    src = tmp_path / 'layout.c'
    src.write_text('\n'.join(lines))
    exe = tmp_path / 'layout'
    subprocess.check_call(['gcc', '-o', str(exe), str(src)])
    got = [int(x) for x in subprocess.check_output([str(exe)]).split()]
    want = []
    for cname, (cls, fields) in checks.items():
        want.append(ctypes.sizeof(cls))
        want.extend(getattr(cls, f).offset for f in fields)
    assert got == want


def test_version_and_error_without_gpu():
    from taichi_three_amd import _abi
    lib = _abi.load()
    assert lib.tina_version() == 1
    # argument validation happens before any device call
    assert lib.tina_tonemap_aces(None, 10, None) == _abi.TINA_ERR_INVALID
    assert b'NULL' in lib.tina_last_error(None)
