"""The C-ABI libraries load without a GPU and export every declared symbol;
descriptor validation and error paths work on the host."""
import ctypes as C
import re
from pathlib import Path

import numpy as np
import pytest

from atrc_amd import capi, scenes

ROOT = Path(__file__).resolve().parent.parent


def declared(header: str) -> set:
    text = (ROOT / "include" / header).read_text()
    return set(re.findall(r"^\s*(?:int|void|const char \*|const agx_scene_desc \*|void \*)\s*\*?\s*(\w+)\s*\(", text,
                          re.M))


def test_gpu_library_exports_every_declared_symbol():
    lib = capi.load_library()
    names = declared("atrc_gpu.h") | declared("atrc_gpu_debug.h")
    assert {"agx_create", "agx_render", "agx_trace_rays", "agx_upload_scene"} <= names
    for n in names:
        assert hasattr(lib, n), n
    assert set(capi.EXPORTS) == names
    assert lib.agx_abi_version() == capi.AGX_ABI_VERSION


def test_host_library_exports_every_declared_symbol():
    lib = capi.load_host()
    for n in declared("atrc_host.h"):
        assert hasattr(lib, n), n


def test_null_arguments_are_rejected():
    lib = capi.load_library()
    assert lib.agx_create(0, None) == capi.AGX_E_INVALID_ARG
    assert lib.agx_upload_scene(None, None) == capi.AGX_E_INVALID_ARG
    assert lib.agx_render(None, None, None, None, None, None, None) == capi.AGX_E_INVALID_ARG
    assert lib.agx_render_batch(None, None, 0, 0, None) == capi.AGX_E_INVALID_ARG
    assert b"null" in lib.agx_last_error()
    assert lib.agx_render_batch_graph(None, None, 0, 0, None) == capi.AGX_E_INVALID_ARG
    assert b"null" in lib.agx_last_error()


def test_descriptor_validation():
    s, _, _ = scenes.cornell_box()
    d, keep = s.to_desc()
    summ = capi.debug_scene_summary(d)
    assert summ["entities"] == 8 and summ["lights"] == 1 and summ["blas_prims"] == 24
    assert summ["max_tlas_depth"] >= 1
    # bad material index
    s.entities[0].material = 99
    d, keep = s.to_desc()
    with pytest.raises(capi.AgxError) as e:
        capi.debug_scene_summary(d)
    assert e.value.code == capi.AGX_E_INVALID_ARG
    # more lights than the kernels support
    s2 = scenes.Scene()
    m = s2.add_material(scenes.Material())
    for i in range(20):
        s2.add(scenes.Entity(capi.GEOM_QUAD, m, emit=(1, 1, 1),
                             points=((i, 0, 0), (i + 1, 0, 0), (i + 1, 1, 0), (i, 1, 0))))
    d, keep = s2.to_desc()
    with pytest.raises(capi.AgxError) as e:
        capi.debug_scene_summary(d)
    assert e.value.code == capi.AGX_E_UNSUPPORTED
    # wrong ABI version
    d.abi_version = 999
    with pytest.raises(capi.AgxError):
        capi.debug_scene_summary(d)


def test_ctypes_structures_match_the_c_headers(tmp_path):
    """Every ctypes mirror in atrc_amd/capi.py has the size and field offsets
    that the C compiler gives the structure in include/*.h."""
    import shutil
    import subprocess

    if shutil.which("gcc") is None:
        pytest.skip("no C compiler")
    structs = [getattr(capi, n) for n in dir(capi)
               if isinstance(getattr(capi, n), type) and issubclass(getattr(capi, n), C.Structure)
               and getattr(capi, n) is not C.Structure]
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "atrc_gpu.h"', '#include "atrc_gpu_debug.h"',
             '#include "atrc_host.h"', 'int main(void) {']
    for st in structs:
        lines.append(f'printf("{st.__name__} size %zu\\n", sizeof({st.__name__}));')
        for name, _ in st._fields_:
            lines.append(f'printf("{st.__name__} {name} %zu\\n", offsetof({st.__name__}, {name}));')
    lines += ['return 0;', '}']
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c11", f"-I{ROOT / 'include'}", "-o", str(exe), str(src)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    got = {tuple(l.split()[:2]): int(l.split()[2]) for l in out if l.strip()}
    for st in structs:
        assert got[(st.__name__, "size")] == C.sizeof(st), st.__name__
        for name, _ in st._fields_:
            assert got[(st.__name__, name)] == getattr(st, name).offset, (st.__name__, name)


def test_fp16_grids_only_when_exact():
    """A heterogeneous medium's grid is stored in FP16 exactly when FP16 holds
    every texel (FP16 zeros, normals and subnormals); one texel off the FP16
    grid keeps the whole grid FP32."""
    rng = np.random.default_rng(5)
    exact = rng.uniform(0, 2, (8, 8, 8)).astype(np.float16).astype(np.float32)
    exact[0, 0, :4] = [0.0, -0.0, 2.0 ** -24, 3 * 2.0 ** -20]   # zero, signed zero, FP16 subnormals
    off = exact.copy()
    off[3, 4, 5] = np.float32(1.0) + np.float32(2.0 ** -20)      # needs more than FP16's 10 mantissa bits
    tiny = exact.copy()
    tiny[1, 1, 1] = np.float32(2.0 ** -25)                       # below FP16's smallest subnormal
    for grid, f16 in ((exact, True), (off, False), (tiny, False)):
        s, cam, tr = scenes.absorbing_slab()
        s.add_medium(scenes.Medium(kind=capi.MEDIUM_HETEROGENEOUS, density=scenes.Texture3D(grid=grid)))
        desc, keep = s.to_desc()
        sm = capi.debug_scene_summary(desc)
        # density grid only (albedo and g are constant textures)
        assert (sm["grid_texels_f16"], sm["grid_texels_f32"]) == ((512, 0) if f16 else (0, 512))
