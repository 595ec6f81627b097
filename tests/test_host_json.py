"""The C++ host (libatrc_host.so): Atrc JSON -> configuration tree -> factory
-> agx_scene_desc, without a GPU.

Semantics checked against Atrc's factory and config code:
  src/factory/src/utility/config_cvt.cpp:39-105   (floats via std::to_string: 6 decimals)
  src/tracer/src/utility/config.cpp:36-59, 195-314 (vec3 of 1 or 3 values; transform arrays)
  src/factory/include/agz/factory/context.h:179-208 ("reference" creator)
  src/factory/src/creator/scene_creators.cpp:39-43  (entity types ending in "//" are skipped)
"""
import ctypes as C
import json
import math

import numpy as np
import pytest

from atrc_amd import capi, scenes
from oracle import py_oracle as po


def quad_json(A, B, C_, D, transform=None, emit=None, albedo=0.5):
    g = {"type": "quad", "A": A, "B": B, "C": C_, "D": D}
    if transform is not None:
        g["transform"] = transform
    e = {"type": "geometric", "geometry": g,
         "material": {"type": "ideal_diffuse", "albedo": {"type": "constant", "texel": [albedo]}}}
    if emit is not None:
        e["emit_radiance"] = emit
    return e


def wrap(entities, extra_scene=None, renderer=None):
    sc = {"type": "default", "entities": entities}
    sc.update(extra_scene or {})
    return json.dumps({"scene": sc, "rendering": {
        "width": 8, "height": 6, "camera": {"type": "thin_lens", "pos": [0, -5, 0], "dst": [0, 0, 0], "up": [0, 0, 1],
                                            "fov": 45},
        "renderer": renderer or {"type": "gpu_pt", "spp": 4}, "reporter": {"type": "noout"}}})


def entity(desc, i):
    return desc.entities[i]


def test_transform_array_applies_last_first():
    # [translate(1,0,0), rotate_z(90 deg)]: rotate first, then translate
    t = [{"type": "translate", "offset": [1, 0, 0]}, {"type": "rotate_z", "deg": 90}]
    h = capi.HostScene(wrap([quad_json([1, 0, 0], [1, 1, 0], [0, 1, 0], [0, 0, 0], transform=t)]), ".")
    e = entity(h.desc(), 0)
    np.testing.assert_allclose(list(e.p[0]), [1, 1, 0], atol=1e-6)   # (1,0,0) -> (0,1,0) -> (1,1,0)
    np.testing.assert_allclose(list(e.p[1]), [0, 1, 0], atol=1e-6)   # (1,1,0) -> (-1,1,0) -> (0,1,0)
    t2 = [{"type": "scale", "ratio": [2, 3, 4]}, {"type": "rotate", "axis": [0, 0, 1], "rad": math.pi / 2}]
    h2 = capi.HostScene(wrap([quad_json([1, 0, 0], [1, 1, 0], [0, 1, 0], [0, 0, 0], transform=t2)]), ".")
    np.testing.assert_allclose(list(entity(h2.desc(), 0).p[0]), [0, 3, 0], atol=1e-6)


def test_vec3_broadcast_and_six_decimal_floats():
    # a 1-element array is a vec3 of three equal values; |x| < 5e-7 reads back as 0
    h = capi.HostScene(wrap([quad_json([0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0], emit=[2.5],
                                       albedo=4e-7)]), ".")
    e = entity(h.desc(), 0)
    assert list(e.emit_radiance) == [2.5, 2.5, 2.5]
    assert list(h.desc().materials[0].color) == [0.0, 0.0, 0.0]
    # 0.1234567 -> "0.123457"
    h = capi.HostScene(wrap([quad_json([0.1234567, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0])]), ".")
    assert entity(h.desc(), 0).p[0][0] == np.float32(0.123457)


def test_skip_commented_entities_reference_and_aggregate():
    ents = [quad_json([0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0]),
            {"type": "geometric//", "geometry": {"type": "nonsense"}},
            {"type": "reference", "name": ["lib", "lamp"]}]
    extra = {"lib": {"lamp": quad_json([0, 0, 2], [0, 1, 2], [1, 1, 2], [1, 0, 2], emit=[3, 3, 3])},
             "aggregate": {"type": "bvh", "max_leaf_size": 3},
             "env": {"type": "native_sky", "top": [1, 1, 1], "bottom": [0.2, 0.2, 0.2]}}
    h = capi.HostScene(wrap(ents, extra), ".")
    d = h.desc()
    assert d.entity_count == 2 and d.aggregate == capi.AGGREGATE_BVH and d.max_leaf_size == 3
    assert list(entity(d, 1).emit_radiance) == [3, 3, 3]
    assert d.env.kind == capi.ENV_NATIVE_SKY
    s = h.session(0)
    assert (s.width, s.height) == (8, 6) and abs(s.camera.film_aspect - 8 / 6) < 1e-7
    assert abs(s.camera.fov_rad - math.radians(45)) < 1e-6 and s.filter_radius == 0.5


@pytest.mark.parametrize("bad,msg", [
    ({"type": "geometric", "geometry": {"type": "unknown_geom"}, "material": {"type": "ideal_black"}},
     "unknown creator type name: unknown_geom"),
    ({"type": "geometric", "geometry": {"type": "quad", "A": [0, 0], "B": [1, 0, 0], "C": [1, 1, 0], "D": [0, 1, 0]},
      "material": {"type": "ideal_black"}}, "invalid array size"),
    ({"type": "geometric", "geometry": {"type": "sphere", "radius": 1},
      "material": {"type": "phong"}}, "not supported by gpu_pt"),
])
def test_errors_carry_the_creation_hierarchy(bad, msg):
    with pytest.raises(RuntimeError) as ei:
        capi.HostScene(wrap([bad]), ".")
    assert msg in str(ei.value) and "in creating object with factory" in str(ei.value)


def test_cornell_json_equals_python_scene(tmp_path):
    """The Cornell scene written as Atrc JSON (+ .bm meshes) and read back by the
    host yields the same descriptor as the Python-built scene, so the oracle
    renders identical images from both."""
    s, cam, tr = scenes.cornell_box()
    text = scenes.to_atrc_json(s, cam, tr, 24, 24, 2, tmp_path)
    h = capi.HostScene(text, str(tmp_path))
    hd = h.desc()
    pd, keep = s.to_desc()
    assert hd.entity_count == pd.entity_count
    for i in range(pd.entity_count):
        a, b = hd.entities[i], pd.entities[i]
        assert a.geometry == b.geometry and list(a.emit_radiance) == list(b.emit_radiance)
        assert [list(x) for x in a.p] == [list(x) for x in b.p]
        if a.geometry == capi.GEOM_TRIANGLE_BVH:
            n = a.tri_count * 3 * 8
            ta = np.ctypeslib.as_array(C.cast(a.tris, C.POINTER(C.c_float)), (n,))
            tb = np.ctypeslib.as_array(C.cast(b.tris, C.POINTER(C.c_float)), (n,))
            np.testing.assert_array_equal(ta, tb)
    p = scenes.render_params(24, 24, 2, cam, seed=3, **tr)
    ia = po.OracleScene(hd, 32).render(p, threads=4)["image"]
    ib = po.OracleScene(pd, 32).render(p, threads=4)["image"]
    np.testing.assert_array_equal(ia, ib)


def _medium_box_json(material, med_in, med_out=None):
    e = {"type": "geometric",
         "geometry": {"type": "quad", "A": [0, 0, 0], "B": [1, 0, 0], "C": [1, 1, 0], "D": [0, 1, 0]},
         "material": material, "med_in": med_in}
    if med_out is not None:
        e["med_out"] = med_out
    return e


def test_media_and_texture3d_creators(tmp_path):
    # image3d "real" binary file: int32 width, height, depth, then float texels in (z, y, x) order
    W, H, D = 3, 2, 4
    grid = np.arange(W * H * D, dtype=np.float32).reshape(D, H, W) / 10
    with open(tmp_path / "dens.bin", "wb") as f:
        f.write(np.array([W, H, D], np.int32).tobytes())
        f.write(grid.tobytes())
    # gray8 ASCII: Atrc's operator>> on uint8_t reads single characters
    (tmp_path / "alb.txt").write_text("1 1 2\n AZ")
    homo = {"type": "homogeneous", "sigma_a": [0.1, 0.2, 0.3], "sigma_s": [0.5], "g": 0.25,
            "max_scattering_count": 7}
    het = {"type": "heterogeneous",
           "transform": [{"type": "scale", "ratio": [2, 2, 2]}],
           "density": {"type": "image3d", "format": "real", "binary_filename": "${scene-directory}/dens.bin",
                       "sampler": "nearest", "wrap_u": "repeat", "wrap_w": "mirror", "inv_gamma": 2},
           "albedo": {"type": "image3d", "format": "gray8", "ascii_filename": "${scene-directory}/alb.txt"},
           "g": {"type": "constant", "texel": [0.3]},
           "white_for_indirect": 1}
    inv = {"type": "invisible_surface"}
    ents = [_medium_box_json(inv, homo), _medium_box_json(inv, het, {"type": "void"}),
            _medium_box_json({"type": "ideal_black"}, {"type": "void"})]
    h = capi.HostScene(wrap(ents), str(tmp_path))
    d = h.desc()
    assert d.medium_count == 2
    assert d.materials[d.entities[0].material].kind == capi.MAT_INVISIBLE_SURFACE
    assert (d.entities[0].medium_in, d.entities[0].medium_out) == (0, -1)
    assert (d.entities[1].medium_in, d.entities[1].medium_out) == (1, -1)
    assert (d.entities[2].medium_in, d.entities[2].medium_out) == (-1, -1)
    m0, m1 = d.media[0], d.media[1]
    assert m0.kind == capi.MEDIUM_HOMOGENEOUS and m0.max_scattering_count == 7
    np.testing.assert_allclose(list(m0.sigma_a), [0.1, 0.2, 0.3], rtol=1e-6)
    np.testing.assert_allclose(list(m0.sigma_s), [0.5, 0.5, 0.5], rtol=1e-6)
    assert m0.g == pytest.approx(0.25)
    assert m1.kind == capi.MEDIUM_HETEROGENEOUS and m1.white_for_indirect == 1
    assert m1.max_scattering_count == 2 ** 31 - 1
    np.testing.assert_allclose(list(m1.local_to_world), [2, 0, 0, 0, 0, 2, 0, 0, 0, 0, 2, 0], atol=1e-6)
    np.testing.assert_allclose(list(m1.world_to_local), [0.5, 0, 0, 0, 0, 0.5, 0, 0, 0, 0, 0.5, 0], atol=1e-6)
    t = m1.density
    assert (t.kind, t.width, t.height, t.depth, t.channels, t.linear) == (capi.TEX3D_GRID, W, H, D, 1, 0)
    assert list(t.wrap) == [capi.WRAP_REPEAT, capi.WRAP_CLAMP, capi.WRAP_MIRROR] and t.inv_gamma == 2
    np.testing.assert_array_equal(np.ctypeslib.as_array(t.data, (W * H * D,)), grid.ravel())
    a = m1.albedo
    assert (a.width, a.height, a.depth, a.channels, a.linear) == (1, 1, 2, 1, 1)
    np.testing.assert_allclose(np.ctypeslib.as_array(a.data, (2,)), [ord("A") / 255, ord("Z") / 255], rtol=1e-6)
    assert m1.g3.kind == capi.TEX3D_CONSTANT and m1.g3.value[0] == pytest.approx(0.3)
    # the oracle accepts the same descriptor and reads the texels through it
    o = po.OracleScene(d, 64)
    v = o.tex3d(1, "density", [[0.5 / 3, 0.25, 0.125]])[0]
    assert v[0] == pytest.approx(float(grid[0, 0, 0]) ** 2)


def test_texture3d_common_params():
    # inv_u flips u, uvw_perm permutes, then the user transform (applied first)
    het = {"type": "heterogeneous", "transform": [],
           "density": {"type": "image3d", "format": "real", "ascii_filename": "${scene-directory}/d.txt",
                       "inv_u": 1, "uvw_perm": [2, 0, 1], "transform": [{"type": "translate", "offset": [0.1, 0, 0]}]},
           "albedo": {"type": "constant", "texel": [1]}, "g": {"type": "constant", "texel": [0]}}
    import tempfile, os
    with tempfile.TemporaryDirectory() as td:
        with open(os.path.join(td, "d.txt"), "w") as f:
            f.write("1 1 1 0.5")
        h = capi.HostScene(wrap([_medium_box_json({"type": "invisible_surface"}, het)]), td)
        m = np.array(list(h.desc().media[0].density.transform)).reshape(3, 4)
    # p -> T: p + (0.1,0,0) -> Perm: (p_z, p_x + 0.1, p_y) -> inv_u: (1 - p_z, p_x + 0.1, p_y)
    p = np.array([0.2, 0.3, 0.7])
    np.testing.assert_allclose(m[:, :3] @ p + m[:, 3], [1 - 0.7, 0.3, 0.3], atol=1e-6)
    with pytest.raises(RuntimeError, match="wrap"):
        bad = dict(het, density=dict(het["density"], wrap_v="bogus"))
        with tempfile.TemporaryDirectory() as td:
            with open(os.path.join(td, "d.txt"), "w") as f:
                f.write("1 1 1 0.5")
            capi.HostScene(wrap([_medium_box_json({"type": "invisible_surface"}, bad)]), td)


def test_disney_creator_defaults():
    # material_creators.cpp:56-112: required base_color / metallic / roughness, the
    # rest default (transmission 0, ior 1.5, specular_scale 1, clearcoat_gloss 1,
    # transmission_roughness = roughness)
    c = lambda v: {"type": "constant", "texel": v if isinstance(v, list) else [v]}
    mat = {"type": "disney", "base_color": c([0.8, 0.4, 0.2]), "metallic": c(0.3), "roughness": c(0.45),
           "clearcoat": c(0.5)}
    e = {"type": "geometric", "geometry": {"type": "sphere", "radius": 1}, "material": mat}
    h = capi.HostScene(wrap([e]), ".")   # keeps the descriptor alive
    m = h.desc().materials[0]
    assert m.kind == capi.MAT_DISNEY
    np.testing.assert_allclose(list(m.color), [0.8, 0.4, 0.2], rtol=1e-6)
    np.testing.assert_allclose(list(m.color2), [1, 1, 1])
    assert (m.metallic, m.roughness, m.transmission, m.ior) == pytest.approx((0.3, 0.45, 0.0, 1.5))
    assert m.transmission_roughness == pytest.approx(0.45) and m.clearcoat_gloss == 1.0
    assert (m.clearcoat, m.sheen, m.sheen_tint, m.specular_tint, m.anisotropic) == pytest.approx((0.5, 0, 0, 0, 0))
    mat2 = dict(mat, transmission_roughness=c(0.1), ior=c(1.33), specular_scale=c([0.5, 0.6, 0.7]))
    e2 = dict(e, material=mat2)
    h2 = capi.HostScene(wrap([e2]), ".")
    m2 = h2.desc().materials[0]
    assert (m2.transmission_roughness, m2.ior) == pytest.approx((0.1, 1.33))
    np.testing.assert_allclose(list(m2.color2), [0.5, 0.6, 0.7], rtol=1e-6)


def test_gaussian_film_filter_reaches_the_session():
    text = json.loads(wrap([quad_json([0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0])]))
    text["rendering"]["film_filter"] = {"type": "gaussian", "radius": 1.5, "alpha": 2}
    h = capi.HostScene(json.dumps(text), ".")
    s = h.session(0)
    assert s.filter == capi.FILTER_GAUSSIAN and s.filter_radius == 1.5 and s.filter_alpha == 2.0
    text["rendering"]["film_filter"] = {"type": "gaussian", "radius": 1.5, "alpha": -1}
    with pytest.raises(RuntimeError) as ei:
        capi.HostScene(json.dumps(text), ".")
    assert "invalid alpha value" in str(ei.value)


def write_hdr(path, mant, expo, rle):
    """Radiance RGBE file from integer mantissas (h, w, 3) and exponents (h, w)."""
    h, w, _ = mant.shape
    out = bytearray(b"#?RADIANCE\nFORMAT=32-bit_rle_rgbe\n\n" + f"-Y {h} +X {w}\n".encode())
    for y in range(h):
        px = np.concatenate([mant[y], expo[y][:, None]], axis=1).astype(np.uint8)
        if not rle:
            out += px.tobytes()
            continue
        out += bytes([2, 2, w >> 8, w & 255])
        for c in range(4):
            ch = px[:, c]
            x = 0
            while x < w:
                run = 1
                while x + run < w and run < 127 and ch[x + run] == ch[x]:
                    run += 1
                if run >= 3:
                    out += bytes([128 + run, int(ch[x])])
                    x += run
                else:
                    n = min(128, w - x)
                    # literal block up to the next run of 3
                    k = 0
                    while k < n and not (x + k + 2 < w and ch[x + k] == ch[x + k + 1] == ch[x + k + 2]):
                        k += 1
                    k = max(1, k)
                    out += bytes([k]) + bytes(int(v) for v in ch[x:x + k])
                    x += k
    path.write_bytes(bytes(out))


@pytest.mark.parametrize("rle", [False, True])
def test_ibl_hdr_texture_is_read_exactly(tmp_path, rle):
    rng = np.random.default_rng(3)
    h, w = 5, 24
    mant = rng.integers(0, 256, (h, w, 3))
    mant[..., 0] = rng.integers(128, 256, (h, w))
    mant[1, 3:9] = (200, 100, 50)          # runs for the encoder
    expo = rng.integers(120, 140, (h, w))
    write_hdr(tmp_path / "env.hdr", mant, expo, rle)
    expect = (mant * np.ldexp(1.0, expo - 136)[..., None]).astype(np.float32)
    text = json.loads(wrap([quad_json([0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0])]))
    text["scene"]["env"] = {"type": "ibl", "tex": {"type": "hdr", "filename": "${scene-directory}/env.hdr",
                                                    "sample": "nearest", "wrap_u": "repeat"},
                            "no_importance_sampling": 1}
    hs = capi.HostScene(json.dumps(text), str(tmp_path))
    env = hs.desc().env
    assert env.kind == capi.ENV_IBL and env.no_importance_sampling == 1
    t = env.tex
    assert (t.kind, t.width, t.height, t.linear) == (capi.TEX2D_IMAGE, w, h, 0)
    assert list(t.wrap) == [capi.WRAP_REPEAT, capi.WRAP_CLAMP]
    got = np.ctypeslib.as_array(t.data, shape=(h * w * 3,)).reshape(h, w, 3)
    np.testing.assert_array_equal(got, expect)


def test_ibl_checker_and_transform2():
    text = json.loads(wrap([quad_json([0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0])]))
    text["scene"]["env"] = {"type": "ibl", "tex": {
        "type": "checker_board", "grid_count": 4, "color1": [1, 0, 0], "color2": [0, 0, 1], "inv_u": 1,
        "transform": [{"type": "translate", "offset": [0.25, 0]}, {"type": "scale", "ratio": 2}]}}
    hs = capi.HostScene(json.dumps(text), ".")
    t = hs.desc().env.tex
    assert t.kind == capi.TEX2D_CHECKER and t.grid_count == 4
    # inv_u (u -> 1 - u) applied after the user transform: translate(0.25) * scale(2): u -> 2u + 0.25
    np.testing.assert_allclose(list(t.transform), [-2, 0, 0.75, 0, 2, 0], atol=1e-6)


def test_double_sided_and_transform_wrapper(tmp_path):
    """double_sided (double_sided.cpp) sets the entity's double_sided flag around
    any geometry; transform_wrapper (transform_wrapper.cpp) is composed into
    the wrapped geometry: quad corners and mesh vertices are moved, a sphere's
    local-to-world is premultiplied."""
    tr = [{"type": "translate", "offset": [0, 0, 2]}, {"type": "rotate_z", "deg": 90}]
    quad = {"type": "quad", "A": [1, 0, 0], "B": [1, 1, 0], "C": [0, 1, 0], "D": [0, 0, 0]}
    sphere = {"type": "sphere", "radius": 0.5, "transform": [{"type": "translate", "offset": [3, 0, 0]}]}
    tri = np.array([[0, 0, 0, 0, 0, 1, 0, 0], [1, 0, 0, 0, 0, 1, 0, 0], [0, 1, 0, 0, 0, 1, 0, 0]], np.float32)
    scenes.write_bin_mesh(tmp_path / "t.bm", tri)
    mesh = {"type": "triangle_bvh", "filename": "${scene-directory}/t.bm"}
    mat = {"type": "ideal_diffuse", "albedo": {"type": "constant", "texel": [0.5]}}
    ents = [
        {"type": "geometric", "geometry": {"type": "double_sided", "internal": quad}, "material": mat},
        {"type": "geometric", "geometry": {"type": "transform_wrapper", "internal": quad, "transform": tr},
         "material": mat},
        {"type": "geometric", "geometry": {"type": "transform_wrapper", "transform": tr,
                                           "internal": {"type": "double_sided", "internal": sphere}},
         "material": mat},
        {"type": "geometric", "geometry": {"type": "transform_wrapper", "internal": mesh, "transform": tr},
         "material": mat},
    ]
    h = capi.HostScene(wrap(ents), str(tmp_path))
    d = h.desc()
    assert [entity(d, i).double_sided for i in range(4)] == [1, 0, 1, 0]
    rot = lambda p: (-p[1], p[0], p[2] + 2)       # rotate_z(90) first, then translate (0, 0, 2)
    np.testing.assert_allclose(list(entity(d, 1).p[0]), rot((1, 0, 0)), atol=1e-6)
    np.testing.assert_allclose(list(entity(d, 1).p[2]), rot((0, 1, 0)), atol=1e-6)
    l2w = np.array(list(entity(d, 2).local_to_world)).reshape(3, 4)
    np.testing.assert_allclose(l2w @ [0, 0, 0, 1], rot((3, 0, 0)), atol=1e-6)
    np.testing.assert_allclose(l2w[:, :3], [[0, -1, 0], [1, 0, 0], [0, 0, 1]], atol=1e-6)
    m = entity(d, 3)
    v = np.array([[m.tris[k].pos[j] for j in range(3)] for k in range(3)])
    np.testing.assert_allclose(v, [rot(p) for p in tri[:, :3]], atol=1e-6)
    with pytest.raises(RuntimeError, match="double_sided"):
        capi.HostScene(wrap([{"type": "geometric", "material": mat, "geometry": {
            "type": "double_sided", "internal": {"type": "double_sided", "internal": quad}}}]), ".")


def test_stl_meshes_binary_and_ascii(tmp_path):
    """triangle_bvh from STL files (§ mesh formats): binary and ASCII STL give the
    same world-space triangles, each corner carrying the facet normal (or the
    cross product when the file's normal is zero)."""
    import struct
    rng = np.random.default_rng(3)
    v = rng.uniform(-1, 1, (4, 3, 3)).astype(np.float32)
    normals = np.zeros((4, 3), np.float32)
    normals[0] = (0, 0, 1)                              # a stated normal is kept as it is
    with open(tmp_path / "m.stl", "wb") as f:
        f.write(b"solid binary header".ljust(80, b" "))
        f.write(struct.pack("<I", len(v)))
        for n, t in zip(normals, v):
            f.write(struct.pack("<12fH", *n, *t.reshape(-1), 0))
    lines = ["solid m"]
    for n, t in zip(normals, v):
        lines.append(f"facet normal {n[0]:.9g} {n[1]:.9g} {n[2]:.9g}")
        lines.append(" outer loop")
        lines += [f"  vertex {p[0]:.9g} {p[1]:.9g} {p[2]:.9g}" for p in t]
        lines += [" endloop", "endfacet"]
    lines.append("endsolid m")
    (tmp_path / "a.stl").write_text("\n".join(lines) + "\n")
    mat = {"type": "ideal_diffuse", "albedo": {"type": "constant", "texel": [0.5]}}
    ents = [{"type": "geometric", "material": mat,
             "geometry": {"type": "triangle_bvh", "filename": "${scene-directory}/" + f}} for f in ("m.stl", "a.stl")]
    h = capi.HostScene(wrap(ents), str(tmp_path))   # keeps the descriptor's memory alive
    d = h.desc()
    for i in range(2):
        e = entity(d, i)
        assert e.tri_count == 4
        pos = np.array([[e.tris[k].pos[j] for j in range(3)] for k in range(12)], np.float32).reshape(4, 3, 3)
        nor = np.array([[e.tris[k].nor[j] for j in range(3)] for k in range(12)], np.float32).reshape(4, 3, 3)
        np.testing.assert_array_equal(pos, v)
        np.testing.assert_array_equal(nor[0], np.tile([0, 0, 1], (3, 1)))
        for t in range(1, 4):
            c = np.cross(v[t, 1] - v[t, 0], v[t, 2] - v[t, 0])
            np.testing.assert_allclose(nor[t], np.tile(c, (3, 1)), rtol=1e-5, atol=1e-6)
