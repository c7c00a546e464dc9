"""The scene-file front-end (lightmetrica_amd.scene_xml) on the CPU: the reference's own
transform known answers (test.math.transform.cpp:43-53, test.primitives.cpp:146-176), the
room written as a scene file with transformed unit cubes rendering the same image as the
generator's room through the oracle, and the load errors the reference reports."""
import numpy as np
import pytest

from lightmetrica_amd import scene_xml, scenes


def nanon(assets: str, scene: str, extra: str = "") -> str:
    return f'<nanon version="1.0.dev"><assets>{assets}</assets><scene type="qbvh">{scene}</scene>{extra}</nanon>'


MESH = """<triangle_meshes><triangle_mesh id="mesh1" type="raw">
  <positions>0 0 0  1 0 0  1 1 0</positions><normals>0 0 1  0 0 1  0 0 1</normals><faces>0 1 2</faces>
</triangle_mesh></triangle_meshes>
<bsdfs><bsdf id="bsdf1" type="diffuse"><diffuse_reflectance><color>0.5 0.5 0.5</color></diffuse_reflectance></bsdf></bsdfs>
<films><film id="film1" type="hdr"><width>8</width><height>8</height></film></films>
<cameras><camera id="cam1" type="perspective"><film ref="film1"/><fovy>45</fovy></camera></cameras>"""
CAMERA_NODE = """<node><transform><lookat><position>0 0 5</position><center>0 0 0</center><up>0 1 0</up></lookat>
</transform><camera ref="cam1"/></node>"""


def test_rotate_known_answer():
    """Rotate(90, z) (1, 0, 0) = (0, 1, 0)."""
    np.testing.assert_allclose(scene_xml.rotate(90.0, (0, 0, 1)) @ [1, 0, 0, 1], [0, 1, 0, 1], atol=1e-12)


def test_transform_by_matrix_known_answer():
    """A root transform given as a column-major matrix times the node's: Load_WithTransformByMatrix."""
    root = """<root><transform><matrix>1 0 0 0  0 1 0 0  0 0 1 0  1 2 3 1</matrix></transform>
      <node id="node1"><transform><matrix>2 0 0 0  0 2 0 0  0 0 2 0  0 0 0 1</matrix></transform>
        <triangle_mesh ref="mesh1"/><bsdf ref="bsdf1"/></node>""" + CAMERA_NODE + "</root>"
    f = scene_xml.load_scene_xml(nanon(MESH, root))
    want = np.array([2, 0, 0, 0, 0, 2, 0, 0, 0, 0, 2, 0, 1, 2, 3, 1], dtype=np.float64).reshape(4, 4).T
    np.testing.assert_allclose(f.transforms[f.primitive_ids["node1"]], want)
    # the mesh is flattened with it: (1, 1, 0) -> (3, 4, 3)
    np.testing.assert_allclose(f.scene.positions[2], [3, 4, 3])


def test_translate_rotate_scale_known_answer():
    """translate * rotate * scale: Load_WithTransform, T (1, 1, 1, 1) = (2 sqrt 2 + 1, 3, 1, 1)."""
    root = """<root><node id="node1"><transform><translate>1 1 1</translate>
        <rotate><angle>45</angle><axis>0 1 0</axis></rotate><scale>2 2 2</scale></transform>
        <triangle_mesh ref="mesh1"/><bsdf ref="bsdf1"/></node>""" + CAMERA_NODE + "</root>"
    f = scene_xml.load_scene_xml(nanon(MESH, root))
    t = f.transforms[f.primitive_ids["node1"]] @ np.ones(4)
    np.testing.assert_allclose(t, [2 * np.sqrt(2) + 1, 3, 1, 1], atol=1e-12)


def quad_mesh(mid, corner, a, b):
    corner, a, b = (np.asarray(v, float) for v in (corner, a, b))
    n = np.cross(a, b)
    n /= np.linalg.norm(n)
    P = [corner, corner + a, corner + a + b, corner + b]
    pos = " ".join(f"{x:.9g}" for p in P for x in p)
    nrm = " ".join(f"{x:.9g}" for _ in range(4) for x in n)
    return (f'<triangle_mesh id="{mid}" type="raw"><positions>{pos}</positions><normals>{nrm}</normals>'
            f'<faces>0 1 2 0 2 3</faces></triangle_mesh>')


def unit_cube_mesh(mid):
    """The unit cube centred at the origin, the face order and winding of the room generator's boxes."""
    h = (0.5, 0.5, 0.5)
    faces = [((h[0], -h[1], -h[2]), (0, 1, 0), (0, 0, 1)), ((-h[0], -h[1], -h[2]), (0, 0, 1), (0, 1, 0)),
             ((-h[0], h[1], -h[2]), (0, 0, 1), (1, 0, 0)), ((-h[0], -h[1], -h[2]), (1, 0, 0), (0, 0, 1)),
             ((-h[0], -h[1], h[2]), (1, 0, 0), (0, 1, 0)), ((-h[0], -h[1], -h[2]), (0, 1, 0), (1, 0, 0))]
    pos, nrm, fc = [], [], []
    for k, (c, a, b) in enumerate(faces):
        c, a, b = (np.asarray(v, float) for v in (c, a, b))
        n = np.cross(a, b)
        pos += [c, c + a, c + a + b, c + b]
        nrm += [n] * 4
        fc += [4 * k, 4 * k + 1, 4 * k + 2, 4 * k, 4 * k + 2, 4 * k + 3]
    return (f'<triangle_mesh id="{mid}" type="raw"><positions>{" ".join(f"{x:.9g}" for p in pos for x in p)}'
            f'</positions><normals>{" ".join(f"{x:.9g}" for p in nrm for x in p)}</normals>'
            f'<faces>{" ".join(map(str, fc))}</faces></triangle_mesh>')


def room_xml(W=48, H=48, spp=8) -> str:
    quads = [("floor", (0, 0, 0), (0, 0, 1), (1, 0, 0), "white"), ("ceiling", (0, 1, 0), (1, 0, 0), (0, 0, 1), "white"),
             ("back", (0, 0, 0), (1, 0, 0), (0, 1, 0), "white"), ("left", (0, 0, 0), (0, 1, 0), (0, 0, 1), "red"),
             ("right", (1, 0, 0), (0, 0, 1), (0, 1, 0), "green")]
    meshes = "".join(quad_mesh(n, c, a, b) for n, c, a, b, _ in quads)
    meshes += unit_cube_mesh("cube") + quad_mesh("lightquad", (0.375, 0.999, 0.375), (0.25, 0, 0), (0, 0, 0.25))
    col = {"white": "0.75 0.75 0.75", "red": "0.63 0.06 0.04", "green": "0.14 0.45 0.09", "black": "0 0 0"}
    bsdfs = "".join(f'<bsdf id="{k}" type="diffuse"><diffuse_reflectance><color>{v}</color></diffuse_reflectance></bsdf>'
                    for k, v in col.items())
    assets = (f"<triangle_meshes>{meshes}</triangle_meshes><bsdfs>{bsdfs}</bsdfs>"
              f'<films><film id="film" type="hdr"><width>{W}</width><height>{H}</height></film></films>'
              '<cameras><camera id="cam" type="perspective"><film ref="film"/><fovy>60</fovy></camera></cameras>'
              '<lights><light id="light" type="area"><luminance>17 12 4</luminance></light></lights>')
    nodes = "".join(f'<node><triangle_mesh ref="{n}"/><bsdf ref="{m}"/></node>' for n, *_, m in quads)
    for ctr, size, ang in (((0.35, 0.15, 0.35), (0.3, 0.3, 0.3), 18.0), ((0.68, 0.30, 0.62), (0.3, 0.6, 0.3), -20.0)):
        nodes += (f'<node><transform><translate>{" ".join(map(str, ctr))}</translate><rotate><angle>{ang}</angle>'
                  f'<axis>0 1 0</axis></rotate><scale>{" ".join(map(str, size))}</scale></transform>'
                  f'<triangle_mesh ref="cube"/><bsdf ref="white"/></node>')
    nodes += '<node><triangle_mesh ref="lightquad"/><bsdf ref="black"/><light ref="light"/></node>'
    nodes += ('<node><transform><lookat><position>0.5 0.5 1.95</position><center>0.5 0.5 0</center><up>0 1 0</up>'
              '</lookat></transform><camera ref="cam"/></node>')
    extra = ('<renderer type="pt"><rr_depth>3</rr_depth><max_path_vertices>9</max_path_vertices>'
             '<sampler type="random"><rng>sfmt</rng><rng_seed>1</rng_seed></sampler></renderer>'
             f'<render_scheduler type="mt"><num_samples>{W * H * spp}</num_samples></render_scheduler>')
    return nanon(assets, f"<root>{nodes}</root>", extra)


def test_room_scene_file_matches_generator(oracle):
    f = scene_xml.load_scene_xml(room_xml())
    ref = scenes.room_scene(48, 48, 8)
    assert f.renderer == "pt" and f.scheduler == "mt"
    assert f.scene.num_triangles == ref.num_triangles
    rc = f.scene.render
    assert (rc.num_samples, rc.max_path_vertices, rc.rr_depth, rc.seed) == (48 * 48 * 8, 9, 3, 1)
    np.testing.assert_allclose(f.scene.camera.view, ref.camera.view, atol=1e-7)
    # same triangles (the cubes through translate * rotate * scale)
    np.testing.assert_allclose(np.sort(f.scene.positions[f.scene.faces].reshape(-1, 9), axis=0),
                               np.sort(ref.positions[ref.faces].reshape(-1, 9), axis=0), atol=2e-7)
    a, _ = oracle.Oracle(f.scene, fp64=True, accel="bvh").render("pt", threads=0)
    b, _ = oracle.Oracle(ref, fp64=True, accel="bvh").render("pt", threads=0)
    rel = abs(a.sum() - b.sum()) / b.sum()
    close = (np.abs(a - b) <= 1e-6 * np.maximum(np.abs(b), 1e-2 * np.abs(b).mean())).mean()
    assert rel < 1e-4 and close > 0.99, (rel, close)


@pytest.mark.parametrize("mutate,message", [
    (lambda s: s.replace('version="1.0.dev"', 'version="0.9"'), "different version"),
    (lambda s: s.replace('type="diffuse"><diffuse_reflectance><color>0.75', 'type="glossy"><diffuse_reflectance><color>0.75'),
     "not supported"),
    (lambda s: s.replace('<bsdf ref="black"/><light ref="light"/></node>',
                         '<bsdf ref="black"/><light ref="light"/><node/></node>'), "Leaf node cannot have children"),
    (lambda s: s.replace('<bsdf ref="black"/>', ''), "missing 'bsdf'"),
    (lambda s: s.replace("<nanon", "<lightmetrica").replace("</nanon>", "</lightmetrica>"), "missing <nanon>"),
])
def test_load_errors(mutate, message):
    with pytest.raises(scene_xml.SceneFormatError, match=message):
        scene_xml.load_scene_xml(mutate(room_xml()))


@pytest.mark.gpu
def test_render_command_line(tmp_path):
    """The scene file through the command-line renderer: the saved .pfm equals the GPU render
    of the loaded scene, and both agree with the oracle drawing the same streams."""
    import subprocess
    import sys

    from lightmetrica_amd import film_io
    from oracle import lmref

    xml = tmp_path / "room.xml"
    xml.write_text(room_xml(48, 48, 8))
    out = tmp_path / "room.pfm"
    p = subprocess.run([sys.executable, "-m", "lightmetrica_amd.render", str(xml), str(out), "--film-precision", "fp64"],
                       stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=300)
    print(p.stdout)
    assert p.returncode == 0
    img = film_io.read_pfm(out).astype(np.float64)
    sc = scene_xml.load_scene_xml(xml).scene
    ref, _ = lmref.Oracle(sc, fp64=True, accel="bvh").render("pt", threads=0)
    rel = abs(img.sum() - ref.sum()) / ref.sum()
    assert rel < 5e-3
