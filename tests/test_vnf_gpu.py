"""A .vnf scene file built and rendered by the library (vgpu_vnf_load / vgpu_vnf_build, then render
passes and the output drivers) against the oracle's restatement of the same PreRender
(core/core.go:36-89, the nodes' PreRender methods, camera.go:80-207) feeding the oracle renderer.

The scene exercises what PreRender decides: a PolyMesh with PolyCount fans, per-polygon ShaderIdx
and Normals on their own NormalIdx; a moving PolyMesh with per-corner normals on both keys; a Sphere
geom; Tri, Disk and Sphere lights (each adding its own geometry); an Included file (read at
PreRender) holding a ShaderStd; a LookAt camera with two From keys, a Roll and a thin lens; a
GaussianFilter made parseable by a NodeDef token; OutputHDR and OutputFloat nodes.  A second scene
holds a GeomInstance with two transform keys, a PolyMesh with its own Transform and a Proc reading
another file through the "vnf" handler.

Bar: with both sides on the IEEE 1/sqrt estimate, every pass's samples and the film are
bit-identical, and the files written are byte-identical to the oracle's writers on the oracle's film.
"""
import numpy as np
import pytest

import oracle
import oracle_vnf
from vermeer_amd import output, scenes
from vermeer_amd.vnf import VnfFile, render_file

pytestmark = pytest.mark.gpu

W, H, PASSES = 96, 64, 4


def _scene(tmp_path):
    inc = tmp_path / "materials.vnf"
    inc.write_text('ShaderStd { Name "metal" Spec1Colour rgb 0.9 0.8 0.6 Spec1Strength float 1 '
                   'Spec1Roughness float 0.2 Spec1FresnelModel "Metal" '
                   'Spec1FresnelRefl rgb 0.9 0.7 0.5 Spec1FresnelEdge rgb 0.95 0.9 0.8 }\n')
    text = f"""# scene for the loader parity test
Globals {{ XRes {W} YRes {H} MaxIter {PASSES} }}
Include {{ Filename "{inc}" }}
Camera {{
    Name "camera"
    Type "LookAt"
    From 2 1 point 0.2 1.0 4.0   -0.2 1.1 3.8
    To 1 1 point 0 0.8 0
    Roll 2 1 float 0 0.05
    Up 0 1 0
    Fov 50
    Focal 3.5
    Radius 0.01
}}
ShaderStd {{ Name "white" DiffuseColour rgb 0.8 0.8 0.8 DiffuseStrength float 1 }}
ShaderStd {{ Name "red" DiffuseColour rgb 0.8 0.2 0.1 DiffuseStrength float 1 DiffuseRoughness float 0.4 }}
ShaderStd {{ Name "glow" EmissionColour rgb 1 0.9 0.8 EmissionStrength float 12 DiffuseStrength float 1 }}
PolyMesh {{
    Name "room"
    Verts 1 10 point -3 0 -3  3 0 -3  3 0 3  -3 0 3   -3 3 -3  3 3 -3   -3 0 -3  -3 3 -3  -3 3 3  -3 0 3
    PolyCount 3 int 4 4 4
    FaceIdx 12 int 0 3 2 1   0 1 5 4   6 9 8 7
    ShaderIdx 3 int 0 1 0
    Shader 2 string "white" "red"
    Normals 1 3 vec3 0 1 0  0 0 1  1 0 0
    NormalIdx 12 int 0 0 0 0  1 1 1 1  2 2 2 2
}}
PolyMesh {{
    Name "mover"
    RayBias 0.0001
    Verts 2 3 point  0.5 0.2 0.5  1.5 0.2 0.5  1.0 1.2 0.2    0.7 0.2 0.5  1.7 0.2 0.5  1.2 1.2 0.2
    Shader 1 string "metal"
    Normals 2 3 vec3 0 0 1  0 0.2 1  0.2 0 1    0 0 1  0.1 0.1 1  0 0.1 1
}}
Sphere {{ Name "ball" P -0.8 0.6 0.3 Radius 0.6 Shader "red" }}
TriLight {{ Name "tri" P0 -0.5 2.9 -0.5 P1 0.5 2.9 -0.5 P2 0 2.9 0.5 Shader "glow" Samples 1 }}
DiskLight {{ Name "disk" P 1.5 2.5 1.5 Up 1 0 0 LookAt 1.5 0 1.5 Radius 0.3 Shader "glow" Segments 12 Samples 1 }}
SphereLight {{ Name "bulb" P -1.5 2.2 1.0 Radius 0.15 Shader "glow" }}
GaussianFilter {{ NodeDef x Name "gauss" Width 2 Res 17 }}
OutputHDR {{ Filename "{tmp_path / 'out.hdr'}" }}
OutputFloat {{ Filename "{tmp_path / 'out.float'}" }}
"""
    p = tmp_path / "scene.vnf"
    p.write_text(text)
    return p


def _nan_equal(a, b):
    a = np.asarray(a, np.float32)
    b = np.asarray(b, np.float32)
    return bool(np.all((a.view(np.uint32) == b.view(np.uint32)) | (np.isnan(a) & np.isnan(b))))


def test_vnf_scene_bitexact(tmp_path):
    print("\n[case] vnf scene: build and render", flush=True)
    path = _scene(tmp_path)
    f = VnfFile(path)
    assert f.info() == {"width": W, "height": H, "max_iter": PASSES, "errors": 1}   # the NodeDef token
    R = f.build(0)
    oracle.set_normalize_ieee(True)    # before the oracle's PreRender: the camera basis normalizes too
    try:
        built = oracle_vnf.build(str(path))
        assert len(built["camera"]["keys"]) == 2 and built["filter"][0] == "gaussian"
        assert [L["type"] for L in built["desc"]["lights"]] == ["tri", "disk", "sphere"]
        o = oracle_vnf.renderer(built)
        scr = scenes.pixel_scrambles(W, H, seed=4)
        for it in range(1, PASSES + 1):
            print(f"[case] vnf pass {it}", flush=True)
            ng, gs = R.render_pass(scr, it, want_sample=True)
            no, os_ = o.render_pass(scr, it, want_sample=True)
            assert ng == no, f"pass {it}: shadow rays {ng} vs {no}"
            assert _nan_equal(gs, os_), f"pass {it}: samples differ"
        assert _nan_equal(R.film, o.film)
        assert (R.film > 0).mean() > 0.5
    finally:
        oracle.set_normalize_ieee(False)
    R.close()
    f.close()


def test_vnf_render_file_writes_outputs(tmp_path):
    print("\n[case] vnf render_file and output nodes", flush=True)
    path = _scene(tmp_path)
    oracle.set_normalize_ieee(True)
    try:
        film = render_file(path, maxiter=2, seed=9, log=lambda *a, **k: None)
        o = oracle_vnf.renderer(oracle_vnf.build(str(path)))
        scr = scenes.pixel_scrambles(W, H, seed=9)
        for it in (1, 2):
            o.render_pass(scr, it)
    finally:
        oracle.set_normalize_ieee(False)
    assert _nan_equal(film, o.film)
    ref_hdr, ref_float = tmp_path / "ref.hdr", tmp_path / "ref.float"
    oracle.write_hdr(o.film, str(ref_hdr))
    oracle.write_float(o.film, str(ref_float))
    assert (tmp_path / "out.hdr").read_bytes() == ref_hdr.read_bytes()
    assert (tmp_path / "out.float").read_bytes() == ref_float.read_bytes()


CAMERA = ('Camera { Name "camera" Type "LookAt" From 1 1 point 0 1.2 4 To 1 1 point 0 0.5 0 Roll 0 0 float '
          'Up 0 1 0 Fov 50 Focal 3.5 }\n')


def _instance_scene(tmp_path):
    sub = tmp_path / "held.vnf"
    sub.write_text('ShaderStd { Name "blue" DiffuseColour rgb 0.1 0.2 0.8 DiffuseStrength float 1 '
                   'Spec1Strength float 0.3 Spec1Roughness float 0.3 }\n'
                   'PolyMesh { Name "fin_a" Verts 1 3 point -0.3 0 0  0.3 0 0  0 0.7 0 Shader 1 string "blue" }\n'
                   'PolyMesh { Name "fin_b" Verts 1 3 point -0.3 0 -0.4  0.3 0 -0.4  0 0.5 -0.4 Shader 1 string "white" }\n'
                   'Sphere { Name "knob" P 0 0.9 -0.2 Radius 0.15 Shader "blue" }\n'
                   'TriLight { Name "dropped" P0 0 3 0 P1 1 3 0 P2 0 3 1 Shader "blue" Samples 1 }\n')
    text = f"""Globals {{ XRes {W} YRes {H} MaxIter {PASSES} }}
{CAMERA}
ShaderStd {{ Name "white" DiffuseColour rgb 0.8 0.8 0.8 DiffuseStrength float 1 }}
ShaderStd {{ Name "red" DiffuseColour rgb 0.8 0.2 0.1 DiffuseStrength float 1 }}
ShaderStd {{ Name "glow" EmissionColour rgb 1 0.9 0.8 EmissionStrength float 12 DiffuseStrength float 1 }}
PolyMesh {{
    Name "floor"
    Verts 1 4 point -3 0 -3  3 0 -3  3 0 3  -3 0 3
    PolyCount 1 int 4
    FaceIdx 4 int 0 3 2 1
    Shader 1 string "white"
}}
PolyMesh {{
    Name "proto"
    Verts 1 4 point 0 0 0  0.5 0 0  0 0.5 0  0 0 0.5
    FaceIdx 12 int 0 2 1  0 1 3  0 3 2  1 2 3
    Shader 1 string "red"
}}
GeomInstance {{
    Name "spun"
    Geom "proto"
    BMin 1 1 point -2 -0.5 -1
    BMax 1 1 point 2 1.5 2
    Transform 2 matrix 1 0 0 0.8  0 1 0 0.1  0 0 1 0.4  0 0 0 1
                       0.866 0 0.5 1.1  0 1 0 0.3  -0.5 0 0.866 0.5  0 0 0 1
}}
PolyMesh {{
    Name "turned"
    Verts 1 3 point -1.5 0.1 -0.5  -0.9 0.1 -0.5  -1.5 0.7 -0.5
    FaceIdx 3 int 0 1 2
    Shader 1 string "white"
    Transform 1 matrix 0 -1 0 -0.8  1 0 0 1.6  0 0 1 0  0 0 0 1
}}
Proc {{
    Name "held"
    Handler "vnf"
    Data "{sub}"
    BMin 1 1 point -1 -0.5 -1.5
    BMax 1 1 point 1 1.5 1.5
    Transform 1 matrix 1 0 0 -0.2  0 1 0 0  0 0 1 1.2  0 0 0 1
}}
TriLight {{ Name "tri" P0 -0.5 2.9 -0.5 P1 0.5 2.9 -0.5 P2 0 2.9 0.5 Shader "glow" Samples 1 }}
"""
    p = tmp_path / "instances.vnf"
    p.write_text(text)
    return p


def test_vnf_instances_procs_transforms_bitexact(tmp_path):
    """A GeomInstance with two transform keys, a PolyMesh with its own Transform (a quarter turn about
    its box centre, so that rays entering its untransformed bounds find it) and a Proc whose "vnf" handler
    reads a file of two meshes and a sphere (its TriLight dropped), under a moved Proc transform.  Camera
    rays hit each of them."""
    print("\n[case] vnf instances, procs and transforms", flush=True)
    path = _instance_scene(tmp_path)
    f = VnfFile(path)
    R = f.build(0)
    oracle.set_normalize_ieee(True)
    try:
        built = oracle_vnf.build(str(path))
        kinds = [("instance" if "instance" in m else "proc" if "proc" in m else "hidden" if m.get("hidden")
                  else "geom") for m in built["desc"]["meshes"]]
        assert kinds == ["geom", "geom", "instance", "geom", "hidden", "hidden", "hidden", "proc", "geom"]
        o = oracle_vnf.renderer(built)
        # camera rays reach the floor, the prototype, the instance, the turned mesh and the proc
        rays = scenes.pinhole_rays(W, H, fov_deg=50.0, eye=(0.0, 1.2, 4.0), target=(0.0, 0.5, 0.0))
        rays["time"] = np.linspace(0, 1, len(rays), dtype=np.float32)
        hit_slots = set(np.unique(o.scene.trace(rays)[0]["geom"]).tolist())
        assert {0, 1, 2, 3, 4} <= hit_slots
        scr = scenes.pixel_scrambles(W, H, seed=6)
        for it in range(1, PASSES + 1):
            print(f"[case] vnf instance pass {it}", flush=True)
            ng, gs = R.render_pass(scr, it, want_sample=True)
            no, os_ = o.render_pass(scr, it, want_sample=True)
            assert ng == no, f"pass {it}: shadow rays {ng} vs {no}"
            assert _nan_equal(gs, os_), f"pass {it}: samples differ"
        assert _nan_equal(R.film, o.film)
        assert (R.film > 0).mean() > 0.3
    finally:
        oracle.set_normalize_ieee(False)
    R.close()
    f.close()


def _obj_scene(tmp_path):
    (tmp_path / "mats.mtl").write_text(
        "newmtl red\nKd 0.8 0.15 0.1\nillum 1\n"
        "newmtl shiny\nKd 0.2 0.3 0.6\nKs 0.9 0.9 0.8\nNs 300\nillum 2\n"
        "newmtl glow\nKe 2 1.8 1.5\nKd 0.5 0.5 0.5\nillum 1\n")
    # a pyramid on a square base (quads and triangles, corner normals, relative indices), and a small
    # emissive quad
    (tmp_path / "thing.obj").write_text(
        "mtllib mats.mtl\n"
        "v -0.5 0 -0.5\nv 0.5 0 -0.5\nv 0.5 0 0.5\nv -0.5 0 0.5\nv 0 0.9 0\n"
        "vn 0 -1 0\nvn 0 0.5 -1\nvn 1 0.5 0\nvn 0 0.5 1\nvn -1 0.5 0\n"
        "usemtl shiny\nf 1//1 4//1 3//1 2//1\n"
        "usemtl red\nf 1//2 2//2 5//2\nf 2//3 3//3 5//3\nf -3//4 -2//4 -1//4\nf 4//5 1//5 5//5\n"
        "v 0.8 0.6 -0.2\nv 1.0 0.6 -0.2\nv 1.0 0.8 -0.2\nv 0.8 0.8 -0.2\n"
        "usemtl glow\nf -4 -3 -2 -1\n")
    text = f"""Globals {{ XRes {W} YRes {H} MaxIter {PASSES} }}
{CAMERA}
ShaderStd {{ Name "white" DiffuseColour rgb 0.8 0.8 0.8 DiffuseStrength float 1 }}
ShaderStd {{ Name "light" EmissionColour rgb 1 0.9 0.8 EmissionStrength float 12 DiffuseStrength float 1 }}
PolyMesh {{
    Name "floor"
    Verts 1 4 point -3 0 -3  3 0 -3  3 0 3  -3 0 3
    PolyCount 1 int 4
    FaceIdx 4 int 0 3 2 1
    Shader 1 string "white"
}}
Proc {{
    Name "obj"
    Handler "wfobj"
    Data "{tmp_path / 'thing.obj'}"
    BMin 1 1 point -1.5 -0.5 -1.5
    BMax 1 1 point 1.5 1.5 1.5
    Transform 1 matrix 1 0 0 0.1  0 1 0 0  0 0 1 0.3  0 0 0 1
}}
TriLight {{ Name "tri" P0 -0.5 2.9 -0.5 P1 0.5 2.9 -0.5 P2 0 2.9 0.5 Shader "light" Samples 1 }}
"""
    p = tmp_path / "obj.vnf"
    p.write_text(text)
    return p


def test_vnf_wfobj_proc_bitexact(tmp_path):
    """A Proc with the wfobj handler: an .obj mesh (quads and triangles, corner normals, relative
    indices) with an .mtl of three materials (Kd / Ks / Ns / illum / Ke), under a moved Proc
    transform; rendered bit-exact against the oracle's reading of the same files."""
    print("\n[case] vnf wfobj proc", flush=True)
    path = _obj_scene(tmp_path)
    f = VnfFile(path)
    R = f.build(0)
    oracle.set_normalize_ieee(True)
    try:
        built = oracle_vnf.build(str(path))
        assert len(built["desc"]["materials"]) == 5   # white, light and the three .mtl materials
        o = oracle_vnf.renderer(built)
        scr = scenes.pixel_scrambles(W, H, seed=12)
        for it in range(1, PASSES + 1):
            print(f"[case] vnf wfobj pass {it}", flush=True)
            ng, gs = R.render_pass(scr, it, want_sample=True)
            no, os_ = o.render_pass(scr, it, want_sample=True)
            assert ng == no, f"pass {it}: shadow rays {ng} vs {no}"
            assert _nan_equal(gs, os_), f"pass {it}: samples differ"
        assert _nan_equal(R.film, o.film)
        rays = scenes.pinhole_rays(W, H, fov_deg=50.0, eye=(0.0, 1.2, 4.0), target=(0.0, 0.5, 0.0))
        assert (o.scene.trace(rays)[0]["geom"] == 1).sum() > 100   # the Proc is in view
    finally:
        oracle.set_normalize_ieee(False)
    R.close()
    f.close()


def test_vnf_build_errors(tmp_path):
    from vermeer_amd import _lib
    cases = {
        "nocam.vnf": 'ShaderStd { Name "s" }\n',
        "noshader.vnf": 'TriLight { Name "t" P0 0 0 0 P1 1 0 0 P2 0 1 0 Shader "none" Samples 1 }\n',
        "handler.vnf": ('ShaderStd { Name "s" }\nProc { Name "p" Handler "ply" Data "x.ply" '
                        'BMin 1 1 point 0 0 0 BMax 1 1 point 1 1 1 }\n'),
        "noobj.vnf": (CAMERA + 'Proc { Name "p" Handler "wfobj" Data "no/such.obj" '
                      'BMin 1 1 point 0 0 0 BMax 1 1 point 1 1 1 }\n'),
        "forward.vnf": (CAMERA + 'ShaderStd { Name "s" }\nGeomInstance { Name "i" Geom "m" BMin 1 1 point 0 0 0 '
                        'BMax 1 1 point 1 1 1 Transform 1 matrix 1 0 0 0 0 1 0 0 0 0 1 0 0 0 0 1 }\n'
                        'PolyMesh { Name "m" Verts 1 3 point 0 0 0 1 0 0 0 1 0 Shader 1 string "s" }\n'),
        "quad.vnf": 'ShaderStd { Name "s" }\nQuadLight { Name "q" P 0 0 0 U 1 0 0 V 0 1 0 Shader "s" }\n',
    }
    for name, text in cases.items():
        print(f"[case] vnf build error {name}", flush=True)
        p = tmp_path / name
        p.write_text(text)
        f = VnfFile(p)
        with pytest.raises(_lib.VgpuError):
            f.build(0)
        f.close()
