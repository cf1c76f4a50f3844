"""Pins the oracle's restatement of PreRender for .vnf files (oracle/oracle_vnf.py build()) with known
answers worked from the Go source:

  polymesh/init.go:12-134   PolyCount fans (0, i, i+1), ShaderIdx per polygon, NormalIdx per corner
                            (FaceIdx when absent), FaceIdx-only and index-free meshes
  camera.go:80-98           Aspect from the frame (float32 XRes / YRes), TanThetaFocal =
                            float32(tan(float64(Fov/2 * Pi / 180))) * Focal with Pi = 3.14159265358
  core/core.go:69-89        AddNode order decides geoms and lights; a light's PreRender adds its
                            geometry as a new node, prerendered in the next round
  core/render.go:146-163    the camera is the node named by Globals.Camera, else "camera"
  instance.go:116-146,      a GeomInstance takes its geom by name and its bounds from BMin[0] / BMax[0];
  proc.go:140-172,          a Proc with the "vnf" handler reads its file at PreRender: the file's shaders
  proc/vnf/vnf.go:19-56     join the scene, its geoms belong to the Proc alone (hidden from the top
                            level), its other nodes are dropped; a Proc without Transform holds the identity
"""
import math

import numpy as np
import pytest

import oracle_vnf


def _build(tmp_path, text):
    p = tmp_path / "s.vnf"
    p.write_text(text)
    return oracle_vnf.build(str(p))


CAM = ('Camera { Name "camera" Type "LookAt" From 1 1 point 0 0 5 To 1 1 point 0 0 0 Roll 0 0 float '
       'Up 0 1 0 Fov 50 Focal 3.5 }\n')
SH = 'ShaderStd { Name "a" DiffuseStrength float 1 }\nShaderStd { Name "b" Spec1Strength float 1 }\n'


def test_polycount_fans_and_indices(tmp_path):
    b = _build(tmp_path, 'Globals { XRes 30 YRes 20 }\n' + CAM + SH +
               'PolyMesh { Name "m" Verts 1 7 point 0 0 0 1 0 0 1 1 0 0 1 0 2 0 0 3 0 0 3 1 0 '
               'PolyCount 2 int 4 3 FaceIdx 7 int 0 1 2 3 4 5 6 ShaderIdx 2 int 1 0 '
               'Shader 2 string "a" "b" Normals 1 2 vec3 0 0 1 0 0 -1 NormalIdx 7 int 0 1 0 1 1 1 0 }\n')
    m = b["desc"]["meshes"][0]
    assert m["idx"].tolist() == [[0, 1, 2], [0, 2, 3], [4, 5, 6]]
    assert m["shaderidx"].tolist() == [1, 1, 0]                      # per polygon, repeated per triangle
    assert m["normal_idx"].tolist() == [0, 1, 0, 0, 0, 1, 1, 1, 0]   # NormalIdx at corners 0,1,2 0,2,3 4,5,6
    assert m["shaders"].tolist() == [0, 1]                           # materials in first-use order
    assert b["camera"]["aspect"] == np.float32(1.5)
    ttf = np.float32(math.tan(float(np.float32(np.float32(25.0) * np.float32(3.14159265358)) / np.float32(180))))
    assert b["camera"]["tan_theta_focal"] == np.float32(ttf * np.float32(3.5))
    assert abs(b["camera"]["tan_theta_focal"] - math.tan(math.radians(25)) * 3.5) < 1e-6
    assert (b["width"], b["height"], b["max_iter"]) == (30, 20, 0)


def test_faceidx_only_and_implicit(tmp_path):
    b = _build(tmp_path, CAM + SH +
               'PolyMesh { Name "f" Verts 1 4 point 0 0 0 1 0 0 1 1 0 0 1 0 FaceIdx 6 int 0 1 2 2 3 0 '
               'ShaderIdx 2 int 0 257 Shader 1 string "a" Normals 1 4 vec3 0 0 1 0 0 1 0 0 1 0 0 1 }\n'
               'PolyMesh { Name "v" Verts 1 3 point 0 0 1 1 0 1 1 1 1 Shader 1 string "b" }\n')
    f, v = b["desc"]["meshes"]
    assert f["idx"].tolist() == [[0, 1, 2], [2, 3, 0]]
    assert f["normal_idx"].tolist() == [0, 1, 2, 2, 3, 0]              # FaceIdx when NormalIdx is absent
    assert f["shaderidx"].tolist() == [0, 1]                         # uint8(257) = 1
    assert v["idx"].tolist() == [[0, 1, 2]] and "normal_list" not in v
    assert (b["width"], b["height"], b["max_iter"]) == (1024, 1024, 16)   # defaultGlobals without Globals


def test_lights_add_geometry_in_order(tmp_path):
    b = _build(tmp_path, CAM + SH +
               'SphereLight { Name "s" P 0 3 0 Radius 0.5 Shader "a" }\n'
               'PolyMesh { Name "m" Verts 1 3 point 0 0 0 1 0 0 0 1 0 Shader 1 string "b" }\n'
               'TriLight { Name "t" P0 0 2 0 P1 1 2 0 P2 0 2 1 Shader "a" Samples 2 }\n'
               'DiskLight { Name "d" P 0 2 0 Up 1 0 0 LookAt 0 0 0 Radius 1 Shader "a" Segments 4 }\n')
    meshes, lights = b["desc"]["meshes"], b["desc"]["lights"]
    # the file's geoms first, then the light geometry in the order the lights were prerendered
    assert [("sphere" if m.get("sphere") else len(m["idx"])) for m in meshes] == [1, "sphere", 1, 4]
    assert [(L["type"], L["geom"], L["samples"]) for L in lights] == [("sphere", 1, 1), ("tri", 2, 2), ("disk", 3, 1)]
    assert meshes[1]["radius"] == 0.5 and meshes[1]["centre"].tolist() == [0, 3, 0]
    tri = meshes[2]
    assert tri["verts"].tolist() == [[0, 2, 0], [1, 2, 0], [0, 2, 1]]
    # Tri.createMesh's normal: normalize((P1 - P0) x (P2 - P0)) = (0, -1, 0) on every corner
    assert tri["normal_list"].tolist() == [[0, -1, 0]] * 3
    disk = meshes[3]
    rim = np.linalg.norm(disk["verts"][1::3] - np.array([0, 2, 0], np.float32), axis=1)
    assert np.allclose(rim, 1.0, atol=1e-6) and np.allclose(disk["verts"][0::3], [0, 2, 0])
    # materials in first-use order over the geoms: "b" on the file's mesh, then "a" on the sphere
    assert b["desc"]["materials"]["has_spec1_strength"].tolist() == [1, 0]
    assert b["desc"]["materials"]["has_diffuse_strength"].tolist() == [0, 1]


def test_camera_selection_and_errors(tmp_path):
    two = CAM.replace('"camera"', '"c2"').replace("Fov 50", "Fov 90")
    b = _build(tmp_path, 'Globals { Camera "c2" MaxIter 5 }\n' + CAM + two + SH)
    assert b["camera"]["tan_theta_focal"] == np.float32(np.float32(math.tan(float(np.float32(45) * np.float32(3.14159265358) / np.float32(180)))) * np.float32(3.5))
    assert b["max_iter"] == 5
    with pytest.raises(oracle_vnf.Unsupported, match="no camera"):
        _build(tmp_path, 'Globals { Camera "none" }\n' + CAM + SH)
    with pytest.raises(oracle_vnf.Unsupported, match="shader"):
        _build(tmp_path, CAM + 'PolyMesh { Name "m" Verts 1 3 point 0 0 0 1 0 0 0 1 0 Shader 1 string "x" }\n')
    # a PolyMesh Transform is kept as given (row-major in the file, stored column-major)
    b = _build(tmp_path, CAM + SH + 'PolyMesh { Name "m" Verts 1 3 point 0 0 0 1 0 0 0 1 0 Shader 1 string "a" '
                                    'Transform 1 matrix 1 0 0 1 0 1 0 0 0 0 1 0 0 0 0 1 }\n')
    t = b["desc"]["meshes"][0]["transform"]
    assert t.shape == (1, 16) and t[0, 12] == 1.0 and t[0, 3] == 0.0


def test_include_is_read_at_prerender(tmp_path):
    inc = tmp_path / "inc.vnf"
    inc.write_text(SH + 'Globals { XRes 8 YRes 8 MaxIter 2 }\n')
    b = _build(tmp_path, 'Globals { XRes 40 YRes 10 }\n' + CAM + f'Include {{ Filename "{inc}" }}\n'
               'PolyMesh { Name "m" Verts 1 3 point 0 0 0 1 0 0 0 1 0 Shader 1 string "b" }\n')
    # the frame was created from the first file's Globals; the included Globals still sets MaxIter
    assert (b["width"], b["height"], b["max_iter"]) == (40, 10, 2)
    assert b["camera"]["aspect"] == np.float32(4.0)
    assert b["desc"]["materials"]["has_spec1_strength"].tolist() == [1]


def test_instances_and_procs(tmp_path):
    sub = tmp_path / "sub.vnf"
    sub.write_text('ShaderStd { Name "c" DiffuseStrength float 1 }\n'
                   'PolyMesh { Name "k1" Verts 1 3 point 0 0 0 1 0 0 0 1 0 Shader 1 string "c" }\n'
                   'Sphere { Name "k2" P 0 0 1 Radius 0.5 Shader "a" }\n'
                   'TriLight { Name "dropped" P0 0 2 0 P1 1 2 0 P2 0 2 1 Shader "c" Samples 1 }\n')
    b = _build(tmp_path, CAM + SH +
               'PolyMesh { Name "proto" Verts 1 3 point 0 0 0 1 0 0 0 1 0 Shader 1 string "a" }\n'
               'GeomInstance { Name "i" Geom "proto" BMin 1 1 point 3 -1 -1 BMax 1 1 point -1 3 1 '
               'Transform 2 matrix 1 0 0 1 0 1 0 0 0 0 1 0 0 0 0 1  1 0 0 2 0 1 0 0 0 0 1 0 0 0 0 1 }\n'
               f'Proc {{ Name "p" Handler "vnf" Data "{sub}" BMin 1 1 point -2 -2 -2 BMax 1 1 point 2 2 2 }}\n')
    m = b["desc"]["meshes"]
    assert "idx" in m[0]                                                   # proto
    inst = m[1]["instance"]
    assert inst["geom"] == 0 and inst["matrices"].shape == (2, 16) and inst["matrices"][1, 12] == 2.0
    assert inst["bmin"].tolist() == [3, -1, -1] and inst["bmax"].tolist() == [-1, 3, 1]   # grown at commit
    assert m[2].get("hidden") and m[3].get("hidden") and m[3]["sphere"]   # the Proc's geoms, in file order
    assert m[4]["proc"]["geoms"] == [2, 3] and m[4]["proc"]["matrices"] is None
    assert len(m) == 5 and b["desc"]["lights"] == []                       # the sub-file's TriLight is dropped
    assert len(b["desc"]["materials"]) == 2                                # "a" and the sub-file's "c"
    with pytest.raises(oracle_vnf.Unsupported, match="handler"):
        _build(tmp_path, CAM + SH + 'Proc { Name "p" Handler "ply" Data "x.ply" BMin 1 1 point 0 0 0 BMax 1 1 point 1 1 1 }\n')
    with pytest.raises(oracle_vnf.Unsupported, match="Unable to find geom"):
        _build(tmp_path, CAM + SH + 'GeomInstance { Name "i" Geom "none" BMin 1 1 point 0 0 0 BMax 1 1 point 1 1 1 '
                                    'Transform 1 matrix 1 0 0 0 0 1 0 0 0 0 1 0 0 0 0 1 }\n')
    with pytest.raises(oracle_vnf.Unsupported, match="follows"):
        _build(tmp_path, CAM + SH + 'GeomInstance { Name "i" Geom "later" BMin 1 1 point 0 0 0 BMax 1 1 point 1 1 1 '
                                    'Transform 1 matrix 1 0 0 0 0 1 0 0 0 0 1 0 0 0 0 1 }\n'
                                    'PolyMesh { Name "later" Verts 1 3 point 0 0 0 1 0 0 0 1 0 Shader 1 string "a" }\n')


def test_wfobj_handler_parses_as_the_reference(tmp_path):
    """builtin/geom/proc/wfobj/parser.go and mtlparser.go: face fields p, p/t, p//n, p/t/n with
    negative (relative) and zero indices, a '#' ending a line, usemtl by the rest of the line (last
    match), a repeated mtllib read once, a missing one skipped, an .mtl whose bad line drops all its
    shaders, Kd / Ks / Ns / illum / Ke / d 0, and the material strengths illum sets."""
    (tmp_path / "a.mtl").write_text(
        "newmtl red\nKd 0.8 0.1 0.1\nillum 1\n"
        "newmtl shiny\nKs 0.9 0.9 0.9\nNs 250\nillum 5\n"
        "newmtl lamp\nKe 3 4 0\nd 0\n")
    (tmp_path / "bad.mtl").write_text("newmtl gone\nKd 1 1 x\n")
    (tmp_path / "m.obj").write_text(
        "# comment\nmtllib a.mtl\nmtllib a.mtl\nmtllib missing.mtl\nmtllib bad.mtl\n"
        "v 0 0 0\nv 1 0 0\nv 1 1 0\nv 0 1 0\nv 0.5 0.5 1e0\n"
        "vn 0 0 1\nvn 0 0 -1\nvt 0 0\nvt 1 0\n"
        "usemtl shiny\nf 1/1/1 2/2/1 3//1 4 # a quad\n"
        "usemtl nothing\nf -1 -4 -3\n"
        "usemtl red\nf 0 2//2 5/-1/-2\n"
        "f 1 2\n")
    b = _build(tmp_path, CAM + f'Proc {{ Name "p" Handler "wfobj" Data "{tmp_path / "m.obj"}" '
                                'BMin 1 1 point -1 -1 -1 BMax 1 1 point 2 2 2 }\n')
    mesh = b["desc"]["meshes"][0]
    assert mesh["hidden"] and b["desc"]["meshes"][1]["proc"]["geoms"] == [0]
    # the quad fans to (0,1,2) (0,2,3); "-1 -4 -3" is (5,2,3) counted from the end -> 0-based 4,1,2;
    # "0" reads as the first vertex; "f 1 2" is a polygon of two vertices: no triangle
    assert mesh["idx"].tolist() == [[0, 1, 2], [0, 2, 3], [4, 1, 2], [0, 1, 4]]
    # usemtl "nothing" keeps the previous material (shiny = 1); "red" selects 0
    assert mesh["shaderidx"].tolist() == [1, 1, 1, 0]
    # normals per corner: vn 1 for the quad ("4" has n = 0 -> 1), the second face has none -> 1,
    # "2//2" -> 2, "5/-1/-2" -> the second-last normal = 1
    assert mesh["normal_idx"].tolist() == [0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 1, 0]
    mats = b["desc"]["materials"]
    # materials in the mesh's Shader order, red, shiny, lamp (bad.mtl's shaders are dropped)
    assert len(mats) == 3
    red, shiny, lamp = mats[0], mats[1], mats[2]
    assert lamp["has_emission_colour"] and lamp["emission_strength"] == np.float32(5)
    assert red["has_diffuse_colour"] and np.allclose(red["diffuse"], [0.8, 0.1, 0.1])
    assert red["has_diffuse_strength"] and red["diffuse_strength"] == np.float32(0.5) and not red["has_spec1_strength"]
    assert shiny["spec1_roughness"] == np.float32(1) - np.float32(250) / np.float32(1000)
    assert shiny["diffuse_strength"] == np.float32(0.2) and shiny["spec1_strength"] == np.float32(0.5)
    assert shiny["fresnel_model"] == 2                                   # illum 5: "Metal"


def test_wfobj_mtl_values_and_errors(tmp_path):
    from oracle_vnf import _parse_mtl, _parse_obj, _MtlError, _go_float32, _go_atoi
    (tmp_path / "c.mtl").write_text("newmtl lamp\nKe 3 4 0\nd 0\nNs 4096\nillum 12\nnewmtl tex\nmap_Kd  a b.png\n")
    lamp, tex = _parse_mtl(str(tmp_path / "c.mtl"), str(tmp_path))
    p = lamp["p"]
    assert p["EmissionColour"]["value"] == ("c", [0.6000000238418579, 0.800000011920929, 0.0])
    assert p["EmissionStrength"]["value"][1][0] == 5.0
    assert p["DiffuseColour"]["value"][1] == [0.5, 0.5, 0.5]            # d 0: 'dissolve'
    assert p["Spec1Roughness"]["value"][1][0] == 0.0                      # Max(0, 1 - 4096/2048)
    assert p["DiffuseStrength"]["value"][1][0] == 0.5                     # from d 0; illum 12 sets nothing
    # map_Kd: the rest of the line after one separator, joined to the directory
    assert tex["p"]["DiffuseColour"]["value"] == ("tex", str(tmp_path / " a b.png"))
    (tmp_path / "e.mtl").write_text("Kd 1 1 1\n")
    with pytest.raises(oracle_vnf.Unsupported):                          # a nil dereference in Go
        _parse_mtl(str(tmp_path / "e.mtl"), str(tmp_path))
    (tmp_path / "f.mtl").write_text("newmtl a\nKd x 1 1\nillum q\n")
    with pytest.raises(_MtlError):                                         # Kd's first error is ignored,
        _parse_mtl(str(tmp_path / "f.mtl"), str(tmp_path))                # illum's ends the file
    assert _go_float32("1e") == (False, 0) and _go_float32("0x1p-2") == (True, 0.25)
    assert _go_float32("1e39")[0] is False and np.isinf(_go_float32("1e39")[1])
    assert _go_float32("+nan")[0] is False and _go_float32("-Inf")[1] == -np.inf
    assert _go_atoi("+7") == (True, 7) and _go_atoi("1.5") == (False, 0)
