"""The .vnf parser (nodes/lex.go, nodes/parser.go) in libvermeer_gpu.so against the oracle's Python
restatement, on the CPU: both dump the parsed nodes in one text form (floats as float32 bits) and keep
the same error log.  The quirks the reference has are pinned by known answers worked from the Go
source: '#' inside a string starts a comment, a number Go's strconv rejects ends the token stream, a
dropped node swallows the next node's body, an unknown parameter skips to the next non-keyword
token, GaussianFilter's untagged NodeDef makes it unparseable, matrices are given by rows.
"""
import struct

import numpy as np
import pytest

import oracle_vnf


def _lib_parse(path):
    from vermeer_amd import _lib
    from vermeer_amd.vnf import VnfFile
    try:
        f = VnfFile(path)
    except _lib.VgpuError as e:
        return None, str(e)
    out = (f.nodes_text(), f.info()["errors"], f.messages())
    f.close()
    return out, None


def _parse_both(tmp_path, text, name="s.vnf"):
    p = tmp_path / name
    p.write_bytes(text.encode() if isinstance(text, str) else text)
    lib, err = _lib_parse(str(p))
    nodes, nerr, log, fatal = oracle_vnf.parse_file(str(p))
    if fatal:
        assert lib is None, "the library accepted a file the reference exits on"
        return None
    assert lib is not None, err
    dump = oracle_vnf.dump(nodes)
    assert lib[0] == dump
    assert lib[1] == nerr
    assert lib[2] == "".join(line + "\n" for line in log)
    return nodes, dump, nerr, log


def _f(x):
    return "%08x" % struct.unpack("<I", struct.pack("<f", x))[0]


SCENE = """# quickstart-like scene
Globals {
\tXRes 64
\tYRes 48
\tMaxIter 3
}
Camera {
\tName "camera"
\tType "LookAt"
\tRoll 2 1 float 0 0.1
\tFrom 2 1 point 0 0.85 4 0 0.85 4
\tTo 1 1 point 0 0.85 -1
\tRadius 0.0
\tFocal  3.5
\tFov 35
\tUp 0 1 0
}
ShaderStd {
\tName "mtl2"
\tDiffuseStrength float 1
\tDiffuseColour rgb 0.8 0.5 0.2
\tSpec1FresnelModel "Metal"
\tSpec1FresnelRefl rgbtex "tex.png"
}
ShaderStd { Name "lightmtl" EmissionColour rgb 1 1 1 EmissionStrength float 10 }
PolyMesh {
\tName "quad"
\tRayBias 0.001
\tVerts 2 4 point -1 0.5 0  -1 1 0   0 1 0  0 0.5 0    -1 0.53 0  -1 1.03 0   0 1.03 0  0 0.53 0
\tPolyCount 1 int 4
\tFaceIdx 4 int 3 2 1 0
\tTransform 1 matrix 1 0 0 0
\t             0 1 0 0
\t             0 0 1 0
\t             0 0 0 1
\tShader 1 string "mtl2"
\tCalcNormals 1
\tNormals 1 2 vec3 0 0 1 0 0 -1
\tNormalIdx 4 int 0 0 1 1
\tUV 1 2 vec2 0 0 1 1
}
TriLight { Name "l1" Shader "lightmtl" P0 0 1.57 0 P1 0.15 1.57 1 P2 1 1.57 0.15 Samples 1 }
DiskLight { Name "d1" Shader "lightmtl" P 0 2 0 Up 1 0 0 LookAt 0 0 0 Radius 0.5 }
SphereLight { Name "s1" Shader "lightmtl" P 0 3 0 Radius 0.25 }
AiryFilter { Name "f" }
OutputHDR { Filename "out.hdr" }
OutputFloat { Filename "out.float" }
Include { Filename "other.vnf" }
"""


def test_scene_parses_identically(tmp_path):
    nodes, dump, nerr, _ = _parse_both(tmp_path, SCENE)
    assert nerr == 0
    assert [n["type"] for n in nodes] == ["Globals", "Camera", "ShaderStd", "ShaderStd", "PolyMesh", "TriLight",
                                         "DiskLight", "SphereLight", "AiryFilter", "OutputHDR", "OutputFloat",
                                         "Include"]
    # a matrix is given by rows and stored transposed; an identity stays an identity
    assert "  Transform: keys 1 per 1 " + " ".join(_f(1.0 if i % 5 == 0 else 0.0) for i in range(16)) in dump
    # defaults of the constructors for parameters not given: not in the dump, present in the nodes
    disk = nodes[6]["p"]
    assert disk["Segments"]["value"] == 20 and not disk["Segments"]["present"]
    assert nodes[7]["p"]["Samples"]["value"] == 1 and not nodes[7]["p"]["Samples"]["present"]
    assert nodes[8]["p"]["Peak"]["value"] == 4.0
    assert '  Spec1FresnelRefl: tex "tex.png"' in dump
    assert "  DiffuseStrength: " + " ".join([_f(1.0)] * 3) in dump     # a float map fills all three channels


def test_matrix_rows(tmp_path):
    nodes, dump, _, _ = _parse_both(tmp_path, 'Camera { Name "c" Type "Matrix" From 0 0 point To 0 0 point Roll 0 0 float '
                                              'Up 0 1 0 WorldToLocal 1 matrix 1 2 3 4 5 6 7 8 9 10 11 12 13 14 15 16 }')
    e = nodes[0]["p"]["WorldToLocal"]["value"][2]
    # element i of the stored (column-major) matrix is row i % 4, column i // 4 of what the file lists
    assert e == [1.0, 5.0, 9.0, 13.0, 2.0, 6.0, 10.0, 14.0, 3.0, 7.0, 11.0, 15.0, 4.0, 8.0, 12.0, 16.0]


def test_comments_and_skipped_characters(tmp_path):
    text = ('Globals { # a comment { } "\n XRes 12 ; @ YRes 7# trailing\n MaxIter 2 }\n'
            'OutputHDR { Filename "a#b" }\n'        # '#' inside a string starts a comment: the string gets '\n'
            '" }\n'
            'OutputFloat { Filename "c" }  # last line without a newline')
    nodes, dump, nerr, _ = _parse_both(tmp_path, text)
    g = nodes[0]["p"]
    assert (g["XRes"]["value"], g["YRes"]["value"], g["MaxIter"]["value"]) == (12, 7, 2)
    assert nodes[1]["p"]["Filename"]["value"] == "a\n"
    assert nodes[2]["p"]["Filename"]["value"] == "c"
    assert nerr == 0


def test_dropped_node_swallows_the_next(tmp_path):
    # GaussianFilter's NodeDef is required and has no parseable type: the node is dropped after its '}',
    # and the recovery loop then skips to the next '}', the end of the Camera
    text = ('GaussianFilter { Name "g" Width 2 Res 17 }\n'
            'Camera { Name "c" Type "LookAt" From 1 1 point 0 0 0 To 1 1 point 0 0 1 Roll 0 0 float Up 0 1 0 }\n'
            'ShaderStd { Name "s" }\n'
            'TriLight { Name "t" P0 0 0 0 }\n'        # missing required parameters: dropped, swallows the next
            'OutputHDR { Filename "x.hdr" }\n'
            'OutputFloat { Filename "y" }\n')
    nodes, _, nerr, log = _parse_both(tmp_path, text)
    assert [n["type"] for n in nodes] == ["ShaderStd", "OutputFloat"]
    assert nerr == 4 and "required field NodeDef" in log[0]


def test_gaussian_filter_with_a_nodedef_token(tmp_path):
    # giving NodeDef (any one token) marks it present, with an error for its type: the filter then parses
    text = 'GaussianFilter { NodeDef x Name "g" Width 1.5 Res 9 }\nOutputHDR { Filename "z" }\n'
    nodes, _, nerr, log = _parse_both(tmp_path, text)
    assert [n["type"] for n in nodes] == ["GaussianFilter", "OutputHDR"]
    assert nerr == 1 and "Invalid type for param" in log[0]
    assert nodes[0]["p"]["Width"]["value"] == 1.5 and nodes[0]["p"]["Res"]["value"] == 9


def test_unknown_parameter_skips_keywords_and_values(tmp_path):
    text = 'ShaderStd { Name "s" Colour rgb 1 2 3 float 4 DiffuseStrength float 0.5 Bogus "x" 7 }\n'
    nodes, dump, nerr, _ = _parse_both(tmp_path, text)
    assert nerr == 2
    assert nodes[0]["p"]["DiffuseStrength"]["value"] == ("c", [0.5, 0.5, 0.5])


def test_wrong_value_types(tmp_path):
    text = ('Globals { XRes "abc" YRes 64.9 UseProgress 2 MaxGoRoutines -7.5 Camera 12 }\n'
            'ShaderStd { Name "s" DiffuseColour 0.5 }\n'          # no map type: the number is left in the stream
            'PolyMesh { Name "m" Verts 1 1 point 1 2 Shader 1 string "s" FaceIdx x }\n')
    nodes, dump, nerr, log = _parse_both(tmp_path, text)
    g = nodes[0]["p"]
    assert g["XRes"]["value"] == 256 and g["XRes"]["present"]       # not set, but present
    assert g["YRes"]["value"] == 64 and g["UseProgress"]["value"] is True
    assert g["MaxGoRoutines"]["value"] == -7 and g["Camera"]["value"] is None
    assert nodes[1]["p"]["DiffuseColour"]["value"] is None
    assert "  DiffuseColour: nil" in dump
    # the point array fails at "Shader" (not a number): Error parsing field, and the name is consumed,
    # so "1" "string" "s" follow as stray tokens
    assert any("Error parsing field Verts" in x for x in log)


def test_numbers(tmp_path):
    text = ('TriLight { Name "t" Shader "s" Samples 3 P0 16777217 -16777219 1152921504606846977 '
            'P1 0.1 -.5 1e-50 P2 3.4028235677973366e38 1E3 -0 }\n')
    nodes, _, _, _ = _parse_both(tmp_path, text)
    L = nodes[0]["p"]
    # integers convert to float32 with one rounding (to nearest even)
    assert L["P0"]["value"][2] == [16777216.0, -16777220.0, float(2 ** 60)]
    assert L["P1"]["value"][2] == [float(np.float32(0.1)), -0.5, 0.0]
    # the float32 midpoint above FLT_MAX rounds (to even) to infinity
    assert L["P2"]["value"][2] == [float("inf"), 1000.0, 0.0]


@pytest.mark.parametrize("value", ["1e+3", "99999999999999999999", "1-2", "1e400", "1e"])
def test_rejected_numbers_are_end_of_file_tokens(tmp_path, value):
    # strconv rejects the run ("1e+3" stops at '+'): the lexer returns its end-of-file token.  As a
    # value the parameter is left unset (but present) and parsing goes on with what follows the run
    nodes, _, nerr, _ = _parse_both(tmp_path, f'Globals {{ XRes {value} YRes 9 }}')
    g = nodes[0]["p"]
    assert g["XRes"]["present"] and g["XRes"]["value"] == 256 and g["YRes"]["value"] == 9
    assert nerr == (1 if value == "1e+3" else 0)       # the '3' after '+' is a stray integer
    # where a parameter name is expected it is an end of file inside the node: the reference reports
    # it until its error limit exits
    assert _parse_both(tmp_path, f'Globals {{ {value} }} OutputHDR {{ Filename "a" }}') is None


def test_end_of_file_inside_a_node(tmp_path):
    assert _parse_both(tmp_path, 'Globals { XRes 5 ') is None


def test_rejected_number_at_top_level(tmp_path):
    nodes, _, nerr, _ = _parse_both(tmp_path, 'OutputHDR { Filename "a" } 1e+3 OutputFloat { Filename "b" }')
    assert [n["type"] for n in nodes] == ["OutputHDR"] and nerr == 0


def test_duplicate_and_unregistered(tmp_path):
    text = ('Globals { XRes 10 XRes 20 }\n'       # dropped at the second XRes; recovery skips to its own '}'
            'Teapot { Size 3 }\n'                 # not registered: skipped to its '}'
            'Globals { XRes 30 }\n')
    nodes, _, nerr, log = _parse_both(tmp_path, text)
    assert [n["p"]["XRes"]["value"] for n in nodes] == [30]
    assert nerr == 3 and "Node type Teapot not registered" in log[2]


def test_slices(tmp_path):
    text = ('PolyMesh { Name "m" Verts 0 3 point 0 0 0 1 0 0 0 1 0 Shader 2 string "a" "b" '
            'FaceIdx 3 int 0 1 2 ShaderIdx 1 "int" 4294967297 PolyCount 0 int }\n')
    nodes, dump, nerr, _ = _parse_both(tmp_path, text)
    p = nodes[0]["p"]
    assert p["ShaderIdx"]["value"] == [1]                 # int32 conversion wraps
    assert p["PolyCount"]["value"] == []                  # read, empty: not nil
    assert p["Verts"]["value"][:2] == (0, 3) and len(p["Verts"]["value"][2]) == 9
    assert nerr == 0


def test_too_many_errors(tmp_path):
    assert _parse_both(tmp_path, "Globals { " + "Bad 1 " * 11 + "}") is None
    nodes, _, nerr, _ = _parse_both(tmp_path, "Globals { " + "Bad 1 " * 10 + "}")
    assert nerr == 10 and len(nodes) == 1


def test_camera_default_names(tmp_path):
    text = ('Camera { Type "LookAt" From 1 1 point 0 0 0 To 1 1 point 0 0 1 Roll 0 0 float Up 0 1 0 }\n'
            'Camera { Type "LookAt" From 1 1 point 0 0 0 To 1 1 point 0 0 1 Roll 0 0 float Up 0 1 0 }\n')
    # Name is required even though the constructor sets one: both are dropped (the second is swallowed)
    nodes, _, nerr, _ = _parse_both(tmp_path, text)
    assert nodes == [] and nerr == 2


def test_every_registered_type(tmp_path):
    text = ('DebugShader { Name "d" Sides 2 Colour rgb 1 0 0 }\n'
            'QuadLight { Name "q" P 0 0 0 U 1 0 0 V 0 1 0 Shader "d" }\n'
            'Sphere { Name "sp" P 0 0 0 Shader "d" Radius 2 }\n'
            'Proc { Name "p" Handler "wfobj" Data "x.obj" BMin 1 1 point 0 0 0 BMax 1 1 point 1 1 1 }\n'
            'GeomInstance { Name "i" Geom "sp" BMin 1 1 point 0 0 0 BMax 1 1 point 1 1 1 Transform 1 matrix '
            '1 0 0 0 0 1 0 0 0 0 1 0 0 0 0 1 }\n')
    nodes, _, nerr, _ = _parse_both(tmp_path, text)
    assert [n["type"] for n in nodes] == ["DebugShader", "QuadLight", "Sphere", "Proc", "GeomInstance"]
    assert nodes[1]["p"]["Samples"]["value"] == 1 and nodes[2]["p"]["Radius"]["value"] == 2.0
    assert nerr == 0


def test_file_errors(tmp_path):
    from vermeer_amd import _lib
    from vermeer_amd.vnf import VnfFile
    with pytest.raises(_lib.VgpuError):
        VnfFile(str(tmp_path / "missing.vnf"))
    (tmp_path / "bad.vnf").write_text("Globals { XRes 0 }")
    with pytest.raises(_lib.VgpuError, match="positive"):
        VnfFile(str(tmp_path / "bad.vnf"))
