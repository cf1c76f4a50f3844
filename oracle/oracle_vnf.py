"""Oracle (test infrastructure only): .vnf parsing and PreRender restated in Python.

  nodes/lex.go          Lex: next (line reads, '#' comments), lex, token, num, str
  nodes/parser.go       parse, node, parseParam and the typed readers (int32slice, rgb, floatmap,
                        vec3, pointarray, vec3array, vec2array, float32array, matrixarray)
  nodes/register.go     createNode; the node types' constructor defaults from their Register calls
  core/core.go:38-89    PreRender rounds, AddNode (geoms, lights, filter, globals, node names)
  builtin/geom/polymesh/init.go:12-134   fan triangulation of PolyCount polygons
  builtin/light/{triangle,disk,sphere}.go PreRender: the geometry each light adds
  builtin/camera/camera.go:80-207        Aspect, TanThetaFocal, LocalToWorld keys
  builtin/geom/instance, proc, proc/vnf  GeomInstance, Proc with the vnf and wfobj handlers
  builtin/geom/proc/wfobj/*.go           the .obj and .mtl readers

dump() writes the parsed nodes in the same text form as the library's vgpu_vnf_text(VGPU_VNF_NODES),
so the two parsers can be compared; build() gives a scene description, camera and filter for
OracleRenderer.
"""
import math
import os
import struct

import numpy as np

EOF, TOK, STR, FLT, INT, OBRACE, CBRACE, OCURLY, CCURLY, COMMA = range(10)

KEYWORDS = ("int", "float", "vec2", "vec3", "point", "rgb", "rgbtex", "matrix")

# Field name, kind, required: the exported, settable fields of each node struct in declaration order
# (a `node:"-"` tag skips a field, `node:",opt"` makes it optional, `node:"Name"` renames it).
SCHEMAS = {
    "Globals": [("NodeDef", "struct", False), ("XRes", "int", False), ("YRes", "int", False),
                ("UseProgress", "bool", False), ("MaxGoRoutines", "int", False), ("Camera", "string", False),
                ("MaxIter", "int", False), ("Output", "string", False)],
    "PolyMesh": [("Name", "string", True), ("RayBias", "float", False), ("Verts", "point", True),
                 ("PolyCount", "ints", False), ("FaceIdx", "ints", False), ("Shader", "strings", True),
                 ("ShaderIdx", "ints", False), ("CalcNormals", "bool", False), ("IsVisible", "bool", False),
                 ("Transform", "matrix", False), ("UV", "vec2", False), ("UVIdx", "ints", False),
                 ("Normals", "vec3", False), ("NormalIdx", "ints", False)],
    "Camera": [("Name", "string", True), ("Type", "string", True), ("From", "point", True), ("To", "point", True),
               ("Roll", "float32array", True), ("Up", "v3", True), ("Aspect", "float", False), ("Fov", "float", False),
               ("Focal", "float", False), ("WorldToLocal", "matrix", False), ("LocalToWorld", "matrix", False),
               ("L", "float", False), ("R", "float", False), ("T", "float", False), ("B", "float", False),
               ("Radius", "float", False)],
    "ShaderStd": [("Name", "string", True), ("EmissionColour", "map", False), ("EmissionStrength", "map", False),
                  ("Sides", "int", False), ("DiffuseColour", "map", False), ("DiffuseStrength", "map", False),
                  ("DiffuseRoughness", "map", False), ("Spec1Colour", "map", False), ("Spec1Strength", "map", False),
                  ("Spec1Roughness", "map", False), ("Spec1FresnelModel", "string", False),
                  ("Spec1FresnelRefl", "map", False), ("Spec1FresnelEdge", "map", False), ("IOR", "map", False)],
    "DebugShader": [("Name", "string", True), ("Sides", "int", False), ("Colour", "map", True)],
    "TriLight": [("Name", "string", True), ("P0", "v3", True), ("P1", "v3", True), ("P2", "v3", True),
                 ("Shader", "string", True), ("Samples", "int", True)],
    "DiskLight": [("Name", "string", True), ("P", "v3", True), ("Up", "v3", True), ("LookAt", "v3", True),
                  ("Radius", "float", True), ("Shader", "string", True), ("Segments", "int", False),
                  ("Samples", "int", False)],
    "SphereLight": [("Name", "string", True), ("P", "v3", True), ("Radius", "float", True), ("Shader", "string", True),
                    ("Samples", "int", False)],
    "QuadLight": [("Name", "string", True), ("P", "v3", True), ("U", "v3", True), ("V", "v3", True),
                  ("Shader", "string", True), ("Samples", "int", False)],
    "Sphere": [("Name", "string", True), ("RayBias", "float", False), ("P", "v3", True), ("Radius", "float", True),
               ("Shader", "string", True)],
    "GaussianFilter": [("NodeDef", "struct", True), ("Name", "string", True), ("Width", "float", True),
                       ("Res", "int", True)],
    "AiryFilter": [("Name", "string", True), ("Width", "float", False), ("Res", "int", False), ("Peak", "float", False)],
    "OutputHDR": [("Filename", "string", True)],
    "OutputFloat": [("Filename", "string", True)],
    "Include": [("NodeDef", "struct", False), ("Name", "string", False), ("Filename", "string", True)],
    "Proc": [("Name", "string", True), ("Handler", "string", True), ("Data", "string", True),
             ("BMin", "point", True), ("BMax", "point", True), ("Transform", "matrix", False)],
    "GeomInstance": [("Name", "string", True), ("Geom", "string", True), ("BMin", "point", True),
                     ("BMax", "point", True), ("Transform", "matrix", True)],
}

ARRAY_WIDTH = {"point": 3, "vec3": 3, "vec2": 2, "float32array": 1, "matrix": 16}
ARRAY_TYPE = {"point": "point", "vec3": "vec3", "vec2": "vec2", "float32array": "float", "matrix": "matrix"}


def f32(x):
    with np.errstate(over="ignore"):
        return float(np.float32(x))


def i64_to_f32(v):
    """float32(int64) on amd64 (CVTSI2SS): one rounding, to nearest even."""
    a = abs(v)
    if a < 1 << 24:
        return float(v)
    shift = a.bit_length() - 24
    q, r = divmod(a, 1 << shift)
    half = 1 << (shift - 1)
    if r > half or (r == half and q & 1):
        q += 1
    return float(q << shift) * (-1 if v < 0 else 1)


def go_int64(f):
    """int64(float64) on amd64 (CVTTSD2SQ): truncation, -2^63 for NaN or out of range."""
    return int(f) if -2.0 ** 63 <= f < 2.0 ** 63 else -2 ** 63


def wrap32(v):
    return (v + 2 ** 31) % 2 ** 32 - 2 ** 31


class Lex:
    def __init__(self, data):
        self.lines = data.split(b"\n")
        self.nl = [True] * (len(self.lines) - 1) + [False]   # the last piece has no newline
        self.li = -1
        self.line = b""
        self.peek = 0
        self.peek_token = None
        self.line_number = 0

    def read_line(self):
        self.line_number += 1
        self.li += 1
        if self.li >= len(self.lines):
            return False
        piece = self.lines[self.li] + (b"\n" if self.nl[self.li] else b"")
        if not piece:
            return False
        self.line = piece
        return True

    def next(self):
        if self.peek:
            r, self.peek = self.peek, 0
            return r
        while not self.line:
            if not self.read_line():
                return 0
        c, self.line = self.line[0], self.line[1:]
        if c == ord("#"):
            if not self.read_line():
                return 0
            return ord("\n")
        return c

    @staticmethod
    def alpha(c):
        return c == ord("_") or (c < 128 and chr(c).isalpha())

    def lex(self):
        while True:
            c = self.next()
            if self.alpha(c):
                return self.token(c)
            if c == 0:
                return EOF, None
            if c == ord("-") or ord("0") <= c <= ord("9"):
                return self.num(c)
            ch = chr(c)
            if ch == "[":
                return OBRACE, None
            if ch == "]":
                return CBRACE, None
            if ch == "{":
                return OCURLY, None
            if ch == "}":
                return CCURLY, None
            if ch == ",":
                return COMMA, None
            if ch == '"':
                return self.string()

    def string(self):
        b = bytearray()
        while True:
            c = self.next()
            if c == 0:
                return EOF, None
            if c == ord('"'):
                return STR, b.decode("utf-8", "replace")
            b.append(c)

    def token(self, c):
        b = bytearray([c])
        while True:
            c = self.next()
            if c == 0:
                return EOF, None
            if chr(c) in '"[]{},':
                self.peek = c
                break
            if chr(c) in " \t\r\n":
                break
            b.append(c)
        return TOK, b.decode("utf-8", "replace")

    def num(self, c):
        b = chr(c)
        is_float = False
        while True:
            c = self.next()
            if c and (chr(c).isdigit() and c < 128 or chr(c) == "-"):
                b += chr(c)
            elif c and chr(c) in ".eE":
                is_float = True
                b += chr(c)
            else:
                break
        if c != 0:
            self.peek = c
        if is_float:
            v = go_parse_float(b)
            return (EOF, None) if v is None else (FLT, v)
        v = go_parse_int(b)
        return (EOF, None) if v is None else (INT, v)

    def Lex(self):
        if self.peek_token is not None:
            t, self.peek_token = self.peek_token, None
            return t
        return self.lex()

    def Peek(self):
        t = self.lex()
        self.peek_token = t
        return t


def go_parse_int(s):
    """strconv.ParseInt(s, 10, 64)."""
    body = s[1:] if s.startswith("-") else s
    if not body or not body.isdigit():
        return None
    v = int(s)
    return v if -2 ** 63 <= v < 2 ** 63 else None


def go_parse_float(s):
    """strconv.ParseFloat(s, 64) over digits, '-', '.', 'e', 'E'."""
    import re
    if not re.fullmatch(r"-?(\d+\.?\d*|\.\d+)([eE]-?\d+)?", s):
        return None
    v = float(s)
    return None if math.isinf(v) else v


class Parser:
    def __init__(self, data, filename, counters):
        self.lex = Lex(data)
        self.filename = filename
        self.nerrors = 0
        self.log = []
        self.fatal = False
        self.counters = counters   # {"camera": n}

    def errorf(self, msg):
        self.log.append(f"{self.filename}:{self.lex.line_number}: {msg}")
        self.nerrors += 1
        if self.nerrors > 10:
            self.fatal = True

    def number(self):
        t, v = self.lex.Lex()
        if t == INT:
            return i64_to_f32(v)
        if t == FLT:
            return f32(v)
        return None

    def slice_(self, p, strings):
        t, v = self.lex.Lex()
        if t != INT:
            return
        count = v
        t, v = self.lex.Lex()
        if t != TOK and v != ("string" if strings else "int"):
            return
        out = []
        for _ in range(count):
            t, v = self.lex.Lex()
            if t != (STR if strings else INT):
                return
            out.append(v if strings else wrap32(v))
        p["value"] = out

    def map_(self, p, what):
        self.lex.Lex()
        if what == "rgbtex":
            t, v = self.lex.Lex()
            if t != STR:
                return
            p["value"] = ("tex", v)
            return
        c = []
        for _ in range(3 if what == "rgb" else 1):
            x = self.number()
            if x is None:
                return
            c.append(x)
        if len(c) == 1:
            c = c * 3
        p["value"] = ("c", c)

    def array(self, p, kind):
        t, v = self.lex.Lex()
        if t != INT:
            return False
        keys, per = v, 1
        if kind != "matrix":
            t, v = self.lex.Lex()
            if t != INT:
                return False
            per = v
        t, v = self.lex.Lex()
        if t != TOK and v != ARRAY_TYPE[kind]:
            return False
        w = ARRAY_WIDTH[kind]
        e = []
        for _ in range((keys or 1) * per):
            m = []
            for _ in range(w):
                x = self.number()
                if x is None:
                    return False
                m.append(x)
            if kind == "matrix":   # the file lists rows; Matrix4Transpose stores columns
                m = [m[c * 4 + r] for r in range(4) for c in range(4)]
            e += m
        p["value"] = (keys, per, e)
        return True

    def param(self, p):
        kind = p["kind"]
        if kind == "int":
            t, v = self.lex.Lex()
            if t == FLT:
                p["value"] = go_int64(v)
            elif t == INT:
                p["value"] = v
        elif kind == "float":
            t, v = self.lex.Lex()
            if t in (FLT, INT):
                p["value"] = float(v)
        elif kind == "bool":
            t, v = self.lex.Lex()
            if t == INT:
                p["value"] = v != 0
        elif kind == "string":
            t, v = self.lex.Lex()
            if t == STR:
                p["value"] = v
        elif kind in ("ints", "strings"):
            t, v = self.lex.Peek()
            if t == INT:
                self.slice_(p, kind == "strings")
            else:
                self.errorf("Invalid token for param (expecting length of slice)")
                self.lex.Lex()
        elif kind == "map":
            t, v = self.lex.Peek()
            if t == TOK and v in ("rgb", "float", "rgbtex"):
                self.map_(p, v)
        elif kind == "v3":
            c = []
            for _ in range(3):
                x = self.number()
                if x is None:
                    return False
                c.append(x)
            p["value"] = (0, 0, c)
        elif kind in ARRAY_WIDTH:
            return self.array(p, kind)
        elif kind == "struct":
            self.errorf("Invalid type for param (core.NodeDef)")
            self.lex.Lex()
        return True

    def skip_to_token(self):
        while True:
            t, v = self.lex.Peek()
            if t == CCURLY or (t == TOK and v not in KEYWORDS) or t == EOF:
                return
            self.lex.Lex()

    def node(self, type_):
        if type_ not in SCHEMAS:
            return None
        n = {"type": type_, "file": self.filename, "line": self.lex.line_number,
             "p": {name: {"kind": k, "req": r, "present": False, "value": None} for name, k, r in SCHEMAS[type_]}}
        defaults(type_, n["p"], self.counters)
        while True:
            if self.fatal:
                return None
            t, v = self.lex.Lex()
            if t == TOK:
                f = n["p"].get(v)
                if f is None:
                    self.errorf(f'Field "{v}" not found in node {type_}')
                    self.skip_to_token()
                    continue
                if f["present"]:
                    self.errorf(f"Field {v} already found in {type_}")
                    return None
                if not self.param(f):
                    self.errorf(f"Error parsing field {v}")
                f["present"] = True
            elif t == CCURLY:
                for name, f in n["p"].items():
                    if f["req"] and not f["present"]:
                        self.errorf(f"node: required field {name} not found in {type_}")
                        return None
                return n
            elif t == EOF:
                self.errorf(f"node: end of file inside node {type_}")
                self.fatal = True
                return None
            else:
                self.errorf(f'node: Parse error, invalid token in object "{t}"')

    def parse(self):
        out = []
        while not self.fatal:
            t, v = self.lex.Lex()
            if t != TOK:
                break
            type_ = v
            if self.lex.Lex()[0] != OCURLY:
                self.errorf("Invalid token in node preamble")
            n = self.node(type_)
            if n is None:
                self.errorf(f"Node is nil: {'<nil>' if type_ in SCHEMAS else 'Node type ' + type_ + ' not registered.'}")
                while True:
                    u = self.lex.Lex()[0]
                    if u in (CCURLY, EOF):
                        break
                continue
            out.append(n)
        return out


def defaults(type_, p, counters):
    """The constructors' field values (their nodes.Register functions)."""
    def put(k, v):
        p[k]["value"] = v
    if type_ == "Globals":
        put("XRes", 256), put("YRes", 256), put("MaxGoRoutines", 5)
    elif type_ == "PolyMesh":
        put("IsVisible", True)
    elif type_ == "Camera":
        put("Focal", 12.0), put("Fov", 90.0), put("Name", f"camera<{counters['camera']}>")
        counters["camera"] += 1
    elif type_ == "DiskLight":
        put("Segments", 20), put("Samples", 1)
    elif type_ == "SphereLight":
        put("Radius", 1.0), put("Samples", 1)
    elif type_ == "QuadLight":
        put("Samples", 1)
    elif type_ == "TriLight":
        put("Samples", 1)
    elif type_ == "Sphere":
        put("Radius", 1.0)
    elif type_ == "GaussianFilter":
        put("Res", 17), put("Width", 2.0)
    elif type_ == "AiryFilter":
        put("Res", 49), put("Width", 6.0), put("Peak", 4.0)
    elif type_ == "OutputHDR":
        put("Filename", "out.hdr")
    elif type_ == "OutputFloat":
        put("Filename", "out.float")


def parse_file(path, counters=None):
    """nodes.Parse: (nodes, nerrors, log lines, fatal)."""
    counters = counters if counters is not None else {"camera": 0}
    with open(path, "rb") as f:
        p = Parser(f.read(), str(path), counters)
    nodes = p.parse()
    return nodes, p.nerrors, p.log, p.fatal


def _hex(x):
    return "%08x" % struct.unpack("<I", struct.pack("<f", x))[0]


def dump(nodes):
    out = []
    for n in nodes:
        out.append(n["type"])
        for name, k, _ in SCHEMAS[n["type"]]:
            f = n["p"][name]
            if not f["present"]:
                continue
            v = f["value"]
            s = f"  {name}:"
            if k == "int":
                s += f" {v if v is not None else 0}"
            elif k == "float":
                s += " " + _hex(f32(v if v is not None else 0.0))
            elif k == "bool":
                s += " 1" if v else " 0"
            elif k == "string":
                s += f' "{v if v is not None else ""}"'
            elif k == "ints":
                s += "".join(f" {x}" for x in (v or []))
            elif k == "strings":
                s += "".join(f' "{x}"' for x in (v or []))
            elif k == "map":
                if v is None:
                    s += " nil"
                elif v[0] == "tex":
                    s += f' tex "{v[1]}"'
                else:
                    s += "".join(" " + _hex(x) for x in v[1])
            elif k == "struct":
                pass
            else:
                keys, per, e = v if v is not None else (0, 0, [])
                s += f" keys {keys} per {per}" + "".join(" " + _hex(x) for x in e)
            out.append(s)
    return "\n".join(out) + ("\n" if out else "")


# ------------------------------------------------------------------ PreRender

def _val(n, k, default=None):
    v = n["p"][k]["value"] if k in n["p"] else None
    return default if v is None else v


def _name(n):
    t = n["type"]
    if t == "Globals":
        return "<globals>"
    if t in ("OutputHDR", "OutputFloat"):
        return t + "<>"
    return _val(n, "Name", "")


def _triangulate(n):
    """PolyMesh.init (init.go:12-134): (idxp, normalidx or None, shaderidx or None)."""
    verts = _val(n, "Verts")
    pc, fi = _val(n, "PolyCount"), _val(n, "FaceIdx")     # None: not read (nil in Go)
    ni, si = _val(n, "NormalIdx", []), _val(n, "ShaderIdx", [])
    normals = _val(n, "Normals")
    has_n = normals is not None and len(normals[2]) > 0
    idxp, nidx, sidx = [], [], []
    if pc is not None:
        base = 0
        for k, cnt in enumerate(pc):
            i = 0
            for _ in range(cnt - 2):
                i += 1
                for q in (base, base + i, base + i + 1):
                    idxp.append(fi[q])
                    if has_n:
                        nidx.append(ni[q] if ni else fi[q])
                if si:
                    sidx.append(si[k] & 0xFF)
            base += cnt
    elif fi is not None:
        for j, x in enumerate(fi):
            idxp.append(x)
            if has_n:
                nidx.append(ni[j] if ni else x)
        sidx = [x & 0xFF for x in si]
    else:
        for j in range(verts[1]):
            idxp.append(j)
            if has_n:
                nidx.append(ni[j] if ni else j)
        sidx = [x & 0xFF for x in si]
    return idxp, (nidx if has_n else None), (sidx if sidx else None)


def _blank(type_, path, line=1):
    return {"type": type_, "file": str(path), "line": line,
            "p": {name: {"kind": k, "req": r, "present": False, "value": None} for name, k, r in SCHEMAS[type_]}}


class _ObjLine:
    """lineScanner (wfobj/mtlparser.go:33-88) over bytes: token() and rest()."""

    def __init__(self, line):
        self.s, self.pos = line, 0

    def token(self):
        s, n = self.s, len(self.s)
        while self.pos < n and s[self.pos] in " \t":
            self.pos += 1
        if self.pos >= n or s[self.pos] in "#\0":
            self.pos = n
            return ""
        b = self.pos
        while self.pos < n and s[self.pos] not in " \t\0":
            self.pos += 1
        tok = s[b:self.pos]
        if self.pos < n:
            self.pos += 1
        return tok

    def rest(self):
        return self.s[self.pos:]


def _go_atoi(t):
    """strconv.Atoi: (ok, value); 0 on a syntax error, the int64 limit out of range."""
    import re
    if not re.fullmatch(r"[+-]?\d+", t):
        return False, 0
    v = int(t)
    if v > 2 ** 63 - 1:
        return False, 2 ** 63 - 1
    if v < -2 ** 63:
        return False, -2 ** 63
    return True, v


def _go_float32(t):
    """strconv.ParseFloat(t, 32): (ok, float32 value); 0 on a syntax error, +-Inf out of range.  The
    decimal value reaches float32 by way of float64 (ParseFloat rounds once; they differ only for
    values within 2^-29 relative of a float32 halfway point)."""
    import re
    body = t[1:] if t[:1] in "+-" else t
    low = body.lower()
    if low in ("inf", "infinity"):
        return True, np.float32(-np.inf if t.startswith("-") else np.inf)
    if low == "nan":
        return (t == body), (np.float32(np.nan) if t == body else np.float32(0))
    if re.fullmatch(r"0x([0-9a-f]+\.?[0-9a-f]*|\.[0-9a-f]+)p[+-]?\d+", low):
        v = float.fromhex(t)
    elif re.fullmatch(r"(\d+\.?\d*|\.\d+)(e[+-]?\d+)?", low):
        v = float(t)
    else:
        return False, np.float32(0)
    with np.errstate(over="ignore"):
        f = np.float32(v)
    return bool(np.isfinite(f)), f


def _go_clean(path):
    """path/filepath.Clean for '/'-separated paths."""
    if not path:
        return "."
    rooted = path.startswith("/")
    parts = []
    for e in path.split("/"):
        if e in ("", "."):
            continue
        if e == "..":
            if parts and parts[-1] != "..":
                parts.pop()
            elif not rooted:
                parts.append("..")
        else:
            parts.append(e)
    out = ("/" if rooted else "") + "/".join(parts)
    return out or "."


def _go_dir(path):
    k = path.rfind("/")
    return _go_clean(path[: k + 1] if k >= 0 else "")


def _go_join(a, b):
    return _go_clean(a + "/" + b if a and b else (a or b)) if (a or b) else ""


def _set_map(n, k, c):
    n["p"][k]["value"] = ("c", [float(np.float32(x)) for x in c])
    n["p"][k]["present"] = True


def _parse_mtl(path, dirname):
    """parseMTL (wfobj/mtlparser.go:93-296): ShaderStd nodes; raises Unsupported where the reference
    returns an error (its caller then drops the file's shaders) or dereferences nil."""
    with open(path, "rb") as f:
        lines = f.read().decode("latin-1").split("\n")
    out, mtl = [], None

    def need(cmd):
        if mtl is None:
            raise Unsupported(f"{cmd} before newmtl")

    for lineno, line in enumerate(lines, 1):
        if line.endswith("\r"):
            line = line[:-1]
        L = _ObjLine(line)
        cmd = L.token()
        if cmd == "newmtl":
            name = L.token()
            mtl = _blank("ShaderStd", path, lineno)
            mtl["p"]["Name"]["value"] = name
            mtl["p"]["Name"]["present"] = True
            out.append(mtl)
        elif cmd == "d":
            ok, d = _go_float32(L.token())
            if not ok:
                raise _MtlError(cmd)
            if d == 0:
                need(cmd)
                _set_map(mtl, "DiffuseColour", (0.5, 0.5, 0.5))
                _set_map(mtl, "DiffuseStrength", (0.5, 0.5, 0.5))
        elif cmd in ("Ke", "Kd", "Ks"):
            c = [_go_float32(L.token()) for _ in range(3)]
            if not c[2][0]:            # only the last value's error is tested (err is reassigned)
                raise _MtlError(cmd)
            r, g, b = (float(x[1]) for x in c)
            if cmd == "Ke":
                if r == 0 and g == 0 and b == 0:
                    continue
                need(cmd)
                st = math.sqrt(r * r + g * g + b * b)
                _set_map(mtl, "EmissionColour", (r / st, g / st, b / st))
                _set_map(mtl, "EmissionStrength", (st, st, st))
            else:
                need(cmd)
                _set_map(mtl, "DiffuseColour" if cmd == "Kd" else "Spec1Colour", (r, g, b))
        elif cmd == "Ns":
            ok, ns = _go_float32(L.token())
            if not ok:
                raise _MtlError(cmd)
            need(cmd)
            ns = np.float32(ns)    # NsToSpecRoughness (mtlparser.go:25-31); m.Max as MAXSS
            if ns <= np.float32(1000):
                r = np.float32(1) - ns / np.float32(1000)
            else:
                y = np.float32(1) - ns / np.float32(2048)
                r = np.float32(0) if np.float32(0) > y else y
            _set_map(mtl, "Spec1Roughness", (r, r, r))
        elif cmd == "illum":
            ok, m = _go_atoi(L.token())
            if not ok:
                raise _MtlError(cmd)
            if 0 <= m <= 10:
                need(cmd)
                dstr, sstr = {0: (0.5, None), 1: (0.5, None), 4: (0.2, 0.5), 5: (0.2, 0.5)}.get(m, (0.5, 0.5))
                _set_map(mtl, "DiffuseStrength", (dstr,) * 3)
                if sstr is not None:
                    _set_map(mtl, "Spec1Strength", (sstr,) * 3)
                if m == 5:
                    mtl["p"]["Spec1FresnelModel"]["value"] = "Metal"
        elif cmd == "map_Kd":
            need(cmd)
            mtl["p"]["DiffuseColour"]["value"] = ("tex", _go_join(dirname, L.rest()))
            mtl["p"]["DiffuseColour"]["present"] = True
            _set_map(mtl, "DiffuseStrength", (0.5, 0.5, 0.5))
    return out


class _MtlError(Exception):
    pass


def _parse_obj(path):
    """File.parse (wfobj/parser.go:103-263) and File.Init (wfobj.go:21-54): (the PolyMesh node named by
    the file, the ShaderStd nodes of its mtllib files in order)."""
    with open(path, "rb") as f:
        lines = f.read().decode("latin-1").split("\n")
    if lines and lines[-1] == "":
        lines.pop()
    mesh = _blank("PolyMesh", path)
    V, N, UV = [], [], []
    FI, NI, TI, PC, SI = [], [], [], [], []
    shaders, seen, mtl = [], set(), 0
    for line in lines:
        if line.endswith("\r"):
            line = line[:-1]
        L = _ObjLine(line)
        cmd = L.token()
        if cmd == "mtllib":
            mp = _go_join(_go_dir(str(path)), L.rest())
            if mp in seen:
                continue
            try:
                got = _parse_mtl(mp, _go_dir(str(path)))
            except (OSError, _MtlError):
                continue                       # logged by the reference; the file's shaders are dropped
            shaders += got
            if len(shaders) > 255:
                raise Unsupported("wfobj: too many shaders")
            seen.add(mp)
        elif cmd == "usemtl":
            name = L.rest()
            for i, sh in enumerate(shaders):
                if sh["p"]["Name"]["value"] == name:
                    mtl = i
        elif cmd == "f":                       # parseFace (parser.go:41-101)
            nv = 0
            while True:
                fld = L.token()
                if fld == "":
                    break
                parts = fld.split("/")
                p = _go_atoi(parts[0])[1]
                t = _go_atoi(parts[1])[1] if len(parts) > 1 and parts[1] else 0
                n = _go_atoi(parts[2])[1] if len(parts) > 2 and parts[2] else 0
                p = len(V) + p + 1 if p < 0 else (1 if p == 0 else p)
                t = len(UV) + t + 1 if t < 0 else (1 if t == 0 else t)
                n = len(N) + n + 1 if n < 0 else (1 if n == 0 else n)
                FI.append(wrap32(p - 1))
                NI.append(wrap32(n - 1))
                TI.append(wrap32(t - 1))
                nv += 1
            PC.append(nv)
            SI.append(mtl)
        elif cmd in ("v", "vn", "vt"):
            w = 2 if cmd == "vt" else 3
            (V if cmd == "v" else N if cmd == "vn" else UV).append(
                [float(_go_float32(L.token())[1]) for _ in range(w)])
    p = mesh["p"]

    def put(k, v):
        p[k]["value"] = v
        p[k]["present"] = True

    put("Name", str(path))
    put("IsVisible", True)
    put("Verts", (1, len(V), [x for v in V for x in v]))
    if N:
        put("Normals", (1, len(N), [x for v in N for x in v]))
    if UV:
        put("UV", (1, len(UV), [x for v in UV for x in v]))
    if FI:
        put("FaceIdx", FI)
        put("PolyCount", PC)
        put("ShaderIdx", SI)
        if N:
            put("NormalIdx", NI)
        if UV:
            put("UVIdx", TI)
    put("Shader", [sh["p"]["Name"]["value"] for sh in shaders])
    return mesh, shaders


def _light_mesh(L):
    """Tri.createMesh (triangle.go:537-567) / Disk.createMesh (disk.go:224-262) as a PolyMesh node."""
    from vermeer_amd import scenes
    if L["type"] == "TriLight":
        P0, P1, P2 = (np.array(_val(L, k)[2], np.float32) for k in ("P0", "P1", "P2"))
        N = scenes._normalize32(scenes._cross32(P1 - P0, P2 - P0))
        V = np.stack([P0, P1, P2])
        NN = np.stack([N, N, N])
    else:
        V, _, NN = scenes.disk_light_mesh(_val(L, "P")[2], _val(L, "Up")[2], _val(L, "LookAt")[2], _val(L, "Radius"),
                                          segments=_val(L, "Segments"))
    m = {"type": "PolyMesh", "file": L["file"], "line": L["line"],
         "p": {name: {"kind": k, "req": r, "present": False, "value": None} for name, k, r in SCHEMAS["PolyMesh"]}}
    m["p"]["Name"]["value"] = _val(L, "Name", "") + ":<mesh>"
    m["p"]["Shader"]["value"] = [_val(L, "Shader")]
    m["p"]["IsVisible"]["value"] = True
    m["p"]["Verts"]["value"] = (1, len(V), [float(x) for x in np.asarray(V, np.float32).reshape(-1)])
    m["p"]["Normals"]["value"] = (1, len(NN), [float(x) for x in np.asarray(NN, np.float32).reshape(-1)])
    return m


class Unsupported(Exception):
    pass


def build(path, cwd=None):
    """core.PreRender plus Render's set-up for a .vnf file, as inputs of OracleRenderer: a dict with
    desc (meshes, materials, lights), width, height, camera (aspect, tan_theta_focal, focal, radius,
    keys: LocalToWorld matrices), filter (kind, width, res, peak) or None, max_iter, nerrors, outputs."""
    import oracle
    counters = {"camera": 0}
    nodes, nerrors, log, fatal = parse_file(path, counters)
    if fatal:
        raise Unsupported("too many parse errors")
    allnodes, by_name, geoms, lights = [], {}, [], []
    state = {"filter": None, "globals": None}
    light_geom = {}
    proc_geoms = {}   # id(Proc) -> the geoms its handler made (Proc.Geom; not in the scene itself)

    def add(n):
        allnodes.append(n)
        by_name[_name(n)] = n
        t = n["type"]
        if t in ("PolyMesh", "Sphere", "GeomInstance", "Proc"):
            geoms.append(n)
        elif t in ("TriLight", "DiskLight", "SphereLight", "QuadLight"):
            lights.append(n)
        elif t in ("GaussianFilter", "AiryFilter"):
            state["filter"] = n
        elif t == "Globals":
            state["globals"] = n

    for n in nodes:
        add(n)
    g0 = state["globals"]
    width, height = (_val(g0, "XRes"), _val(g0, "YRes")) if g0 else (1024, 1024)

    def shader(name):
        s = by_name.get(name)
        return s if s is not None and s["type"] in ("ShaderStd", "DebugShader") else None

    rnd = list(allnodes)
    while rnd:
        added = []
        for n in rnd:
            t = n["type"]
            if t == "Include":
                sub, ne, lg, fa = parse_file(_val(n, "Filename"), counters)
                if fa:
                    raise Unsupported("too many parse errors in an included file")
                nerrors += ne
                for s in sub:
                    add(s)
                    added.append(s)
                continue
            if t == "QuadLight":
                raise Unsupported(t)
            if t == "GeomInstance":   # Instance.PreRender (builtin/geom/instance/instance.go:116-146)
                g = by_name.get(_val(n, "Geom"))
                if g is None or g["type"] not in ("PolyMesh", "Sphere", "GeomInstance", "Proc"):
                    raise Unsupported(f"Instance: Unable to find geom {_val(n, 'Geom')}")
                continue
            if t == "Proc" and _val(n, "Handler") == "wfobj":
                # File.Init (wfobj.go:21-54): the mesh belongs to the Proc, the .mtl shaders join the scene
                try:
                    mesh, shs = _parse_obj(_val(n, "Data"))
                except OSError:
                    raise Unsupported(f"wfobj: cannot open {_val(n, 'Data')}")
                for x in shs:
                    add(x)
                    added.append(x)
                proc_geoms[id(n)] = [mesh]
                continue
            if t == "Proc":   # Proc.PreRender (builtin/geom/proc/proc.go:140-172) with the "vnf" handler
                if _val(n, "Handler") != "vnf":
                    raise Unsupported(f"Proc handler {_val(n, 'Handler')}")
                sub, ne, lg, fa = parse_file(_val(n, "Data"), counters)
                if fa:
                    raise Unsupported("too many parse errors in a Proc's file")
                nerrors += ne
                # VNFFile.Init (builtin/geom/proc/vnf/vnf.go:19-56): the file's shaders join the scene's
                # nodes; its geoms, prerendered, belong to the Proc alone; other nodes are dropped
                for x in sub:
                    if x["type"] in ("ShaderStd", "DebugShader"):
                        add(x)
                kids = []
                for x in sub:
                    if x["type"] in ("PolyMesh", "Sphere"):
                        for s_ in (_val(x, "Shader") if x["type"] == "PolyMesh" else [_val(x, "Shader")]):
                            if shader(s_) is None:
                                raise Unsupported(f"Unable to find node (shader {s_})")
                        kids.append(x)
                    elif x["type"] in ("GeomInstance", "Proc"):
                        raise Unsupported(f"{x['type']} inside a Proc")
                if not kids:
                    raise Unsupported(f"GeomProc: {_name(n)} no geoms.")
                proc_geoms[id(n)] = kids
                continue
            if t in ("PolyMesh", "Sphere", "TriLight", "DiskLight", "SphereLight"):
                for s in (_val(n, "Shader") if t == "PolyMesh" else [_val(n, "Shader")]):
                    if shader(s) is None:
                        raise Unsupported(f"Unable to find node (shader {s})")
            if t in ("TriLight", "DiskLight"):
                m = _light_mesh(n)
                light_geom[id(n)] = m
                add(m)
                added.append(m)
            elif t == "SphereLight":
                g = {"type": "Sphere", "file": n["file"], "line": n["line"],
                     "p": {name: {"kind": k, "req": r, "present": False, "value": None} for name, k, r in SCHEMAS["Sphere"]}}
                g["p"]["Name"]["value"] = _val(n, "Name", "") + ":<geom>"
                g["p"]["P"]["value"] = _val(n, "P")
                g["p"]["Radius"]["value"] = _val(n, "Radius")
                g["p"]["Shader"]["value"] = _val(n, "Shader")
                light_geom[id(n)] = g
                add(g)
                added.append(g)
        rnd = added

    mats = []

    def material(name):
        s = shader(name)
        if s["type"] == "DebugShader":
            raise Unsupported("DebugShader")
        for i, m in enumerate(mats):
            if m is s:
                return i
        mats.append(s)
        return len(mats) - 1

    meshes, geom_index = [], {}

    def matrices(tr):
        e = tr[2]
        return np.array(e[: len(e) // 16 * 16], np.float32).reshape(-1, 16)

    def box_of(n):   # BMin[0] and BMax[0]: the instance and proc use bounds[0] only (instance.go:64-69)
        return np.array(_val(n, "BMin")[2][:3], np.float32), np.array(_val(n, "BMax")[2][:3], np.float32)

    def add_mesh(g):   # a PolyMesh or Sphere entry (PolyMesh.init, builtin/geom/polymesh/init.go:12-134)
        if g["type"] == "Sphere":
            meshes.append({"sphere": True, "centre": np.array(_val(g, "P")[2], np.float32),
                           "radius": _val(g, "Radius"), "material": material(_val(g, "Shader"))})
            return
        keys, per, ev = _val(g, "Verts")
        keys = max(keys, 1)
        idxp, nidx, sidx = _triangulate(g)
        shaders = np.array([material(s) for s in _val(g, "Shader")], np.int32)
        nv = _val(g, "Normals")
        V = np.array(ev[: keys * per * 3], np.float32).reshape(keys, per, 3)
        mesh = {"idx": np.array(idxp, np.uint32).reshape(-1, 3), "shaders": shaders,
                "shaderidx": None if sidx is None else np.array(sidx, np.uint8),
                "ray_bias": _val(g, "RayBias", 0.0)}
        if keys == 1:
            mesh["verts"] = V[0]
            mesh["normals"] = None
            if nidx is not None:
                mesh["normal_list"] = np.array(nv[2], np.float32).reshape(-1, 3)
                mesh["normal_idx"] = np.array(nidx, np.uint32)
        else:
            mesh["keys"] = keys
            mesh["verts"] = V
            mesh["reference_indexing"] = True
            mesh["normals"] = None
            if nv is not None and nv[2]:
                nk, nper = max(nv[0], 1), nv[1]
                arr = np.array(nv[2], np.float32).reshape(nk, nper, 3)[:, : len(idxp)]
                mesh["normals"] = arr.reshape(nk, -1, 3) if nv[0] == keys else None
        tr = _val(g, "Transform")
        if tr is not None:   # PolyMesh.Trace moves rays whenever Transform is given (trace.go:22)
            mesh["transform"] = matrices(tr)
        meshes.append(mesh)

    for g in geoms:
        if g["type"] == "GeomInstance":
            tgt = by_name[_val(g, "Geom")]
            if id(tgt) not in geom_index:
                raise Unsupported("GeomInstance of a geom that follows it")
            if tgt["type"] not in ("PolyMesh", "Sphere"):
                raise Unsupported("GeomInstance of an instance or proc")
            lo, hi = box_of(g)
            geom_index[id(g)] = len(meshes)
            meshes.append({"instance": {"geom": geom_index[id(tgt)], "matrices": matrices(_val(g, "Transform")),
                                        "bmin": lo, "bmax": hi}})
            continue
        if g["type"] == "Proc":   # its geoms are held by the Proc only: hidden from the top level
            kids = []
            for x in proc_geoms[id(g)]:
                kids.append(len(meshes))
                add_mesh(x)
                meshes[-1]["hidden"] = True
            lo, hi = box_of(g)
            tr = _val(g, "Transform")
            geom_index[id(g)] = len(meshes)
            meshes.append({"proc": {"geoms": kids, "matrices": None if tr is None else matrices(tr),
                                    "bmin": lo, "bmax": hi}})
            continue
        geom_index[id(g)] = len(meshes)
        add_mesh(g)

    from vermeer_amd import scenes
    from vermeer_amd._lib import MATERIAL_DTYPE
    mtab = np.zeros(max(len(mats), 1), MATERIAL_DTYPE)
    for i, s in enumerate(mats):
        def mp(k):
            v = _val(s, k)
            if v is None:
                return None
            if v[0] == "tex":
                raise Unsupported("texture")
            return v[1]
        for key, rec, one in (("EmissionColour", "emission", False), ("EmissionStrength", "emission_strength", True),
                              ("DiffuseColour", "diffuse", False), ("DiffuseStrength", "diffuse_strength", True),
                              ("DiffuseRoughness", "diffuse_roughness", True), ("Spec1Colour", "spec1", False),
                              ("Spec1Strength", "spec1_strength", True), ("Spec1Roughness", "spec1_roughness", True),
                              ("IOR", "ior", True), ("Spec1FresnelRefl", "spec1_fresnel_refl", False),
                              ("Spec1FresnelEdge", "spec1_fresnel_edge", False)):
            c = mp(key)
            if c is None:
                continue
            mtab[i][rec] = c[0] if one else c
            flag = {"emission": "has_emission_colour", "emission_strength": "has_emission_strength",
                    "diffuse": "has_diffuse_colour", "spec1": "has_spec1_colour"}.get(rec, "has_" + rec)
            mtab[i][flag] = 1
        fm = _val(s, "Spec1FresnelModel", "")
        mtab[i]["fresnel_model"] = 2 if fm == "Metal" else 1 if fm == "Dielectric" else 0

    lrec = []
    for L in lights:
        d = {"material": material(_val(L, "Shader")), "samples": _val(L, "Samples"),
             "geom": geom_index[id(light_geom[id(L)])]}
        if L["type"] == "TriLight":
            d.update(type="tri", P0=np.array(_val(L, "P0")[2], np.float32), P1=np.array(_val(L, "P1")[2], np.float32),
                     P2=np.array(_val(L, "P2")[2], np.float32))
        elif L["type"] == "DiskLight":
            d.update(type="disk", P=np.array(_val(L, "P")[2], np.float32), Up=np.array(_val(L, "Up")[2], np.float32),
                     LookAt=np.array(_val(L, "LookAt")[2], np.float32), radius=_val(L, "Radius"))
        else:
            d.update(type="sphere", P=np.array(_val(L, "P")[2], np.float32), radius=_val(L, "Radius"))
        lrec.append(d)

    gl = state["globals"]
    cam_name = "camera"
    max_iter = 16
    if gl is not None:
        if _val(gl, "Camera"):
            cam_name = _val(gl, "Camera")
        max_iter = _val(gl, "MaxIter", 0)
    cam = by_name.get(cam_name)
    if cam is None or cam["type"] != "Camera":
        raise Unsupported("no camera")
    aspect = f32(_val(cam, "Aspect", 0.0))
    if aspect == 0.0:
        aspect = f32(np.float32(width) / np.float32(height))
    fov, focal = np.float32(_val(cam, "Fov")), np.float32(_val(cam, "Focal"))
    ang = (fov / np.float32(2)) * np.float32(3.14159265358) / np.float32(180)
    ttf = f32(np.float32(math.tan(float(ang))) * focal)
    keys_out = []
    ltw = _val(cam, "LocalToWorld")
    if ltw is not None:
        e = ltw[2]
        keys_out += [e[k:k + 16] for k in range(0, len(e) // 16 * 16, 16)]
    if _val(cam, "Type") == "LookAt":
        fr, to, ro = _val(cam, "From"), _val(cam, "To"), _val(cam, "Roll")
        fk = np.array(fr[2], np.float32).reshape(-1, 3)[: fr[0]]
        tk = np.array(to[2], np.float32).reshape(-1, 3)[: to[0]]
        rk = None if ro is None or not ro[2] else np.array(ro[2], np.float32)[: ro[0]]
        mats_k = oracle.camera_lookat_keys(fk, tk, np.array(_val(cam, "Up")[2], np.float32), rk)
        keys_out += [list(np.asarray(m, np.float32).reshape(16)) for m in mats_k]
    else:
        raise Unsupported("matrix camera")
    filt = None
    fn = state["filter"]
    if fn is not None:
        if fn["type"] == "GaussianFilter":
            filt = ("gaussian", _val(fn, "Width"), _val(fn, "Res"), 0.0)
        else:
            filt = ("airy", _val(fn, "Width"), _val(fn, "Res"), _val(fn, "Peak"))
    outputs = [("hdr" if n["type"] == "OutputHDR" else "float", _val(n, "Filename"))
               for n in allnodes if n["type"] in ("OutputHDR", "OutputFloat")]
    return {"desc": {"meshes": meshes, "materials": mtab, "lights": lrec}, "width": width, "height": height,
            "camera": {"aspect": aspect, "tan_theta_focal": ttf, "focal": f32(focal), "radius": f32(_val(cam, "Radius", 0.0)),
                       "keys": np.array(keys_out, np.float32)},
            "filter": filt, "max_iter": max_iter, "nerrors": nerrors, "outputs": outputs}


def renderer(built):
    """An OracleRenderer for build()'s result."""
    import oracle
    c = built["camera"]
    r = oracle.OracleRenderer(built["desc"], built["width"], built["height"], np.eye(4, dtype=np.float32).reshape(-1),
                              c["tan_theta_focal"], c["focal"], aspect=c["aspect"])
    r.set_camera_keys(c["keys"])
    if c["radius"] > 0:
        r.set_lens(c["radius"])
    if built["filter"] is not None:
        kind, w, res, peak = built["filter"]
        r.set_filter(kind, w, res, peak)
    return r
