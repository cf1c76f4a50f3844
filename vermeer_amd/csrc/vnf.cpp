// .vnf scene files: parsing (nodes/lex.go, nodes/parser.go) and the PreRender that turns the nodes
// into a GPU scene and renderer (core/core.go:38-89, the nodes' PreRender methods).
//
// The lexer and parser keep the reference's behaviour, quirks included, because they decide what a
// file means:
//   - '#' discards the rest of its line; characters that start no token (and '\r') are skipped;
//   - a number runs over digits, '-', '.', 'e' and 'E' and is a float if it holds '.', 'e' or 'E'; a
//     run Go's strconv rejects ("1e+3" stops at '+', "1-2", "1e") ends the parse like end of file;
//   - a token is ended by whitespace or one of " [ ] { } , (which is read again); end of file inside a
//     token ends the parse;
//   - a parameter whose value has the wrong token type is consumed and left at its default, but
//     counts as present; an unknown parameter skips to the next token that is not a type keyword;
//   - a node missing a required parameter, or repeating one, is dropped, and the parser then skips to
//     the next '}', which is the end of the following node (the reference does this too);
//   - more than 10 errors stop the reference (os.Exit); here the load fails.
// Every node type of the reference is recognised by name.  What this library does not render fails
// the build with an error: DebugShader, QuadLight, textures, instances of instances or procs, an
// instance placed before its geom, instances and procs inside a Proc's file.
#include <cctype>
#include <cerrno>
#include <climits>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <memory>
#include <sstream>
#include <string>
#include <vector>

#include "scene_internal.h"

using vgpu::fail;

namespace vnf_detail {

// ------------------------------------------------------------------ lexer (nodes/lex.go)

enum Tok { kEof = 0, kToken, kString, kFloat, kInt, kOpenBrace, kCloseBrace, kOpenCurly, kCloseCurly, kComma };

struct Sym {
    double f = 0.0;
    int64_t i = 0;
    std::string s;
};

bool is_alpha(uint32_t c) { return c == '_' || (c < 128 && std::isalpha((int)c)); }

struct Lexer {
    std::istream* in = nullptr;
    std::string line;      // rest of the current line
    size_t pos = 0;
    uint32_t peek = 0;     // a character read back (0: none)
    bool peek_token = false;
    Sym psym;
    int ptoken = kEof;
    int line_number = 0;

    // readLine (lex.go:214-226): false at end of input, which leaves the current line as it was (so
    // the rest of a comment on the last line is read after the end-of-file token, as in the reference)
    bool read_line() {
        line_number++;
        std::string next_line;
        if (!std::getline(*in, next_line)) return false;
        if (!in->eof()) next_line.push_back('\n');   // a last line without '\n' is still returned
        else if (next_line.empty()) return false;
        line.swap(next_line);
        pos = 0;
        return true;
    }

    uint32_t next() {   // next (lex.go:229-261)
        if (peek != 0) { const uint32_t r = peek; peek = 0; return r; }
        while (pos >= line.size()) {
            if (!read_line()) return 0;
        }
        const unsigned char c = (unsigned char)line[pos++];
        if (c == '#') {
            if (!read_line()) return 0;
            return '\n';
        }
        return c;
    }

    int lex_str(Sym& y) {
        std::string b;
        for (;;) {
            const uint32_t c = next();
            if (c == 0) return kEof;
            if (c == '"') break;
            b.push_back((char)c);
        }
        y.s = b;
        return kString;
    }

    int lex_token(uint32_t c, Sym& y) {
        std::string b(1, (char)c);
        for (;;) {
            c = next();
            if (c == 0) return kEof;
            if (c == '"' || c == '[' || c == ']' || c == '{' || c == '}' || c == ',') { peek = c; break; }
            if (c == ' ' || c == '\t' || c == '\r' || c == '\n') break;
            b.push_back((char)c);
        }
        y.s = b;
        return kToken;
    }

    int lex_num(uint32_t c, Sym& y) {
        std::string b(1, (char)c);
        bool is_float = false;
        for (;;) {
            c = next();
            if ((c >= '0' && c <= '9') || c == '-') { b.push_back((char)c); continue; }
            if (c == '.' || c == 'e' || c == 'E') { is_float = true; b.push_back((char)c); continue; }
            break;
        }
        if (c != 0) peek = c;
        if (is_float) {
            double f;
            if (!parse_go_float(b, f)) return kEof;
            y.f = f;
            return kFloat;
        }
        int64_t v;
        if (!parse_go_int(b, v)) return kEof;
        y.i = v;
        return kInt;
    }

    // strconv.ParseInt(s, 10, 64) on the lexer's alphabet: an optional '-', then decimal digits that fit.
    static bool parse_go_int(const std::string& s, int64_t& v) {
        size_t k = s[0] == '-' ? 1 : 0;
        if (k == s.size()) return false;
        for (size_t j = k; j < s.size(); j++) if (s[j] < '0' || s[j] > '9') return false;
        errno = 0;
        char* end = nullptr;
        const long long r = std::strtoll(s.c_str(), &end, 10);
        if (errno == ERANGE || *end) return false;
        v = r;
        return true;
    }

    // strconv.ParseFloat(s, 64) on the lexer's alphabet: -?(d+.?d*|.d+)([eE]-?d+)?, correctly rounded.
    static bool parse_go_float(const std::string& s, double& v) {
        size_t k = s[0] == '-' ? 1 : 0;
        size_t digits = 0;
        while (k < s.size() && std::isdigit((unsigned char)s[k])) { k++; digits++; }
        if (k < s.size() && s[k] == '.') {
            k++;
            while (k < s.size() && std::isdigit((unsigned char)s[k])) { k++; digits++; }
        }
        if (digits == 0) return false;
        if (k < s.size() && (s[k] == 'e' || s[k] == 'E')) {
            k++;
            if (k < s.size() && s[k] == '-') k++;
            size_t ed = 0;
            while (k < s.size() && std::isdigit((unsigned char)s[k])) { k++; ed++; }
            if (ed == 0) return false;
        }
        if (k != s.size()) return false;
        errno = 0;
        v = std::strtod(s.c_str(), nullptr);   // overflow: Go returns +-Inf with an error, the lexer EOF
        if (errno == ERANGE && std::isinf(v)) return false;
        return true;
    }

    int lex(Sym& y) {   // lex (lex.go:73-107)
        for (;;) {
            const uint32_t c = next();
            if (is_alpha(c)) return lex_token(c, y);
            switch (c) {
                case 0: return kEof;
                case '-': case '0': case '1': case '2': case '3': case '4': case '5': case '6': case '7': case '8':
                case '9': return lex_num(c, y);
                case '[': return kOpenBrace;
                case ']': return kCloseBrace;
                case '{': return kOpenCurly;
                case '}': return kCloseCurly;
                case ',': return kComma;
                case '"': return lex_str(y);
                default: break;   // whitespace and every other character
            }
        }
    }

    int Lex(Sym& y) {
        if (peek_token) { peek_token = false; y = psym; return ptoken; }
        return lex(y);
    }
    int Peek(Sym& y) {
        const int t = lex(y);
        peek_token = true;
        psym = y;
        ptoken = t;
        return t;
    }
    void Skip() { Sym y; Lex(y); }
};

// ------------------------------------------------------------------ node schemas (the Go structs)

enum Kind {
    KInt, KFloat, KBool, KString, KInt32Slice, KStringSlice, KMap /* RGBUniform / Float32Uniform */, KVec3,
    KPointArray, KVec3Array, KVec2Array, KFloat32Array, KMatrixArray, KStruct /* core.NodeDef: unparseable */
};

struct FieldDef { const char* name; Kind kind; bool required; };

struct Value {
    Kind kind = KInt;
    bool present = false;
    int64_t i = 0;
    double f = 0.0;
    bool b = false;
    std::string s;
    std::vector<int32_t> ints;
    std::vector<std::string> strs;
    bool slice_set = false;   // a slice was read (Go: non-nil, possibly empty)
    bool map_set = false;   // a map (rgb / float / rgbtex) was given
    bool map_tex = false;   // rgbtex: a texture file
    float c[3] = {0, 0, 0};
    int motion_keys = 0, per_key = 0;
    std::vector<float> elems;   // arrays: 3 floats per point / vec3, 2 per vec2, 1 per float, 16 per matrix
};

struct NodeType { const char* name; std::vector<FieldDef> fields; };

const std::vector<NodeType>& node_types() {
    static const std::vector<NodeType> t = {
        {"Globals", {{"NodeDef", KStruct, false}, {"XRes", KInt, false}, {"YRes", KInt, false},
                     {"UseProgress", KBool, false}, {"MaxGoRoutines", KInt, false}, {"Camera", KString, false},
                     {"MaxIter", KInt, false}, {"Output", KString, false}}},
        {"PolyMesh", {{"Name", KString, true}, {"RayBias", KFloat, false}, {"Verts", KPointArray, true},
                      {"PolyCount", KInt32Slice, false}, {"FaceIdx", KInt32Slice, false}, {"Shader", KStringSlice, true},
                      {"ShaderIdx", KInt32Slice, false}, {"CalcNormals", KBool, false}, {"IsVisible", KBool, false},
                      {"Transform", KMatrixArray, false}, {"UV", KVec2Array, false}, {"UVIdx", KInt32Slice, false},
                      {"Normals", KVec3Array, false}, {"NormalIdx", KInt32Slice, false}}},
        {"Camera", {{"Name", KString, true}, {"Type", KString, true}, {"From", KPointArray, true},
                    {"To", KPointArray, true}, {"Roll", KFloat32Array, true}, {"Up", KVec3, true},
                    {"Aspect", KFloat, false}, {"Fov", KFloat, false}, {"Focal", KFloat, false},
                    {"WorldToLocal", KMatrixArray, false}, {"LocalToWorld", KMatrixArray, false},
                    {"L", KFloat, false}, {"R", KFloat, false}, {"T", KFloat, false}, {"B", KFloat, false},
                    {"Radius", KFloat, false}}},
        {"ShaderStd", {{"Name", KString, true}, {"EmissionColour", KMap, false}, {"EmissionStrength", KMap, false},
                       {"Sides", KInt, false}, {"DiffuseColour", KMap, false}, {"DiffuseStrength", KMap, false},
                       {"DiffuseRoughness", KMap, false}, {"Spec1Colour", KMap, false}, {"Spec1Strength", KMap, false},
                       {"Spec1Roughness", KMap, false}, {"Spec1FresnelModel", KString, false},
                       {"Spec1FresnelRefl", KMap, false}, {"Spec1FresnelEdge", KMap, false}, {"IOR", KMap, false}}},
        {"DebugShader", {{"Name", KString, true}, {"Sides", KInt, false}, {"Colour", KMap, true}}},
        {"TriLight", {{"Name", KString, true}, {"P0", KVec3, true}, {"P1", KVec3, true}, {"P2", KVec3, true},
                      {"Shader", KString, true}, {"Samples", KInt, true}}},
        {"DiskLight", {{"Name", KString, true}, {"P", KVec3, true}, {"Up", KVec3, true}, {"LookAt", KVec3, true},
                       {"Radius", KFloat, true}, {"Shader", KString, true}, {"Segments", KInt, false},
                       {"Samples", KInt, false}}},
        {"SphereLight", {{"Name", KString, true}, {"P", KVec3, true}, {"Radius", KFloat, true},
                         {"Shader", KString, true}, {"Samples", KInt, false}}},
        {"QuadLight", {{"Name", KString, true}, {"P", KVec3, true}, {"U", KVec3, true}, {"V", KVec3, true},
                       {"Shader", KString, true}, {"Samples", KInt, false}}},
        {"Sphere", {{"Name", KString, true}, {"RayBias", KFloat, false}, {"P", KVec3, true}, {"Radius", KFloat, true},
                    {"Shader", KString, true}}},
        {"GaussianFilter", {{"NodeDef", KStruct, true}, {"Name", KString, true}, {"Width", KFloat, true},
                            {"Res", KInt, true}}},
        {"AiryFilter", {{"Name", KString, true}, {"Width", KFloat, false}, {"Res", KInt, false}, {"Peak", KFloat, false}}},
        {"OutputHDR", {{"Filename", KString, true}}},
        {"OutputFloat", {{"Filename", KString, true}}},
        {"Include", {{"NodeDef", KStruct, false}, {"Name", KString, false}, {"Filename", KString, true}}},
        {"Proc", {{"Name", KString, true}, {"Handler", KString, true}, {"Data", KString, true},
                  {"BMin", KPointArray, true}, {"BMax", KPointArray, true}, {"Transform", KMatrixArray, false}}},
        {"GeomInstance", {{"Name", KString, true}, {"Geom", KString, true}, {"BMin", KPointArray, true},
                          {"BMax", KPointArray, true}, {"Transform", KMatrixArray, true}}},
    };
    return t;
}

const NodeType* find_type(const std::string& name) {
    for (const NodeType& t : node_types()) if (name == t.name) return &t;
    return nullptr;
}

// Constructor defaults of the registered node types (their nodes.Register functions).
void set_defaults(const std::string& type, std::map<std::string, Value>& p, int camera_count) {
    auto I = [&](const char* k, int64_t v) { p[k].i = v; };
    auto F = [&](const char* k, double v) { p[k].f = v; };
    if (type == "Globals") { I("XRes", 256); I("YRes", 256); I("MaxGoRoutines", 5); }
    else if (type == "PolyMesh") p["IsVisible"].b = true;
    else if (type == "Camera") { F("Focal", 12); F("Fov", 90); p["Name"].s = "camera<" + std::to_string(camera_count) + ">"; }
    else if (type == "DiskLight") { I("Segments", 20); I("Samples", 1); }
    else if (type == "SphereLight") { F("Radius", 1); I("Samples", 1); }
    else if (type == "QuadLight") I("Samples", 1);
    else if (type == "TriLight") I("Samples", 1);
    else if (type == "Sphere") F("Radius", 1);
    else if (type == "GaussianFilter") { I("Res", 17); F("Width", 2); }
    else if (type == "AiryFilter") { I("Res", 49); F("Width", 6); F("Peak", 4); }
    else if (type == "OutputHDR") p["Filename"].s = "out.hdr";
    else if (type == "OutputFloat") p["Filename"].s = "out.float";
}

struct Node {
    std::string type;
    std::map<std::string, Value> p;
    std::string file;
    int line = 0;

    const Value* get(const char* k) const {
        auto it = p.find(k);
        return it == p.end() ? nullptr : &it->second;
    }
    int64_t i(const char* k) const { const Value* v = get(k); return v ? v->i : 0; }
    double f(const char* k) const { const Value* v = get(k); return v ? v->f : 0.0; }
    bool b(const char* k) const { const Value* v = get(k); return v && v->b; }
    std::string s(const char* k) const { const Value* v = get(k); return v ? v->s : std::string(); }
    bool has(const char* k) const { const Value* v = get(k); return v && v->present; }
    // core.Node.Name() of each type
    std::string name() const {
        if (type == "Globals") return "<globals>";
        if (type == "OutputHDR") return "OutputHDR<>";
        if (type == "OutputFloat") return "OutputFloat<>";
        return s("Name");
    }
};

// ------------------------------------------------------------------ parser (nodes/parser.go)

struct Parser {
    Lexer lex;
    std::string filename;
    int nerrors = 0;
    std::string* log = nullptr;
    bool fatal = false;   // more than 10 errors (the reference exits)
    int* camera_count = nullptr;

    void errorf(const char* fmt, ...) {
        char buf[512];
        va_list ap;
        va_start(ap, fmt);
        vsnprintf(buf, sizeof(buf), fmt, ap);
        va_end(ap);
        if (log) *log += filename + ":" + std::to_string(lex.line_number) + ": " + buf + "\n";
        if (++nerrors > 10) fatal = true;
    }

    static bool is_keyword(const std::string& s) {
        static const char* kw[] = {"int", "float", "vec2", "vec3", "point", "rgb", "rgbtex", "matrix"};
        for (const char* k : kw) if (s == k) return true;
        return false;
    }

    // int64(f) on amd64 (CVTTSD2SQ): truncation, and 0x8000000000000000 for NaN or out of range
    static int64_t go_int64(double f) {
        return (f >= -9223372036854775808.0 && f < 9223372036854775808.0) ? (int64_t)f : INT64_MIN;
    }

    bool number(float& out) {   // TokInt or TokFloat as float32
        Sym y;
        const int t = lex.Lex(y);
        if (t == kInt) { out = (float)y.i; return true; }
        if (t == kFloat) { out = (float)y.f; return true; }
        return false;
    }

    // int32slice / stringslice (parser.go:124-186): count, type token, elements.
    bool slice(Value& v, bool strings) {
        Sym y;
        if (lex.Lex(y) != kInt) return false;
        const int64_t count = y.i;
        if (lex.Lex(y) != kToken && y.s != (strings ? "string" : "int")) return false;   // the type token
        std::vector<int32_t> ints;
        std::vector<std::string> strs;
        for (int64_t i = 0; i < count; i++) {
            const int t = lex.Lex(y);
            if (strings) { if (t != kString) return false; strs.push_back(y.s); }
            else { if (t != kInt) return false; ints.push_back((int32_t)y.i); }
        }
        v.ints = ints;
        v.strs = strs;
        v.slice_set = true;
        return true;
    }

    // rgb / floatmap / rgbtex (parser.go:188-262)
    bool map_value(Value& v, const std::string& what) {
        Sym y;
        lex.Lex(y);   // "rgb" / "float" / "rgbtex"
        if (what == "rgbtex") {
            if (lex.Lex(y) != kString) return false;
            v.map_set = true;
            v.map_tex = true;
            v.s = y.s;
            return true;
        }
        float c[3];
        const int n = what == "rgb" ? 3 : 1;
        for (int k = 0; k < n; k++) if (!number(c[k])) return false;
        if (n == 1) c[1] = c[2] = c[0];
        std::memcpy(v.c, c, sizeof(c));
        v.map_set = true;
        v.map_tex = false;
        return true;
    }

    // pointarray / vec3array / vec2array / float32array / matrixarray (parser.go:287-545)
    bool array(Value& v, int width, bool matrix, const char* type) {
        Sym y;
        if (lex.Lex(y) != kInt) return false;
        const int keys = (int)y.i;
        int per = 1;
        if (!matrix) {
            if (lex.Lex(y) != kInt) return false;
            per = (int)y.i;
        }
        if (lex.Lex(y) != kToken && y.s != type) return false;   // the element type token
        const int k = keys == 0 ? 1 : keys;
        std::vector<float> e;
        for (int64_t j = 0; j < (int64_t)k * per; j++) {
            float m[16];
            for (int c = 0; c < width; c++) if (!number(m[c])) return false;
            if (matrix) {   // Matrix4Transpose: the file lists rows
                for (int r = 0; r < 4; r++) for (int c = 0; c < 4; c++) e.push_back(m[c * 4 + r]);
            } else {
                e.insert(e.end(), m, m + width);
            }
        }
        v.motion_keys = keys;
        v.per_key = per;
        v.elems = e;
        return true;
    }

    bool param(Value& v) {   // parseParam (parser.go:581-688); false: "Error parsing field"
        Sym y;
        switch (v.kind) {
            case KInt: {
                const int t = lex.Lex(y);
                if (t == kFloat) v.i = go_int64(y.f);
                else if (t == kInt) v.i = y.i;
                return true;
            }
            case KFloat: {
                const int t = lex.Lex(y);
                if (t == kFloat) v.f = y.f;
                else if (t == kInt) v.f = (double)y.i;
                return true;
            }
            case KBool:
                if (lex.Lex(y) == kInt) v.b = y.i != 0;
                return true;
            case KString:
                if (lex.Lex(y) == kString) v.s = y.s;
                return true;
            case KInt32Slice:
            case KStringSlice:
                // the reference ignores what int32slice / stringslice return: a bad slice is silent
                if (lex.Peek(y) == kInt) (void)slice(v, v.kind == KStringSlice);
                else { errorf("Invalid token for param (expecting length of slice)"); lex.Skip(); }
                return true;
            case KMap:
                // likewise for rgb / floatmap / rgbtex; any other token is left for the node loop
                if (lex.Peek(y) == kToken && (y.s == "rgb" || y.s == "float" || y.s == "rgbtex")) (void)map_value(v, y.s);
                return true;
            case KVec3: {
                float c[3];
                for (int k = 0; k < 3; k++) if (!number(c[k])) return false;
                v.elems.assign(c, c + 3);
                return true;
            }
            case KPointArray: return array(v, 3, false, "point");
            case KVec3Array: return array(v, 3, false, "vec3");
            case KVec2Array: return array(v, 2, false, "vec2");
            case KFloat32Array: return array(v, 1, false, "float");
            case KMatrixArray: return array(v, 16, true, "matrix");
            case KStruct:
                errorf("Invalid type for param (core.NodeDef)");
                lex.Skip();
                return true;
        }
        return true;
    }

    void skip_to_token() {   // skipToToken (parser.go:762-774)
        Sym y;
        for (;;) {
            const int t = lex.Peek(y);
            if (t == kCloseCurly || (t == kToken && !is_keyword(y.s)) || t == kEof) return;
            lex.Skip();
        }
    }

    // node (parser.go:776-849): nullptr when the node is dropped
    std::unique_ptr<Node> node(const std::string& type) {
        const NodeType* nt = find_type(type);
        if (!nt) return nullptr;
        auto n = std::make_unique<Node>();
        n->type = type;
        n->file = filename;
        n->line = lex.line_number;
        for (const FieldDef& fd : nt->fields) n->p[fd.name].kind = fd.kind;
        set_defaults(type, n->p, type == "Camera" ? (*camera_count)++ : 0);
        Sym y;
        for (;;) {
            if (fatal) return nullptr;
            const int t = lex.Lex(y);
            if (t == kToken) {
                const std::string pname = y.s;
                const FieldDef* fd = nullptr;
                for (const FieldDef& f : nt->fields) if (pname == f.name) fd = &f;
                if (!fd) {
                    errorf("Field \"%s\" not found in node %s", pname.c_str(), type.c_str());
                    skip_to_token();
                    continue;
                }
                Value& v = n->p[pname];
                if (v.present) {
                    errorf("Field %s already found in %s", pname.c_str(), type.c_str());
                    return nullptr;
                }
                if (!param(v)) errorf("Error parsing field %s", pname.c_str());
                v.present = true;
            } else if (t == kCloseCurly) {
                for (const FieldDef& f : nt->fields)
                    if (f.required && !n->p[f.name].present) {
                        errorf("node: required field %s not found in %s", f.name, type.c_str());
                        return nullptr;
                    }
                return n;
            } else if (t == kEof) {
                // the reference reports "invalid token" for ever here, until its error limit exits
                errorf("node: end of file inside node %s", type.c_str());
                fatal = true;
                return nullptr;
            } else {
                errorf("node: Parse error, invalid token in object \"%d\"", t);
            }
        }
    }

    // parse (parser.go:864-903)
    void parse(std::vector<std::unique_ptr<Node>>& out) {
        Sym y;
        for (;;) {
            if (fatal) return;
            const int t = lex.Lex(y);
            if (t != kToken) return;
            const std::string type = y.s;
            if (lex.Lex(y) != kOpenCurly) errorf("Invalid token in node preamble");
            std::unique_ptr<Node> n = node(type);
            if (!n) {
                if (!find_type(type)) errorf("Node is nil: Node type %s not registered.", type.c_str());
                else errorf("Node is nil: <nil>");
                // skip until the next closing brace (the reference loops for ever at end of file)
                for (;;) {
                    const int u = lex.Lex(y);
                    if (u == kCloseCurly || u == kEof) break;
                }
                continue;
            }
            out.push_back(std::move(n));
        }
    }
};

}  // namespace vnf_detail

using namespace vnf_detail;

// ------------------------------------------------------------------ loaded file and PreRender

struct vgpu_vnf {
    std::vector<std::unique_ptr<Node>> nodes;   // core.AddNode order (included files are added at build)
    std::string log;
    int nerrors = 0;
    bool fatal = false;
    int camera_count = 0;
    uint32_t width = 1024, height = 1024;       // the framebuffer PreRender creates (defaultGlobals)
    int32_t max_iter = 16;
    bool built = false;
    std::vector<std::pair<int32_t, std::string>> outputs;   // (VGPU_OUTPUT_*, filename)
};

namespace {

int parse_file(vgpu_vnf* v, const std::string& path, std::vector<std::unique_ptr<Node>>& out) {
    std::ifstream f(path, std::ios::binary);
    if (!f) return fail(VGPU_ERR_ARG, "cannot open %s", path.c_str());
    Parser p;
    p.lex.in = &f;
    p.filename = path;
    p.log = &v->log;
    p.camera_count = &v->camera_count;
    p.parse(out);
    v->nerrors += p.nerrors;
    if (p.fatal) {
        v->fatal = true;
        return fail(VGPU_ERR_ARG, "%s: too many errors (the reference exits here):\n%s", path.c_str(), v->log.c_str());
    }
    return VGPU_OK;
}

// ------------------------------------------------------------------ Wavefront .obj / .mtl (the wfobj Proc handler)
//
// builtin/geom/proc/wfobj/parser.go and mtlparser.go: one PolyMesh with a polygon per "f" line and a
// ShaderStd per "newmtl".  Lines are split as lineScanner does: a token is a run of bytes other than
// space and tab, '#' before a token ends the line, Rest() is what follows the command's first
// separator.  Bytes are read as bytes (the reference decodes UTF-8 and drops invalid bytes).

struct ObjLine {
    std::string s;
    size_t pos = 0;
    std::string token() {
        while (pos < s.size() && (s[pos] == ' ' || s[pos] == '\t')) pos++;
        if (pos >= s.size() || s[pos] == '#' || s[pos] == '\0') { pos = s.size(); return std::string(); }
        const size_t b = pos;
        while (pos < s.size() && s[pos] != ' ' && s[pos] != '\t' && s[pos] != '\0') pos++;
        std::string t = s.substr(b, pos - b);
        if (pos < s.size()) pos++;   // the separator is consumed with the token
        return t;
    }
    std::string rest() const { return pos < s.size() ? s.substr(pos) : std::string(); }
};

// strconv.Atoi: [+-]?digits; 0 (with an error) otherwise; out of range: the int64 limit.
bool go_atoi(const std::string& t, int64_t& out) {
    out = 0;
    size_t i = 0;
    bool neg = false;
    if (i < t.size() && (t[i] == '+' || t[i] == '-')) neg = t[i++] == '-';
    if (i >= t.size()) return false;
    uint64_t v = 0;
    bool over = false;
    for (; i < t.size(); i++) {
        if (t[i] < '0' || t[i] > '9') return false;
        const uint64_t d = (uint64_t)(t[i] - '0');
        if (v > (UINT64_MAX - d) / 10) over = true;
        else v = v * 10 + d;
    }
    const uint64_t lim = neg ? (uint64_t)INT64_MAX + 1 : (uint64_t)INT64_MAX;
    if (over || v > lim) { out = neg ? INT64_MIN : INT64_MAX; return false; }
    out = neg ? (int64_t)(0 - v) : (int64_t)v;
    return true;
}

// strconv.ParseFloat(t, 32): decimal, hexadecimal with a p exponent, inf / infinity / nan; the
// float32 nearest (correctly rounded) as a double; a syntax error gives 0, overflow +-Inf.
bool go_parse_float32(const std::string& t, double& out) {
    out = 0.0;
    size_t i = 0;
    if (i < t.size() && (t[i] == '+' || t[i] == '-')) i++;
    std::string body = t.substr(i);
    std::string low = body;
    for (char& c : low) c = (char)std::tolower((unsigned char)c);
    bool ok = false;
    if (low == "inf" || low == "infinity") ok = true;
    else if (low == "nan") ok = i == 0;
    else if (low.size() > 2 && low[0] == '0' && low[1] == 'x') {
        size_t k = 2, nd = 0;
        while (k < low.size() && std::isxdigit((unsigned char)low[k])) { k++; nd++; }
        if (k < low.size() && low[k] == '.') { k++; while (k < low.size() && std::isxdigit((unsigned char)low[k])) { k++; nd++; } }
        if (nd > 0 && k < low.size() && low[k] == 'p') {
            k++;
            if (k < low.size() && (low[k] == '+' || low[k] == '-')) k++;
            size_t ne = 0;
            while (k < low.size() && std::isdigit((unsigned char)low[k])) { k++; ne++; }
            ok = ne > 0 && k == low.size();
        }
    } else {
        size_t k = 0, nd = 0;
        while (k < low.size() && std::isdigit((unsigned char)low[k])) { k++; nd++; }
        if (k < low.size() && low[k] == '.') { k++; while (k < low.size() && std::isdigit((unsigned char)low[k])) { k++; nd++; } }
        if (nd > 0 && k < low.size() && low[k] == 'e') {
            size_t k0 = k++;
            if (k < low.size() && (low[k] == '+' || low[k] == '-')) k++;
            size_t ne = 0;
            while (k < low.size() && std::isdigit((unsigned char)low[k])) { k++; ne++; }
            if (ne == 0) k = k0;
        }
        ok = nd > 0 && k == low.size();
    }
    if (!ok) return false;
    out = (double)std::strtof(t.c_str(), nullptr);
    return !std::isinf(out) || low.rfind("inf", 0) == 0;
}

// path/filepath.Clean, Dir and Join for '/'-separated paths.
std::string go_clean(const std::string& path) {
    if (path.empty()) return ".";
    const bool rooted = path[0] == '/';
    std::vector<std::string> parts;
    size_t i = 0;
    while (i <= path.size()) {
        size_t j = path.find('/', i);
        if (j == std::string::npos) j = path.size();
        const std::string e = path.substr(i, j - i);
        if (e.empty() || e == ".") {
        } else if (e == "..") {
            if (!parts.empty() && parts.back() != "..") parts.pop_back();
            else if (!rooted) parts.push_back("..");
        } else {
            parts.push_back(e);
        }
        i = j + 1;
    }
    std::string out = rooted ? "/" : "";
    for (size_t k = 0; k < parts.size(); k++) out += (k ? "/" : "") + parts[k];
    return out.empty() ? "." : out;
}
std::string go_dir(const std::string& path) {
    const size_t k = path.rfind('/');
    return go_clean(k == std::string::npos ? std::string() : path.substr(0, k + 1));
}
std::string go_join(const std::string& a, const std::string& b) {
    if (a.empty()) return b.empty() ? std::string() : go_clean(b);
    return go_clean(b.empty() ? a : a + "/" + b);
}

std::unique_ptr<Node> new_node(const char* type, const std::string& file) {
    auto n = std::make_unique<Node>();
    n->type = type;
    for (const FieldDef& fd : find_type(type)->fields) n->p[fd.name].kind = fd.kind;
    n->file = file;
    return n;
}
void set_map(Node& n, const char* k, float r, float g, float b) {
    Value& v = n.p[k];
    v.present = v.map_set = true;
    v.map_tex = false;
    v.c[0] = r; v.c[1] = g; v.c[2] = b;
}

// parseMTL (mtlparser.go:93-296): the ShaderStd nodes of one .mtl file; an error ends the file and
// the reference then drops every shader of it.
int parse_mtl(const std::string& path, const std::string& dir, std::vector<std::unique_ptr<Node>>& out, std::string& err) {
    std::ifstream f(path, std::ios::binary);
    if (!f) { err = "cannot open " + path; return -1; }
    std::string raw;
    Node* mtl = nullptr;
    int lineno = 0;
    while (std::getline(f, raw)) {
        lineno++;
        if (!raw.empty() && raw.back() == '\r') raw.pop_back();
        ObjLine L{raw};
        const std::string cmd = L.token();
        auto need = [&]() -> bool {
            if (mtl) return true;
            err = path + ":" + std::to_string(lineno) + ": " + cmd + " before newmtl (the reference dereferences nil)";
            return false;
        };
        auto three = [&](double* c) -> bool {   // only the last value's error is tested (err is reassigned)
            double x;
            go_parse_float32(L.token(), x); c[0] = x;
            go_parse_float32(L.token(), x); c[1] = x;
            const bool ok = go_parse_float32(L.token(), x);
            c[2] = x;
            if (!ok) err = path + ":" + std::to_string(lineno) + ": bad " + cmd;
            return ok;
        };
        if (cmd == "newmtl") {
            auto n = new_node("ShaderStd", path);
            n->line = lineno;
            n->p["Name"].s = L.token();
            n->p["Name"].present = true;
            mtl = n.get();
            out.push_back(std::move(n));
        } else if (cmd == "d") {
            double d;
            if (!go_parse_float32(L.token(), d)) { err = path + ": bad d"; return -1; }
            if (d == 0.0) {   // 'dissolve'
                if (!need()) return -1;
                set_map(*mtl, "DiffuseColour", 0.5f, 0.5f, 0.5f);
                set_map(*mtl, "DiffuseStrength", 0.5f, 0, 0);
            }
        } else if (cmd == "Ke") {
            double c[3];
            if (!three(c)) return -1;
            if (c[0] == 0.0 && c[1] == 0.0 && c[2] == 0.0) continue;
            if (!need()) return -1;
            const double st = std::sqrt(c[0] * c[0] + c[1] * c[1] + c[2] * c[2]);
            set_map(*mtl, "EmissionColour", (float)(c[0] / st), (float)(c[1] / st), (float)(c[2] / st));
            set_map(*mtl, "EmissionStrength", (float)st, 0, 0);
        } else if (cmd == "Kd" || cmd == "Ks") {
            double c[3];
            if (!three(c)) return -1;
            if (!need()) return -1;
            set_map(*mtl, cmd == "Kd" ? "DiffuseColour" : "Spec1Colour", (float)c[0], (float)c[1], (float)c[2]);
        } else if (cmd == "Ns") {
            double ns;
            if (!go_parse_float32(L.token(), ns)) { err = path + ": bad Ns"; return -1; }
            if (!need()) return -1;
            const float x = (float)ns;   // NsToSpecRoughness (mtlparser.go:25-31), m.Max as MAXSS
            const float y = 1.0f - x / 2048.0f;
            const float r = x <= 1000.0f ? 1.0f - x / 1000.0f : ((0.0f > y) ? 0.0f : y);
            set_map(*mtl, "Spec1Roughness", r, 0, 0);
        } else if (cmd == "illum") {
            int64_t m;
            if (!go_atoi(L.token(), m)) { err = path + ": bad illum"; return -1; }
            if (m >= 0 && m <= 10 && !need()) return -1;
            // the strengths per model (mtlparser.go:204-254)
            if (m == 0 || m == 1) set_map(*mtl, "DiffuseStrength", 0.5f, 0, 0);
            else if (m == 4 || m == 5) { set_map(*mtl, "DiffuseStrength", 0.2f, 0, 0); set_map(*mtl, "Spec1Strength", 0.5f, 0, 0); }
            else if (m >= 2 && m <= 10) { set_map(*mtl, "DiffuseStrength", 0.5f, 0, 0); set_map(*mtl, "Spec1Strength", 0.5f, 0, 0); }
            if (m == 5) { mtl->p["Spec1FresnelModel"].s = "Metal"; mtl->p["Spec1FresnelModel"].present = true; }
        } else if (cmd == "map_Kd") {
            if (!need()) return -1;
            Value& v = mtl->p["DiffuseColour"];
            v.present = v.map_set = v.map_tex = true;
            v.s = go_join(dir, L.rest());
            set_map(*mtl, "DiffuseStrength", 0.5f, 0, 0);
        }
    }
    return 0;
}

// File.parse (parser.go:103-263) and the Init that follows it (wfobj.go:21-54): the mesh (named by the
// file) and the shaders of its mtllib files, in order.
int parse_obj(vgpu_vnf* v, const std::string& filename, std::unique_ptr<Node>& mesh,
              std::vector<std::unique_ptr<Node>>& shaders) {
    std::ifstream f(filename, std::ios::binary);
    if (!f) return fail(VGPU_ERR_ARG, "wfobj: cannot open %s", filename.c_str());
    mesh = new_node("PolyMesh", filename);
    Node& m = *mesh;
    m.line = 1;
    m.p["Name"].s = filename;
    m.p["IsVisible"].b = true;
    Value& V = m.p["Verts"];
    Value& N = m.p["Normals"];
    Value& UV = m.p["UV"];
    Value &FI = m.p["FaceIdx"], &NI = m.p["NormalIdx"], &TI = m.p["UVIdx"], &PC = m.p["PolyCount"], &SI = m.p["ShaderIdx"];
    V.motion_keys = N.motion_keys = UV.motion_keys = 1;
    std::map<std::string, bool> mtllibs;
    uint8_t mtl = 0;
    std::string raw;
    while (std::getline(f, raw)) {
        if (!raw.empty() && raw.back() == '\r') raw.pop_back();
        ObjLine L{raw};
        const std::string cmd = L.token();
        if (cmd == "mtllib") {
            const std::string path = go_join(go_dir(filename), L.rest());
            if (mtllibs.count(path)) continue;
            std::vector<std::unique_ptr<Node>> got;
            std::string err;
            if (parse_mtl(path, go_dir(filename), got, err)) {
                v->log += "wfobj.parse: " + err + "\n";
                continue;
            }
            for (auto& s : got) shaders.push_back(std::move(s));
            if (shaders.size() > 255) return fail(VGPU_ERR_ARG, "wfobj.parse: too many shaders (%s)", filename.c_str());
            mtllibs[path] = true;
        } else if (cmd == "usemtl") {
            const std::string name = L.rest();
            for (size_t i = 0; i < shaders.size(); i++)
                if (shaders[i]->s("Name") == name) mtl = (uint8_t)i;
        } else if (cmd == "f") {   // parseFace (parser.go:41-101)
            int32_t nv = 0;
            for (std::string fld = L.token(); !fld.empty(); fld = L.token()) {
                std::vector<std::string> parts;
                size_t b = 0;
                for (;;) {
                    const size_t e = fld.find('/', b);
                    parts.push_back(fld.substr(b, e == std::string::npos ? std::string::npos : e - b));
                    if (e == std::string::npos) break;
                    b = e + 1;
                }
                int64_t p = 0, t = 0, n = 0;
                go_atoi(parts[0], p);
                if (parts.size() > 1 && !parts[1].empty()) go_atoi(parts[1], t);
                if (parts.size() > 2 && !parts[2].empty()) go_atoi(parts[2], n);
                if (p < 0) p = (int64_t)(V.elems.size() / 3) + p + 1; else if (p == 0) p = 1;
                if (t < 0) t = (int64_t)(UV.elems.size() / 2) + t + 1; else if (t == 0) t = 1;
                if (n < 0) n = (int64_t)(N.elems.size() / 3) + n + 1; else if (n == 0) n = 1;
                FI.ints.push_back((int32_t)(p - 1));
                NI.ints.push_back((int32_t)(n - 1));
                TI.ints.push_back((int32_t)(t - 1));
                nv++;
            }
            PC.ints.push_back(nv);
            SI.ints.push_back((int32_t)mtl);
            FI.slice_set = NI.slice_set = TI.slice_set = PC.slice_set = SI.slice_set = true;
        } else if (cmd == "v" || cmd == "vn" || cmd == "vt") {
            Value& dst = cmd == "v" ? V : (cmd == "vn" ? N : UV);
            const int w = cmd == "vt" ? 2 : 3;
            for (int k = 0; k < w; k++) {
                double x;
                go_parse_float32(L.token(), x);   // an error is logged and the value (0, or +-Inf) kept
                dst.elems.push_back((float)x);
            }
            dst.per_key++;
            dst.present = true;
        }
    }
    // no normals (or UVs): their indices are dropped (parser.go:252-260)
    if (N.elems.empty()) { NI.ints.clear(); NI.slice_set = false; }
    if (UV.elems.empty()) { TI.ints.clear(); TI.slice_set = false; }
    Value& sh = m.p["Shader"];
    sh.slice_set = sh.present = true;
    for (auto& s : shaders) sh.strs.push_back(s->s("Name"));
    return VGPU_OK;
}

bool is_geom(const Node& n) {
    return n.type == "PolyMesh" || n.type == "Sphere" || n.type == "GeomInstance" || n.type == "Proc";
}
bool is_light(const Node& n) {
    return n.type == "TriLight" || n.type == "DiskLight" || n.type == "SphereLight" || n.type == "QuadLight";
}

struct Build {
    vgpu_vnf* v;
    std::vector<Node*> all;        // every node in AddNode order
    std::map<std::string, Node*> by_name;
    std::vector<Node*> geoms, lights;
    Node* filter = nullptr;
    Node* globals = nullptr;
    std::vector<std::unique_ptr<Node>> made;   // nodes created during PreRender (light geometry)
    std::map<Node*, Node*> light_geom;         // light -> the geom its PreRender added
    std::map<Node*, std::vector<Node*>> proc_geoms;   // Proc -> the geoms its handler made (Proc.Geom)

    void add_node(Node* n) {   // core.AddNode (core/core.go:69-89)
        all.push_back(n);
        by_name[n->name()] = n;
        if (is_geom(*n)) geoms.push_back(n);
        else if (is_light(*n)) lights.push_back(n);
        else if (n->type == "GaussianFilter" || n->type == "AiryFilter") filter = n;
        else if (n->type == "Globals") globals = n;
    }

    Node* shader_node(const std::string& name) {   // core.FindNode + the core.Shader assertion
        auto it = by_name.find(name);
        if (it == by_name.end()) return nullptr;
        return (it->second->type == "ShaderStd" || it->second->type == "DebugShader") ? it->second : nullptr;
    }

    static void vset(Value& val, const float* e, int n, int keys, int per) {
        val.present = true;
        val.elems.assign(e, e + n);
        val.motion_keys = keys;
        val.per_key = per;
    }

    // Tri.createMesh (builtin/light/triangle.go:537-567) and Disk.createMesh (disk.go:224-262)
    std::unique_ptr<Node> light_mesh(const Node& L) {
        auto m = std::make_unique<Node>();
        m->type = "PolyMesh";
        for (const FieldDef& fd : find_type("PolyMesh")->fields) m->p[fd.name].kind = fd.kind;
        m->p["Name"].s = L.s("Name") + ":<mesh>";
        m->p["Shader"].strs = {L.s("Shader")};
        m->p["IsVisible"].b = true;
        m->file = L.file;
        m->line = L.line;
        std::vector<float> V, N;
        if (L.type == "TriLight") {
            const float* P0 = L.get("P0")->elems.data();
            const float* P1 = L.get("P1")->elems.data();
            const float* P2 = L.get("P2")->elems.data();
            float e1[3], e2[3], c[3], n[3];
            for (int k = 0; k < 3; k++) { e1[k] = P1[k] - P0[k]; e2[k] = P2[k] - P0[k]; }
            vgpu_detail::vec3_cross(e1, e2, c);
            vgpu_detail::vec3_normalize(c, n);
            V.insert(V.end(), P0, P0 + 3); V.insert(V.end(), P1, P1 + 3); V.insert(V.end(), P2, P2 + 3);
            for (int j = 0; j < 3; j++) N.insert(N.end(), n, n + 3);
        } else {
            const float* P = L.get("P")->elems.data();
            const float* Up = L.get("Up")->elems.data();
            const float* At = L.get("LookAt")->elems.data();
            float d[3] = {At[0] - P[0], At[1] - P[1], At[2] - P[2]}, n[3], t0[3], t[3], b[3];
            vgpu_detail::vec3_normalize(d, n);           // Disk.PreRender (disk.go:89-91)
            vgpu_detail::vec3_cross(n, Up, t0);
            vgpu_detail::vec3_normalize(t0, t);
            vgpu_detail::vec3_cross(n, t, b);
            const int nv = (int)L.i("Segments");
            const float R = (float)L.f("Radius");
            const float pi32 = 3.14159265358f;
            const float dv = 2.0f * pi32 / (float)nv;
            float ang = 0.0f;
            for (int i = 0; i < nv; i++) {
                const float c0 = (float)std::cos((double)ang), s0 = (float)std::sin((double)ang);
                const float a1 = ang + dv;
                const float c1 = (float)std::cos((double)a1), s1 = (float)std::sin((double)a1);
                V.insert(V.end(), P, P + 3);
                for (int k = 0; k < 3; k++) V.push_back(P[k] + (b[k] * (R * c0) + t[k] * (R * s0)));
                for (int k = 0; k < 3; k++) V.push_back(P[k] + (b[k] * (R * c1) + t[k] * (R * s1)));
                for (int j = 0; j < 3; j++) N.insert(N.end(), n, n + 3);
                ang += dv;
            }
        }
        const int nvert = (int)V.size() / 3;
        vset(m->p["Verts"], V.data(), (int)V.size(), 1, nvert);
        vset(m->p["Normals"], N.data(), (int)N.size(), 1, nvert);
        return m;
    }

    int prerender(Node* n, std::vector<Node*>& added) {
        const std::string& t = n->type;
        if (t == "Include") {
            std::vector<std::unique_ptr<Node>> sub;
            int rc = parse_file(v, n->s("Filename"), sub);
            if (rc) return rc;
            for (auto& s : sub) { add_node(s.get()); added.push_back(s.get()); made.push_back(std::move(s)); }
            return VGPU_OK;
        }
        if (t == "GeomInstance") {   // Instance.PreRender (builtin/geom/instance/instance.go:116-146)
            auto it = by_name.find(n->s("Geom"));
            if (it == by_name.end() || !is_geom(*it->second))
                return fail(VGPU_ERR_ARG, "%s:%d: Instance %s: Unable to find geom %s", n->file.c_str(), n->line,
                            n->s("Name").c_str(), n->s("Geom").c_str());
            return VGPU_OK;
        }
        if (t == "Proc" && n->s("Handler") == "wfobj") {
            // File.Init (builtin/geom/proc/wfobj/wfobj.go:21-54): the .obj file's mesh belongs to the Proc;
            // the shaders of its .mtl files join the scene's nodes
            std::unique_ptr<Node> mesh;
            std::vector<std::unique_ptr<Node>> shs;
            int rc = parse_obj(v, n->s("Data"), mesh, shs);
            if (rc) return rc;
            for (auto& x : shs) {
                Node* xp = x.get();
                made.push_back(std::move(x));
                add_node(xp);
                added.push_back(xp);
            }
            proc_geoms[n] = {mesh.get()};
            made.push_back(std::move(mesh));
            return VGPU_OK;
        }
        if (t == "Proc") {   // Proc.PreRender (builtin/geom/proc/proc.go:140-172)
            if (n->s("Handler") != "vnf")
                return fail(VGPU_ERR_ARG, "%s:%d: Proc %s: Unable to find handler %s", n->file.c_str(), n->line,
                            n->s("Name").c_str(), n->s("Handler").c_str());
            // VNFFile.Init (builtin/geom/proc/vnf/vnf.go:19-56): the file's shaders join the scene's nodes;
            // its geoms belong to the Proc alone; its other nodes are dropped
            std::vector<std::unique_ptr<Node>> sub;
            int rc = parse_file(v, n->s("Data"), sub);
            if (rc) return rc;
            std::vector<Node*> subn;
            for (auto& x : sub) { subn.push_back(x.get()); made.push_back(std::move(x)); }
            for (Node* x : subn)
                if (x->type == "ShaderStd" || x->type == "DebugShader") { add_node(x); added.push_back(x); }
            std::vector<Node*> kids;
            for (Node* x : subn) {
                if (x->type == "PolyMesh" || x->type == "Sphere") {
                    std::vector<std::string> names =
                        x->type == "PolyMesh" ? x->get("Shader")->strs : std::vector<std::string>{x->s("Shader")};
                    for (const std::string& sn : names)
                        if (!shader_node(sn))
                            return fail(VGPU_ERR_ARG, "%s:%d: Unable to find node (shader %s)", x->file.c_str(), x->line,
                                        sn.c_str());
                    kids.push_back(x);
                } else if (x->type == "GeomInstance" || x->type == "Proc") {
                    return fail(VGPU_ERR_ARG, "%s:%d: %s inside a Proc's file is not supported by this library",
                                x->file.c_str(), x->line, x->type.c_str());
                }
            }
            if (kids.empty())
                return fail(VGPU_ERR_ARG, "%s:%d: GeomProc: %s no geoms.", n->file.c_str(), n->line, n->s("Name").c_str());
            proc_geoms[n] = kids;
            return VGPU_OK;
        }
        if (t == "QuadLight")
            return fail(VGPU_ERR_ARG, "%s:%d: QuadLight cannot render (its SampleArea panics in the reference, "
                        "builtin/light/quad.go:87-95)", n->file.c_str(), n->line);
        if (is_light(*n) || t == "PolyMesh" || t == "Sphere") {
            std::vector<std::string> names = t == "PolyMesh" ? n->get("Shader")->strs : std::vector<std::string>{n->s("Shader")};
            for (const std::string& s : names)
                if (!shader_node(s))
                    return fail(VGPU_ERR_ARG, "%s:%d: Unable to find node (shader %s)", n->file.c_str(), n->line, s.c_str());
        }
        if (t == "TriLight" || t == "DiskLight") {
            std::unique_ptr<Node> m = light_mesh(*n);
            light_geom[n] = m.get();
            add_node(m.get());
            added.push_back(m.get());
            made.push_back(std::move(m));
        } else if (t == "SphereLight") {   // a sphere geom with the light's shader (sphere.go:51-53)
            auto g = std::make_unique<Node>();
            g->type = "Sphere";
            for (const FieldDef& fd : find_type("Sphere")->fields) g->p[fd.name].kind = fd.kind;
            g->p["Name"].s = n->s("Name") + ":<geom>";
            g->p["P"] = *n->get("P");
            g->p["Radius"] = *n->get("Radius");
            g->p["Shader"].s = n->s("Shader");
            g->file = n->file;
            g->line = n->line;
            light_geom[n] = g.get();
            add_node(g.get());
            added.push_back(g.get());
            made.push_back(std::move(g));
        }
        return VGPU_OK;
    }
};

int material_of(Build& B, const std::string& name, std::vector<Node*>& mats, int32_t& id) {
    Node* s = B.shader_node(name);
    if (!s) return fail(VGPU_ERR_ARG, "Unable to find node (shader %s)", name.c_str());
    if (s->type == "DebugShader") return fail(VGPU_ERR_ARG, "%s:%d: DebugShader is not supported by this library",
                                              s->file.c_str(), s->line);
    for (size_t i = 0; i < mats.size(); i++) if (mats[i] == s) { id = (int32_t)i; return VGPU_OK; }
    mats.push_back(s);
    id = (int32_t)mats.size() - 1;
    return VGPU_OK;
}

// ShaderStd parameters (builtin/shader/std.go:25-47) as a vgpu_material; a map given as rgbtex is a
// texture, which this library does not sample.
int to_material(const Node& s, vgpu_material& m) {
    std::memset(&m, 0, sizeof(m));
    auto map3 = [&](const char* k, float* out, int32_t* has) -> int {
        const Value* v = s.get(k);
        if (!v || !v->map_set) return VGPU_OK;
        if (v->map_tex) return fail(VGPU_ERR_ARG, "%s:%d: %s: textures (rgbtex) are not supported by this library",
                                    s.file.c_str(), s.line, k);
        std::memcpy(out, v->c, 3 * sizeof(float));
        *has = 1;
        return VGPU_OK;
    };
    auto map1 = [&](const char* k, float* out, int32_t* has) -> int {
        float c[3];
        int rc = map3(k, c, has);
        if (!rc && *has) *out = c[0];   // maps.Constant.Float32 returns C[0]
        return rc;
    };
    int rc = 0;
    if ((rc = map3("EmissionColour", m.emission, &m.has_emission_colour))) return rc;
    if ((rc = map1("EmissionStrength", &m.emission_strength, &m.has_emission_strength))) return rc;
    if ((rc = map3("DiffuseColour", m.diffuse, &m.has_diffuse_colour))) return rc;
    if ((rc = map1("DiffuseStrength", &m.diffuse_strength, &m.has_diffuse_strength))) return rc;
    if ((rc = map1("DiffuseRoughness", &m.diffuse_roughness, &m.has_diffuse_roughness))) return rc;
    if ((rc = map3("Spec1Colour", m.spec1, &m.has_spec1_colour))) return rc;
    if ((rc = map1("Spec1Strength", &m.spec1_strength, &m.has_spec1_strength))) return rc;
    if ((rc = map1("Spec1Roughness", &m.spec1_roughness, &m.has_spec1_roughness))) return rc;
    if ((rc = map1("IOR", &m.ior, &m.has_ior))) return rc;
    if ((rc = map3("Spec1FresnelRefl", m.spec1_fresnel_refl, &m.has_spec1_fresnel_refl))) return rc;
    if ((rc = map3("Spec1FresnelEdge", m.spec1_fresnel_edge, &m.has_spec1_fresnel_edge))) return rc;
    const std::string fm = s.s("Spec1FresnelModel");   // ShaderStd.PreRender (std.go:302-311)
    m.fresnel_model = fm == "Metal" ? VGPU_FRESNEL_METAL : fm == "Dielectric" ? VGPU_FRESNEL_DIELECTRIC : 0;
    return VGPU_OK;
}

// PolyMesh.init (builtin/geom/polymesh/init.go:12-134): fan triangulation of PolyCount polygons,
// or FaceIdx as triangles, or the vertices in order; shader indices per polygon or per triangle.
int triangulate(const Node& n, std::vector<uint32_t>& idx, std::vector<uint32_t>& nidx, std::vector<uint8_t>& sidx,
                bool& has_nidx) {
    const Value* pc = n.get("PolyCount");
    const Value* fi = n.get("FaceIdx");
    const Value* ni = n.get("NormalIdx");
    const Value* si = n.get("ShaderIdx");
    const Value* nrm = n.get("Normals");
    const bool normals = nrm && !nrm->elems.empty();
    const bool use_ni = ni && !ni->ints.empty();
    has_nidx = normals;
    auto at = [&](const std::vector<int32_t>& a, size_t k) -> int64_t { return k < a.size() ? a[k] : -1; };
    if (pc && pc->slice_set) {   // mesh.PolyCount != nil
        if (!fi || fi->ints.empty()) return fail(VGPU_ERR_ARG, "%s:%d: PolyCount without FaceIdx", n.file.c_str(), n.line);
        uint32_t base = 0;
        for (size_t k = 0; k < pc->ints.size(); k++) {
            const int32_t cnt = pc->ints[k];
            uint32_t i = 0;
            for (int j = 0; j < cnt - 2; j++) {
                i++;
                const size_t c[3] = {base, base + i, base + i + 1};
                for (size_t q : c) {
                    const int64_t x = at(fi->ints, q);
                    if (x < 0) return fail(VGPU_ERR_ARG, "%s:%d: FaceIdx index %zu out of range", n.file.c_str(), n.line, q);
                    idx.push_back((uint32_t)x);
                    if (normals) {
                        const int64_t y = use_ni ? at(ni->ints, q) : x;
                        if (y < 0) return fail(VGPU_ERR_ARG, "%s:%d: NormalIdx index %zu out of range", n.file.c_str(), n.line, q);
                        nidx.push_back((uint32_t)y);
                    }
                }
                if (si && !si->ints.empty()) {
                    if (k >= si->ints.size()) return fail(VGPU_ERR_ARG, "%s:%d: ShaderIdx too short", n.file.c_str(), n.line);
                    sidx.push_back((uint8_t)si->ints[k]);
                }
            }
            base += (uint32_t)cnt;
        }
    } else if (fi && fi->slice_set) {   // mesh.FaceIdx != nil
        for (size_t j = 0; j < fi->ints.size(); j++) {
            idx.push_back((uint32_t)fi->ints[j]);
            if (normals) nidx.push_back((uint32_t)(use_ni ? at(ni->ints, j) : fi->ints[j]));
        }
        if (si) for (int32_t x : si->ints) sidx.push_back((uint8_t)x);
    } else {
        const int per = n.get("Verts")->per_key;
        for (int j = 0; j < per; j++) {
            idx.push_back((uint32_t)j);
            if (normals) nidx.push_back((uint32_t)(use_ni ? at(ni->ints, j) : j));
        }
        if (si) for (int32_t x : si->ints) sidx.push_back((uint8_t)x);
    }
    if (idx.size() % 3 != 0) return fail(VGPU_ERR_ARG, "%s:%d: %zu indices do not form triangles", n.file.c_str(), n.line, idx.size());
    return VGPU_OK;
}

// One geom into the scene; *geom_out its insertion index.
int add_geom(Build& B, vgpu_scene* sc, Node& n, std::vector<Node*>& mats, int32_t& geom_out) {
    if (n.type == "Sphere") {
        int32_t mat;
        int rc = material_of(B, n.s("Shader"), mats, mat);
        if (rc) return rc;
        return vgpu_scene_add_sphere(sc, n.get("P")->elems.data(), (float)n.f("Radius"), mat, &geom_out);
    }
    const Value* vv = n.get("Verts");
    const int keys = vv->motion_keys < 1 ? 1 : vv->motion_keys;
    const int nverts = vv->per_key;
    if (nverts <= 0 || (int64_t)vv->elems.size() < (int64_t)keys * nverts * 3)
        return fail(VGPU_ERR_ARG, "%s:%d: PolyMesh %s has no vertices", n.file.c_str(), n.line, n.s("Name").c_str());
    std::vector<uint32_t> idx, nidx;
    std::vector<uint8_t> sidx;
    bool has_n = false;
    int rc = triangulate(n, idx, nidx, sidx, has_n);
    if (rc) return rc;
    for (uint32_t x : idx)
        if (x >= (uint32_t)nverts) return fail(VGPU_ERR_ARG, "%s:%d: vertex index %u out of range", n.file.c_str(), n.line, x);
    const uint32_t ntris = (uint32_t)(idx.size() / 3);
    std::vector<int32_t> shaders;
    for (const std::string& s : n.get("Shader")->strs) {
        int32_t id;
        if ((rc = material_of(B, s, mats, id))) return rc;
        shaders.push_back(id);
    }
    if (!sidx.empty() && sidx.size() != ntris)
        return fail(VGPU_ERR_ARG, "%s:%d: ShaderIdx has %zu entries for %u triangles", n.file.c_str(), n.line, sidx.size(), ntris);
    const Value* nv = n.get("Normals");
    if (keys == 1) {
        std::vector<float> verts(vv->elems.begin(), vv->elems.begin() + (size_t)nverts * 3);
        std::vector<float> normals;
        if (has_n) {
            // Normals[normalidx] per corner; where normalidx is not the position index, the corner gets
            // its own copy of the position (identical coordinates, so the same intersections)
            const size_t nn = nv->elems.size() / 3;
            bool same = nn >= (size_t)nverts;
            for (size_t q = 0; q < idx.size() && same; q++) same = nidx[q] == idx[q];
            if (same) {
                normals.assign(nv->elems.begin(), nv->elems.begin() + (size_t)nverts * 3);
            } else {
                std::vector<float> v2, n2;
                for (size_t q = 0; q < idx.size(); q++) {
                    if (nidx[q] >= nn) return fail(VGPU_ERR_ARG, "%s:%d: normal index %u out of range", n.file.c_str(), n.line, nidx[q]);
                    v2.insert(v2.end(), &verts[(size_t)idx[q] * 3], &verts[(size_t)idx[q] * 3] + 3);
                    n2.insert(n2.end(), &nv->elems[(size_t)nidx[q] * 3], &nv->elems[(size_t)nidx[q] * 3] + 3);
                    idx[q] = (uint32_t)q;
                }
                verts.swap(v2);
                normals.swap(n2);
            }
        }
        rc = vgpu_scene_add_shaded_mesh(sc, verts.data(), (uint32_t)(verts.size() / 3), idx.data(), ntris,
                                        normals.empty() ? nullptr : normals.data(), sidx.empty() ? nullptr : sidx.data(),
                                        shaders.data(), (uint32_t)shaders.size(), &geom_out);
    } else {
        // moving mesh: normals per face corner and key at stride Normals.ElemsPerKey (trace.go:694-703)
        std::vector<float> normals;
        uint32_t nkeys_n = 0;
        // the reference uses them only when Normals has as many motion keys as Verts (trace.go:694)
        if (nv && !nv->elems.empty() && nv->motion_keys == vv->motion_keys) {
            const int nk = nv->motion_keys;
            if ((int64_t)nv->per_key < (int64_t)ntris * 3 || (int64_t)nv->elems.size() < (int64_t)nk * nv->per_key * 3)
                return fail(VGPU_ERR_ARG, "%s:%d: moving mesh normals: %d per key for %u triangle corners",
                            n.file.c_str(), n.line, nv->per_key, ntris * 3);
            for (int k = 0; k < nk; k++)
                for (uint32_t c = 0; c < ntris * 3; c++)
                    normals.insert(normals.end(), &nv->elems[((size_t)k * nv->per_key + c) * 3],
                                   &nv->elems[((size_t)k * nv->per_key + c) * 3] + 3);
            nkeys_n = (uint32_t)nv->motion_keys;
        }
        rc = vgpu_scene_add_shaded_motion_mesh(sc, vv->elems.data(), (uint32_t)nverts, (uint32_t)keys, idx.data(), ntris,
                                               normals.empty() ? nullptr : normals.data(), nkeys_n,
                                               sidx.empty() ? nullptr : sidx.data(), shaders.data(),
                                               (uint32_t)shaders.size(), VGPU_MESH_MOTION_REFERENCE_INDEXING, &geom_out);
    }
    if (rc) return rc;
    sc->meshes.back().ray_bias = (float)n.f("RayBias");   // PolyMesh.RayBias (polymesh.go:19)
    // PolyMesh.Trace moves rays into the mesh's space whenever Transform is given (trace.go:22-53)
    const Value* tr = n.get("Transform");
    if (tr && tr->present && tr->elems.size() >= 16)
        return vgpu_scene_set_transform(sc, geom_out, tr->elems.data(), (uint32_t)(tr->elems.size() / 16));
    return VGPU_OK;
}

// GeomInstance and Proc bounds: BMin[0] and BMax[0] (only bounds[0] is used, instance.go:64-69)
int wrapper_box(const Node& n, const float** lo, const float** hi) {
    const Value* a = n.get("BMin");
    const Value* b = n.get("BMax");
    if (!a || !b || a->elems.size() < 3 || b->elems.size() < 3)
        return fail(VGPU_ERR_ARG, "%s:%d: %s %s needs BMin and BMax", n.file.c_str(), n.line, n.type.c_str(),
                    n.s("Name").c_str());
    *lo = a->elems.data();
    *hi = b->elems.data();
    return VGPU_OK;
}

int build(vgpu_vnf* v, int device, vgpu_scene** scene_out, vgpu_render** render_out) {
    Build B;
    B.v = v;
    for (auto& n : v->nodes) B.add_node(n.get());
    // core.PreRender: nodes added during a round are prerendered in the next (core/core.go:38-61)
    std::vector<Node*> round = B.all;
    while (!round.empty()) {
        std::vector<Node*> added;
        for (Node* n : round) {
            int rc = B.prerender(n, added);
            if (rc) return rc;
        }
        round.swap(added);
    }
    // globals as the last Globals node left them (Render reads Camera and MaxIter after PreRender)
    std::string cam_name = "camera";
    if (B.globals) {
        if (!B.globals->s("Camera").empty()) cam_name = B.globals->s("Camera");
        v->max_iter = (int32_t)B.globals->i("MaxIter");
    }
    auto cit = B.by_name.find(cam_name);
    if (cit == B.by_name.end() || cit->second->type != "Camera")
        return fail(VGPU_ERR_ARG, "no camera named \"%s\" (core.ErrNoCamera)", cam_name.c_str());
    const Node& cam = *cit->second;

    vgpu_scene* sc = nullptr;
    int rc = vgpu_scene_create(&sc, device);
    if (rc) return rc;
    std::unique_ptr<vgpu_scene, int (*)(vgpu_scene*)> guard(sc, vgpu_scene_destroy);
    std::vector<Node*> mats;
    std::map<Node*, int32_t> geom_index;
    for (Node* g : B.geoms) {
        int32_t gi = -1;
        const float *lo, *hi;
        if (g->type == "GeomInstance") {
            Node* tgt = B.by_name[g->s("Geom")];
            auto it = geom_index.find(tgt);
            if (it == geom_index.end())
                return fail(VGPU_ERR_ARG, "%s:%d: GeomInstance %s: its geom %s follows it in the scene (not supported "
                            "by this library)", g->file.c_str(), g->line, g->s("Name").c_str(), g->s("Geom").c_str());
            if (tgt->type != "PolyMesh" && tgt->type != "Sphere")
                return fail(VGPU_ERR_ARG, "%s:%d: GeomInstance %s: instances of instances or procs are not supported "
                            "by this library", g->file.c_str(), g->line, g->s("Name").c_str());
            const Value* tr = g->get("Transform");
            if ((rc = wrapper_box(*g, &lo, &hi))) return rc;
            if (!tr || tr->elems.size() < 16)
                return fail(VGPU_ERR_ARG, "%s:%d: GeomInstance %s has no Transform", g->file.c_str(), g->line,
                            g->s("Name").c_str());
            if ((rc = vgpu_scene_add_instance(sc, it->second, tr->elems.data(), (uint32_t)(tr->elems.size() / 16), lo,
                                              hi, &gi)))
                return rc;
        } else if (g->type == "Proc") {
            // the Proc's geoms: in the scene for the Proc to hold, hidden from the top level
            std::vector<int32_t> kids;
            for (Node* x : B.proc_geoms[g]) {
                int32_t k = -1;
                if ((rc = add_geom(B, sc, *x, mats, k))) return rc;
                if ((rc = vgpu_scene_set_hidden(sc, k, 1))) return rc;
                kids.push_back(k);
            }
            const Value* tr = g->get("Transform");
            const bool has_tr = tr && tr->present && tr->elems.size() >= 16;
            if ((rc = wrapper_box(*g, &lo, &hi))) return rc;
            if ((rc = vgpu_scene_add_proc(sc, kids.data(), (uint32_t)kids.size(), has_tr ? tr->elems.data() : nullptr,
                                          has_tr ? (uint32_t)(tr->elems.size() / 16) : 0u, lo, hi, &gi)))
                return rc;
        } else if ((rc = add_geom(B, sc, *g, mats, gi))) {
            return rc;
        }
        geom_index[g] = gi;
    }
    if ((rc = vgpu_scene_commit(sc))) return rc;

    std::vector<vgpu_light> lights;
    for (Node* L : B.lights) {
        vgpu_light l;
        std::memset(&l, 0, sizeof(l));
        int32_t mat;
        if ((rc = material_of(B, L->s("Shader"), mats, mat))) return rc;
        l.material = mat;
        l.samples = (int32_t)L->i("Samples");
        l.geom = geom_index[B.light_geom[L]];
        if (L->type == "TriLight") {
            l.type = VGPU_LIGHT_TRI;
            std::memcpy(l.P0, L->get("P0")->elems.data(), 12);
            std::memcpy(l.P1, L->get("P1")->elems.data(), 12);
            std::memcpy(l.P2, L->get("P2")->elems.data(), 12);
        } else if (L->type == "DiskLight") {
            l.type = VGPU_LIGHT_DISK;
            std::memcpy(l.P0, L->get("P")->elems.data(), 12);
            std::memcpy(l.P1, L->get("Up")->elems.data(), 12);
            std::memcpy(l.P2, L->get("LookAt")->elems.data(), 12);
            l.radius = (float)L->f("Radius");
        } else {
            l.type = VGPU_LIGHT_SPHERE;
            std::memcpy(l.P0, L->get("P")->elems.data(), 12);
            l.radius = (float)L->f("Radius");
        }
        lights.push_back(l);
    }
    std::vector<vgpu_material> mtab(mats.size());
    for (size_t i = 0; i < mats.size(); i++)
        if ((rc = to_material(*mats[i], mtab[i]))) return rc;
    if (mtab.empty()) { mtab.resize(1); std::memset(&mtab[0], 0, sizeof(vgpu_material)); }

    // Camera.PreRender (builtin/camera/camera.go:80-98)
    vgpu_camera c;
    std::memset(&c, 0, sizeof(c));
    for (int i = 0; i < 16; i++) c.M[i] = (i % 5 == 0) ? 1.0f : 0.0f;
    float aspect = (float)cam.f("Aspect");
    if (aspect == 0.0f) aspect = (float)v->width / (float)v->height;   // core.FrameAspect
    const float fov = (float)cam.f("Fov"), focal = (float)cam.f("Focal");
    const float half = fov / 2.0f;
    c.tan_theta_focal = (float)std::tan((double)(half * 3.14159265358f / 180.0f)) * focal;
    c.aspect = aspect;
    c.focal = focal;
    c.radius = (float)cam.f("Radius");
    std::vector<float> keys;
    const Value* ltw = cam.get("LocalToWorld");
    if (ltw && ltw->present) keys = ltw->elems;   // keys the file gave come first
    if (cam.s("Type") == "LookAt") {
        const Value* fr = cam.get("From");
        const Value* to = cam.get("To");
        const Value* ro = cam.get("Roll");
        if (fr->motion_keys < 1 || to->motion_keys < 1 || (int64_t)fr->elems.size() < 3LL * fr->motion_keys ||
            (int64_t)to->elems.size() < 3LL * to->motion_keys)
            return fail(VGPU_ERR_ARG, "%s:%d: camera From/To need at least one key", cam.file.c_str(), cam.line);
        std::vector<float> lk;
        vgpu_detail::lookat_keys(fr->elems.data(), fr->motion_keys, to->elems.data(), to->motion_keys,
                                 cam.get("Up")->elems.data(), ro->elems.empty() ? nullptr : ro->elems.data(),
                                 ro->elems.empty() ? 0 : ro->motion_keys, lk);
        keys.insert(keys.end(), lk.begin(), lk.end());
    } else {   // matrixCalc (camera.go:195-207): the inverses of WorldToLocal
        const Value* wtl = cam.get("WorldToLocal");
        if (wtl && wtl->present)
            for (size_t k = 0; k + 16 <= wtl->elems.size(); k += 16) {
                float inv[16];
                if (!vgpu_detail::mat4_inverse(&wtl->elems[k], inv)) std::memset(inv, 0, sizeof(inv));
                keys.insert(keys.end(), inv, inv + 16);
            }
    }
    if (keys.empty()) return fail(VGPU_ERR_ARG, "%s:%d: the camera has no transform", cam.file.c_str(), cam.line);

    vgpu_render* r = nullptr;
    if ((rc = vgpu_render_create(&r, sc, mtab.data(), (uint32_t)mtab.size(), lights.empty() ? nullptr : lights.data(),
                                 (uint32_t)lights.size(), &c, v->width, v->height)))
        return rc;
    std::unique_ptr<vgpu_render, int (*)(vgpu_render*)> rguard(r, vgpu_render_destroy);
    if ((rc = vgpu_render_set_camera_keys(r, keys.data(), (uint32_t)(keys.size() / 16)))) return rc;
    if (B.filter) {
        const Node& f = *B.filter;
        const bool gauss = f.type == "GaussianFilter";
        if ((rc = vgpu_render_set_filter(r, gauss ? VGPU_FILTER_GAUSSIAN : VGPU_FILTER_AIRY, (float)f.f("Width"),
                                         (int32_t)f.i("Res"), gauss ? 0.0f : (float)f.f("Peak")))) return rc;
    }
    v->outputs.clear();
    for (Node* n : B.all) {
        if (n->type == "OutputHDR") v->outputs.push_back({VGPU_OUTPUT_HDR, n->s("Filename")});
        if (n->type == "OutputFloat") v->outputs.push_back({VGPU_OUTPUT_FLOAT, n->s("Filename")});
    }
    for (auto& m : B.made) v->nodes.push_back(std::move(m));   // keep the PreRender nodes alive for dumps
    v->built = true;
    *scene_out = guard.release();
    *render_out = rguard.release();
    return VGPU_OK;
}

// Text form of a node list, for comparing two parsers: one line per parameter that is present,
// floats as the hex of their float32 bits.
void dump_nodes(const std::vector<std::unique_ptr<Node>>& nodes, std::string& out) {
    char buf[64];
    for (const auto& np : nodes) {
        const Node& n = *np;
        out += n.type + "\n";
        const NodeType* nt = find_type(n.type);
        for (const FieldDef& fd : nt->fields) {
            const Value& v = n.p.at(fd.name);
            if (!v.present) continue;
            out += "  " + std::string(fd.name) + ":";
            auto hexf = [&](float x) { uint32_t b; std::memcpy(&b, &x, 4); std::snprintf(buf, sizeof(buf), " %08x", b); out += buf; };
            switch (v.kind) {
                case KInt: out += " " + std::to_string(v.i); break;
                case KFloat: hexf((float)v.f); break;
                case KBool: out += v.b ? " 1" : " 0"; break;
                case KString: out += " \"" + v.s + "\""; break;
                case KInt32Slice: for (int32_t x : v.ints) out += " " + std::to_string(x); break;
                case KStringSlice: for (const std::string& x : v.strs) out += " \"" + x + "\""; break;
                case KMap:
                    if (!v.map_set) out += " nil";
                    else if (v.map_tex) out += " tex \"" + v.s + "\"";
                    else for (float x : v.c) hexf(x);
                    break;
                case KStruct: break;
                default:
                    out += " keys " + std::to_string(v.motion_keys) + " per " + std::to_string(v.per_key);
                    for (float x : v.elems) hexf(x);
                    break;
            }
            out += "\n";
        }
    }
}

}  // namespace

extern "C" {

int vgpu_vnf_load(const char* path, vgpu_vnf** out) {
    if (!path || !out) return fail(VGPU_ERR_ARG, "NULL argument");
    *out = nullptr;
    auto v = std::make_unique<vgpu_vnf>();
    int rc = parse_file(v.get(), path, v->nodes);
    if (rc) return rc;
    // core.AddNode of the file's nodes: the last Globals node sets the frame PreRender creates
    for (const auto& n : v->nodes)
        if (n->type == "Globals") {
            if (n->i("XRes") <= 0 || n->i("YRes") <= 0)
                return fail(VGPU_ERR_ARG, "%s:%d: Globals XRes and YRes must be positive", n->file.c_str(), n->line);
            v->width = (uint32_t)n->i("XRes");
            v->height = (uint32_t)n->i("YRes");
            v->max_iter = (int32_t)n->i("MaxIter");
        }
    *out = v.release();
    return VGPU_OK;
}

int vgpu_vnf_destroy(vgpu_vnf* v) {
    delete v;
    return VGPU_OK;
}

int vgpu_vnf_info(const vgpu_vnf* v, uint32_t* width, uint32_t* height, int32_t* max_iter, int32_t* nerrors) {
    if (!v) return fail(VGPU_ERR_ARG, "NULL argument");
    if (width) *width = v->width;
    if (height) *height = v->height;
    if (max_iter) *max_iter = v->max_iter;
    if (nerrors) *nerrors = v->nerrors;
    return VGPU_OK;
}

int vgpu_vnf_text(const vgpu_vnf* v, int32_t what, char* buf, uint64_t len, uint64_t* needed) {
    if (!v) return fail(VGPU_ERR_ARG, "NULL argument");
    std::string s;
    if (what == VGPU_VNF_MESSAGES) s = v->log;
    else if (what == VGPU_VNF_NODES) dump_nodes(v->nodes, s);
    else return fail(VGPU_ERR_ARG, "unknown text %d", what);
    if (needed) *needed = s.size() + 1;
    if (buf && len > 0) {
        const size_t n = std::min<size_t>(s.size(), (size_t)len - 1);
        std::memcpy(buf, s.data(), n);
        buf[n] = 0;
    }
    return VGPU_OK;
}

int vgpu_vnf_build(vgpu_vnf* v, int device, vgpu_scene** scene_out, vgpu_render** render_out) {
    if (!v || !scene_out || !render_out) return fail(VGPU_ERR_ARG, "NULL argument");
    if (v->built) return fail(VGPU_ERR_STATE, "the file's nodes were already built (PreRender runs once)");
    *scene_out = nullptr;
    *render_out = nullptr;
    return build(v, device, scene_out, render_out);
}

int vgpu_vnf_output(const vgpu_vnf* v, int32_t i, int32_t* type, char* path, uint64_t len, int32_t* count) {
    if (!v) return fail(VGPU_ERR_ARG, "NULL argument");
    if (count) *count = (int32_t)v->outputs.size();
    if (i < 0) return VGPU_OK;
    if (i >= (int32_t)v->outputs.size()) return fail(VGPU_ERR_ARG, "output %d of %zu", i, v->outputs.size());
    if (type) *type = v->outputs[i].first;
    if (path && len > 0) {
        const std::string& s = v->outputs[i].second;
        const size_t n = std::min<size_t>(s.size(), (size_t)len - 1);
        std::memcpy(path, s.data(), n);
        path[n] = 0;
    }
    return VGPU_OK;
}

}  // extern "C"
