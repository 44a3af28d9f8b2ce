// usd_lite.cpp — SDK-free reader of text USD layers (.usda) into the flat scene of the C ABI.
//
// The reference ingests USD through OpenUSD (src/scene/usd/USDLoader.cpp:1166-1833, loadUSD);
// this file restates what that loader does with one self-contained text layer, so that a caller
// without the USD SDK can hand a scene such as scenes/cornell_box.usda to ahip_prepare:
//   - prims are visited in stage order (depth-first, authored child order; `over`/`class` prims
//     and inactive subtrees are skipped, as UsdStage::Traverse's default predicate does);
//   - Mesh: fan triangulation with three vertices per triangle, world-space positions, normals
//     from `normals` (vertex / faceVarying) or the flat geometric normal, `primvars:st` /
//     `primvars:UVMap` (indexed primvars flattened), GeomSubset splitting and the
//     displayColor fallback (USDLoader.cpp:886-1033, 1035-1080, 1292-1430);
//   - Material: UsdPreviewSurface constants mapped to standard_surface parameters, default
//     Lambert(0.5) at index 0, glass by name when no surface shader (USDLoader.cpp:386-405,
//     526-779);
//   - RectLight / SphereLight / DiskLight as area lights, emissive meshes auto-registered as
//     area lights over their bounding box (USDLoader.cpp:1433-1545, 1748-1826); DistantLight as
//     a directional light (USDLoader.cpp:1553-1575);
//   - Camera: pinhole from focal length and horizontal aperture, selection by explicit path,
//     RenderSettings `camera`, or first camera (USDLoader.cpp:1084-1161, 1635-1722);
//   - upAxis "Z" rotates every point, normal and light axis to Y-up (USDLoader.cpp:58-67).
// Composition (references, payloads, sublayers, inherits, variants), animation (time-sampled
// transforms or geometry), texture-driven material inputs, DomeLight, OpenPBR
// networks and thin-lens cameras have no counterpart on the device path; a layer that uses any
// of them is refused with AHIP_E_UNSUPPORTED and a message naming the prim, never loaded in part.
#include "../../include/anacapa_hip.h"

#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <limits>
#include <map>
#include <memory>
#include <sstream>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <vector>

namespace {

// ----------------------------------------------------------------------------------------------
// errors
// ----------------------------------------------------------------------------------------------
struct LoadError : std::runtime_error {
    int status;
    LoadError(int s, const std::string& m) : std::runtime_error(m), status(s) {}
};
[[noreturn]] void unsupported(const std::string& m) { throw LoadError(AHIP_E_UNSUPPORTED, m); }
[[noreturn]] void invalid(const std::string& m) { throw LoadError(AHIP_E_INVALID_ARG, m); }

// ----------------------------------------------------------------------------------------------
// values and the layer tree
// ----------------------------------------------------------------------------------------------
struct Value {
    enum Kind { None, Number, String, Token, Path, Asset, Tuple, Array, Dict } kind = None;
    double num = 0.0;
    std::string str;
    std::vector<Value> items;
};

struct Property {
    std::string type;                 // "point3f[]", "double3", "rel", ...
    bool isRel = false;
    bool hasValue = false;
    Value value;
    bool hasConnect = false;
    std::vector<std::string> targets; // .connect sources or relationship targets (as written)
    std::vector<std::pair<double, Value>> samples;
    std::map<std::string, Value> meta;
};

struct Prim {
    std::string spec, type, name, path;
    Prim* parent = nullptr;
    std::map<std::string, Value> meta;
    std::map<std::string, Property> props;
    std::vector<std::unique_ptr<Prim>> children;
};

struct Layer {
    std::map<std::string, Value> meta;
    Prim root;                                   // pseudo-root, path "/"
    std::unordered_map<std::string, Prim*> byPath;
};

// ----------------------------------------------------------------------------------------------
// lexer
// ----------------------------------------------------------------------------------------------
enum class Tok { End, Punct, Ident, Number, String, Path, Asset };
struct Token {
    Tok kind = Tok::End;
    std::string text;
    double num = 0.0;
    int line = 0;
};

class Lexer {
public:
    explicit Lexer(const std::string& s) : s_(s) {}
    Token next() {
        skip();
        Token t;
        t.line = line_;
        if (i_ >= s_.size()) return t;
        const char c = s_[i_];
        if (std::strchr("()[]{}=,:;", c)) {
            t.kind = Tok::Punct;
            t.text = std::string(1, c);
            ++i_;
            return t;
        }
        if (c == '"' || c == '\'') return str(t);
        if (c == '<') {
            const size_t e = s_.find('>', i_ + 1);
            if (e == std::string::npos) fail("unterminated path");
            t.kind = Tok::Path;
            t.text = s_.substr(i_ + 1, e - i_ - 1);
            i_ = e + 1;
            return t;
        }
        if (c == '@') {
            const bool triple = s_.compare(i_, 3, "@@@") == 0;
            const size_t b = i_ + (triple ? 3 : 1);
            const size_t e = s_.find(triple ? "@@@" : "@", b);
            if (e == std::string::npos) fail("unterminated asset path");
            t.kind = Tok::Asset;
            t.text = s_.substr(b, e - b);
            i_ = e + (triple ? 3 : 1);
            return t;
        }
        const bool sign = (c == '-' || c == '+');
        const char c1 = i_ + 1 < s_.size() ? s_[i_ + 1] : '\0';
        if (std::isdigit((unsigned char)c) || (c == '.' && std::isdigit((unsigned char)c1)) ||
            (sign && (std::isdigit((unsigned char)c1) || c1 == '.')))
            return number(t);
        if (sign && s_.compare(i_ + 1, 3, "inf") == 0) {
            t.kind = Tok::Number;
            t.num = c == '-' ? -std::numeric_limits<double>::infinity() : std::numeric_limits<double>::infinity();
            i_ += 4;
            return t;
        }
        if (std::isalpha((unsigned char)c) || c == '_') {
            const size_t b = i_;
            while (i_ < s_.size() && (std::isalnum((unsigned char)s_[i_]) || s_[i_] == '_' || s_[i_] == ':' ||
                                      s_[i_] == '.'))
                ++i_;
            t.kind = Tok::Ident;
            t.text = s_.substr(b, i_ - b);
            if (t.text == "inf" || t.text == "nan") {
                t.kind = Tok::Number;
                t.num = t.text == "inf" ? std::numeric_limits<double>::infinity()
                                        : std::numeric_limits<double>::quiet_NaN();
            }
            return t;
        }
        fail(std::string("unexpected character '") + c + "'");
    }
    [[noreturn]] void fail(const std::string& m) const {
        invalid("usda line " + std::to_string(line_) + ": " + m);
    }

private:
    void skip() {
        while (i_ < s_.size()) {
            const char c = s_[i_];
            if (c == '\n') {
                ++line_;
                ++i_;
            } else if (std::isspace((unsigned char)c)) {
                ++i_;
            } else if (c == '#') {
                while (i_ < s_.size() && s_[i_] != '\n') ++i_;
            } else {
                break;
            }
        }
    }
    Token str(Token& t) {
        const char q = s_[i_];
        const std::string q3(3, q);
        t.kind = Tok::String;
        if (s_.compare(i_, 3, q3) == 0) {
            const size_t e = s_.find(q3, i_ + 3);
            if (e == std::string::npos) fail("unterminated string");
            t.text = s_.substr(i_ + 3, e - i_ - 3);
            line_ += (int)std::count(t.text.begin(), t.text.end(), '\n');
            i_ = e + 3;
            return t;
        }
        ++i_;
        while (i_ < s_.size() && s_[i_] != q) {
            char c = s_[i_++];
            if (c == '\n') fail("newline in string");
            if (c == '\\' && i_ < s_.size()) {
                const char e = s_[i_++];
                c = e == 'n' ? '\n' : e == 't' ? '\t' : e;
            }
            t.text.push_back(c);
        }
        if (i_ >= s_.size()) fail("unterminated string");
        ++i_;
        return t;
    }
    Token number(Token& t) {
        const char* b = s_.c_str() + i_;
        char* e = nullptr;
        t.num = std::strtod(b, &e);
        if (e == b) fail("bad number");
        t.kind = Tok::Number;
        t.text.assign(b, (size_t)(e - b));
        i_ += (size_t)(e - b);
        return t;
    }

    const std::string& s_;
    size_t i_ = 0;
    int line_ = 1;
};

// ----------------------------------------------------------------------------------------------
// parser (recursive descent over the text layer grammar)
// ----------------------------------------------------------------------------------------------
class Parser {
public:
    explicit Parser(const std::string& text) : lx_(text) { tok_ = lx_.next(); }

    void parseLayer(Layer& L) {
        L.root.path = "/";
        L.byPath["/"] = &L.root;
        if (isPunct("(")) parseMeta(L.meta);
        while (tok_.kind != Tok::End) parsePrim(L, &L.root);
    }

private:
    bool isPunct(const char* p) const { return tok_.kind == Tok::Punct && tok_.text == p; }
    bool isIdent(const char* p) const { return tok_.kind == Tok::Ident && tok_.text == p; }
    Token take() {
        Token t = tok_;
        tok_ = lx_.next();
        return t;
    }
    void expect(const char* p) {
        if (!isPunct(p)) fail(std::string("expected '") + p + "'");
        take();
    }
    std::string ident(const char* what) {
        if (tok_.kind != Tok::Ident) fail(std::string("expected ") + what);
        return take().text;
    }
    [[noreturn]] void fail(const std::string& m) const {
        invalid("usda line " + std::to_string(tok_.line) + ": " + m +
                (tok_.kind == Tok::End ? " (end of file)" : " near '" + tok_.text + "'"));
    }

    Value parseValue() {
        Value v;
        if (isPunct("(") || isPunct("[")) {
            const std::string close = isPunct("(") ? ")" : "]";
            v.kind = isPunct("(") ? Value::Tuple : Value::Array;
            take();
            while (!isPunct(close.c_str())) {
                v.items.push_back(parseValue());
                if (isPunct(",")) take();
                else if (!isPunct(close.c_str())) fail("expected ',' or '" + close + "'");
            }
            take();
            return v;
        }
        if (isPunct("{")) {  // dictionary-valued metadata: kept opaque
            int depth = 0;
            do {
                if (tok_.kind == Tok::End) fail("unterminated dictionary");
                if (isPunct("{")) ++depth;
                if (isPunct("}")) --depth;
                take();
            } while (depth > 0);
            v.kind = Value::Dict;
            return v;
        }
        Token t = take();
        switch (t.kind) {
            case Tok::Number: v.kind = Value::Number; v.num = t.num; break;
            case Tok::String: v.kind = Value::String; v.str = t.text; break;
            case Tok::Path: v.kind = Value::Path; v.str = t.text; break;
            case Tok::Asset:  // @asset@ optionally followed by </prim/path>
                v.kind = Value::Asset;
                v.str = t.text;
                if (tok_.kind == Tok::Path) take();
                break;
            case Tok::Ident: v.kind = t.text == "None" ? Value::None : Value::Token; v.str = t.text; break;
            default: fail("expected a value");
        }
        return v;
    }

    void parseMeta(std::map<std::string, Value>& meta) {
        expect("(");
        while (!isPunct(")")) {
            if (tok_.kind == Tok::String) {
                meta["doc"] = parseValue();
            } else if (tok_.kind == Tok::Ident) {
                std::string key = take().text;
                if ((key == "prepend" || key == "append" || key == "add" || key == "delete" || key == "reorder") &&
                    tok_.kind == Tok::Ident)
                    key = take().text;
                expect("=");
                meta[key] = parseValue();
            } else if (isPunct(";")) {
                take();
            } else {
                fail("bad metadata entry");
            }
        }
        take();
    }

    void parsePrim(Layer& L, Prim* parent) {
        if (!(isIdent("def") || isIdent("over") || isIdent("class"))) fail("expected def, over or class");
        auto p = std::make_unique<Prim>();
        p->spec = take().text;
        if (tok_.kind == Tok::Ident) p->type = take().text;
        if (tok_.kind != Tok::String) fail("expected prim name");
        p->name = take().text;
        p->parent = parent;
        p->path = (parent->path == "/" ? "" : parent->path) + "/" + p->name;
        if (isPunct("(")) parseMeta(p->meta);
        expect("{");
        while (!isPunct("}")) {
            if (tok_.kind == Tok::End) fail("unterminated prim " + p->path);
            if (isIdent("def") || isIdent("over") || isIdent("class")) {
                parsePrim(L, p.get());
            } else if (isIdent("variantSet")) {
                unsupported("variant sets (prim " + p->path + ")");
            } else if (isIdent("reorder")) {
                unsupported("reorder statements (prim " + p->path + ")");
            } else if (isPunct(";")) {
                take();
            } else {
                parseProperty(*p);
            }
        }
        take();
        Prim* raw = p.get();
        if (!L.byPath.emplace(raw->path, raw).second) invalid("duplicate prim " + raw->path);
        parent->children.push_back(std::move(p));
    }

    static void targetsOf(const Value& v, std::vector<std::string>& out) {
        if (v.kind == Value::Path) out.push_back(v.str);
        for (const Value& it : v.items) targetsOf(it, out);
    }

    void parseProperty(Prim& p) {
        bool listop = false;
        while (isIdent("custom") || isIdent("uniform") || isIdent("varying") || isIdent("config") ||
               isIdent("prepend") || isIdent("append") || isIdent("add") || isIdent("delete")) {
            const std::string q = take().text;
            if (q == "prepend" || q == "append" || q == "add" || q == "delete") listop = true;
        }
        std::string type = ident("a property type");
        if (type == "rel") {
            const std::string name = ident("a relationship name");
            Property& r = p.props[name];
            r.type = "rel";
            r.isRel = true;
            if (isPunct("=")) {
                take();
                Value v = parseValue();
                r.targets.clear();
                targetsOf(v, r.targets);
                r.hasValue = true;
            }
            if (isPunct("(")) parseMeta(r.meta);
            (void)listop;
            return;
        }
        if (isPunct("[")) {
            take();
            expect("]");
            type += "[]";
        }
        std::string name = ident("a property name");
        auto endsWith = [&](const char* suf) {
            const size_t n = std::strlen(suf);
            return name.size() > n && name.compare(name.size() - n, n, suf) == 0;
        };
        if (endsWith(".connect")) {
            name.resize(name.size() - 8);
            Property& a = p.props[name];
            a.type = type;
            expect("=");
            Value v = parseValue();
            a.hasConnect = true;
            a.targets.clear();
            targetsOf(v, a.targets);
            if (isPunct("(")) parseMeta(a.meta);
            return;
        }
        if (endsWith(".timeSamples")) {
            name.resize(name.size() - 12);
            Property& a = p.props[name];
            a.type = type;
            expect("=");
            expect("{");
            while (!isPunct("}")) {
                if (tok_.kind != Tok::Number) fail("expected a time code");
                const double t = take().num;
                expect(":");
                a.samples.emplace_back(t, parseValue());
                if (isPunct(",")) take();
            }
            take();
            std::stable_sort(a.samples.begin(), a.samples.end(),
                             [](const auto& x, const auto& y) { return x.first < y.first; });
            return;
        }
        if (name.find('.') != std::string::npos) unsupported("property form '" + name + "' (prim " + p.path + ")");
        Property& a = p.props[name];
        a.type = type;
        if (isPunct("=")) {
            take();
            a.value = parseValue();
            a.hasValue = a.value.kind != Value::None;
        }
        if (isPunct("(")) parseMeta(a.meta);
    }

    Lexer lx_;
    Token tok_;
};

// ----------------------------------------------------------------------------------------------
// math: USD's row-vector convention (p' = p * M, translation in the last row), in double
// ----------------------------------------------------------------------------------------------
struct M4 {
    double m[4][4];
    static M4 identity() {
        M4 r{};
        for (int i = 0; i < 4; ++i) r.m[i][i] = 1.0;
        return r;
    }
    M4 operator*(const M4& b) const {
        M4 r{};
        for (int i = 0; i < 4; ++i)
            for (int j = 0; j < 4; ++j) {
                double s = 0.0;
                for (int k = 0; k < 4; ++k) s += m[i][k] * b.m[k][j];
                r.m[i][j] = s;
            }
        return r;
    }
    bool operator==(const M4& b) const { return std::memcmp(m, b.m, sizeof m) == 0; }
};

constexpr double kPi = 3.14159265358979323846;

// Rotation by `deg` about a unit axis, built from the unit quaternion (cos θ/2, sin θ/2 · axis)
// and laid out for row vectors (p * R turns p counter-clockwise about the axis).
M4 quatMatrix(double r, double x, double y, double z) {
    const double n = std::sqrt(r * r + x * x + y * y + z * z);
    if (n > 0.0) { r /= n; x /= n; y /= n; z /= n; }
    M4 R = M4::identity();
    R.m[0][0] = 1.0 - 2.0 * (y * y + z * z);
    R.m[0][1] = 2.0 * (x * y + z * r);
    R.m[0][2] = 2.0 * (z * x - y * r);
    R.m[1][0] = 2.0 * (x * y - z * r);
    R.m[1][1] = 1.0 - 2.0 * (z * z + x * x);
    R.m[1][2] = 2.0 * (y * z + x * r);
    R.m[2][0] = 2.0 * (z * x + y * r);
    R.m[2][1] = 2.0 * (y * z - x * r);
    R.m[2][2] = 1.0 - 2.0 * (y * y + x * x);
    return R;
}
M4 axisRotation(int axis, double deg) {
    const double h = deg * (kPi / 180.0) * 0.5;
    const double s = std::sin(h), c = std::cos(h);
    return quatMatrix(c, axis == 0 ? s : 0.0, axis == 1 ? s : 0.0, axis == 2 ? s : 0.0);
}

// General 4x4 inverse (cofactor expansion); throws on a singular matrix.
M4 inverse(const M4& a) {
    const double* m = &a.m[0][0];
    double inv[16];
    inv[0] = m[5] * m[10] * m[15] - m[5] * m[11] * m[14] - m[9] * m[6] * m[15] + m[9] * m[7] * m[14] +
             m[13] * m[6] * m[11] - m[13] * m[7] * m[10];
    inv[4] = -m[4] * m[10] * m[15] + m[4] * m[11] * m[14] + m[8] * m[6] * m[15] - m[8] * m[7] * m[14] -
             m[12] * m[6] * m[11] + m[12] * m[7] * m[10];
    inv[8] = m[4] * m[9] * m[15] - m[4] * m[11] * m[13] - m[8] * m[5] * m[15] + m[8] * m[7] * m[13] +
             m[12] * m[5] * m[11] - m[12] * m[7] * m[9];
    inv[12] = -m[4] * m[9] * m[14] + m[4] * m[10] * m[13] + m[8] * m[5] * m[14] - m[8] * m[6] * m[13] -
              m[12] * m[5] * m[10] + m[12] * m[6] * m[9];
    inv[1] = -m[1] * m[10] * m[15] + m[1] * m[11] * m[14] + m[9] * m[2] * m[15] - m[9] * m[3] * m[14] -
             m[13] * m[2] * m[11] + m[13] * m[3] * m[10];
    inv[5] = m[0] * m[10] * m[15] - m[0] * m[11] * m[14] - m[8] * m[2] * m[15] + m[8] * m[3] * m[14] +
             m[12] * m[2] * m[11] - m[12] * m[3] * m[10];
    inv[9] = -m[0] * m[9] * m[15] + m[0] * m[11] * m[13] + m[8] * m[1] * m[15] - m[8] * m[3] * m[13] -
             m[12] * m[1] * m[11] + m[12] * m[3] * m[9];
    inv[13] = m[0] * m[9] * m[14] - m[0] * m[10] * m[13] - m[8] * m[1] * m[14] + m[8] * m[2] * m[13] +
              m[12] * m[1] * m[10] - m[12] * m[2] * m[9];
    inv[2] = m[1] * m[6] * m[15] - m[1] * m[7] * m[14] - m[5] * m[2] * m[15] + m[5] * m[3] * m[14] +
             m[13] * m[2] * m[7] - m[13] * m[3] * m[6];
    inv[6] = -m[0] * m[6] * m[15] + m[0] * m[7] * m[14] + m[4] * m[2] * m[15] - m[4] * m[3] * m[14] -
             m[12] * m[2] * m[7] + m[12] * m[3] * m[6];
    inv[10] = m[0] * m[5] * m[15] - m[0] * m[7] * m[13] - m[4] * m[1] * m[15] + m[4] * m[3] * m[13] +
              m[12] * m[1] * m[7] - m[12] * m[3] * m[5];
    inv[14] = -m[0] * m[5] * m[14] + m[0] * m[6] * m[13] + m[4] * m[1] * m[14] - m[4] * m[2] * m[13] -
              m[12] * m[1] * m[6] + m[12] * m[2] * m[5];
    inv[3] = -m[1] * m[6] * m[11] + m[1] * m[7] * m[10] + m[5] * m[2] * m[11] - m[5] * m[3] * m[10] -
             m[9] * m[2] * m[7] + m[9] * m[3] * m[6];
    inv[7] = m[0] * m[6] * m[11] - m[0] * m[7] * m[10] - m[4] * m[2] * m[11] + m[4] * m[3] * m[10] +
             m[8] * m[2] * m[7] - m[8] * m[3] * m[6];
    inv[11] = -m[0] * m[5] * m[11] + m[0] * m[7] * m[9] + m[4] * m[1] * m[11] - m[4] * m[3] * m[9] -
              m[8] * m[1] * m[7] + m[8] * m[3] * m[5];
    inv[15] = m[0] * m[5] * m[10] - m[0] * m[6] * m[9] - m[4] * m[1] * m[10] + m[4] * m[2] * m[9] +
              m[8] * m[1] * m[6] - m[8] * m[2] * m[5];
    const double det = m[0] * inv[0] + m[1] * inv[4] + m[2] * inv[8] + m[3] * inv[12];
    if (det == 0.0 || !std::isfinite(det)) invalid("singular transform");
    M4 r;
    for (int i = 0; i < 16; ++i) (&r.m[0][0])[i] = inv[i] / det;
    return r;
}

struct V3 {
    float x = 0.f, y = 0.f, z = 0.f;
    V3 operator+(V3 b) const { return {x + b.x, y + b.y, z + b.z}; }
    V3 operator-(V3 b) const { return {x - b.x, y - b.y, z - b.z}; }
    V3 operator*(float s) const { return {x * s, y * s, z * s}; }
    float length() const { return std::sqrt(x * x + y * y + z * z); }
};
V3 cross(V3 a, V3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
// Types.h:86-89 safeNormalize
V3 safeNormalize(V3 v) {
    const float len = v.length();
    return len > 1e-8f ? v * (1.f / len) : V3{0.f, 0.f, 1.f};
}
// USDLoader.cpp:58-67: Z-up to Y-up (x, y, z) -> (x, z, -y)
V3 upCorrect(V3 v, bool zUp) { return zUp ? V3{v.x, v.z, -v.y} : v; }
// USDLoader.cpp:71-76: transform in double, then round to float
V3 transformPoint(const M4& M, double x, double y, double z, bool zUp) {
    double r[4];
    for (int j = 0; j < 4; ++j) r[j] = x * M.m[0][j] + y * M.m[1][j] + z * M.m[2][j] + M.m[3][j];
    if (r[3] != 1.0 && r[3] != 0.0)
        for (int j = 0; j < 3; ++j) r[j] /= r[3];
    return upCorrect({(float)r[0], (float)r[1], (float)r[2]}, zUp);
}
// USDLoader.cpp:92-98: inverse transpose applied as a direction, then normalized in float
V3 transformNormal(const M4& invT, float nx, float ny, float nz, bool zUp) {
    double r[3];
    for (int j = 0; j < 3; ++j) r[j] = nx * invT.m[0][j] + ny * invT.m[1][j] + nz * invT.m[2][j];
    return safeNormalize(upCorrect({(float)r[0], (float)r[1], (float)r[2]}, zUp));
}
M4 transposeOf(const M4& a) {
    M4 r;
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) r.m[i][j] = a.m[j][i];
    return r;
}

// ----------------------------------------------------------------------------------------------
// value access
// ----------------------------------------------------------------------------------------------
bool asNumber(const Value& v, double& out) {
    if (v.kind == Value::Number) { out = v.num; return true; }
    if (v.kind == Value::Token && (v.str == "true" || v.str == "false")) { out = v.str == "true"; return true; }
    return false;
}
std::vector<double> asNumbers(const Value& v, const std::string& what) {
    std::vector<double> r;
    double d;
    if (asNumber(v, d)) { r.push_back(d); return r; }
    if (v.kind != Value::Tuple) invalid(what + ": expected a number or tuple");
    for (const Value& it : v.items) {
        if (!asNumber(it, d)) invalid(what + ": expected numbers");
        r.push_back(d);
    }
    return r;
}

// Linear interpolation of numbers / tuples between bracketing time samples; values held
// outside the sampled range (USD's default for interpolatable types).
Value lerpValue(const Value& a, const Value& b, double t) {
    double x, y;
    if (asNumber(a, x) && asNumber(b, y)) {
        Value r;
        r.kind = Value::Number;
        r.num = x + (y - x) * t;
        return r;
    }
    if ((a.kind == Value::Tuple || a.kind == Value::Array) && a.kind == b.kind && a.items.size() == b.items.size()) {
        Value r = a;
        for (size_t i = 0; i < a.items.size(); ++i) r.items[i] = lerpValue(a.items[i], b.items[i], t);
        return r;
    }
    return a;
}

class Stage {
public:
    Layer L;
    bool zUp = false;
    double startTime = 0.0, endTime = 1.0;

    const Property* prop(const Prim& p, const std::string& n) const {
        auto it = p.props.find(n);
        return it == p.props.end() ? nullptr : &it->second;
    }
    // Value at the default time: the authored default; time-sampled attributes that these
    // readers take at the default time are refused (animated geometry is not supported).
    const Value* get(const Prim& p, const std::string& n) const {
        const Property* a = prop(p, n);
        if (!a) return nullptr;
        if (!a->samples.empty()) unsupported("animated attribute " + p.path + "." + n);
        return a->hasValue ? &a->value : nullptr;
    }
    // Value at a numeric time: time samples win over the default.
    const Value* getAt(const Prim& p, const std::string& n, double t, Value& tmp) const {
        const Property* a = prop(p, n);
        if (!a) return nullptr;
        if (a->samples.empty()) return a->hasValue ? &a->value : nullptr;
        const auto& S = a->samples;
        if (t <= S.front().first) return &S.front().second;
        if (t >= S.back().first) return &S.back().second;
        for (size_t i = 1; i < S.size(); ++i)
            if (t <= S[i].first) {
                const double u = (t - S[i - 1].first) / (S[i].first - S[i - 1].first);
                tmp = lerpValue(S[i - 1].second, S[i].second, u);
                return &tmp;
            }
        return &S.back().second;
    }
    float getFloat(const Prim& p, const std::string& n, float def) const {
        const Value* v = get(p, n);
        double d;
        return v && asNumber(*v, d) ? (float)d : def;
    }
    bool getBool(const Prim& p, const std::string& n, bool def) const {
        const Value* v = get(p, n);
        double d;
        return v && asNumber(*v, d) ? d != 0.0 : def;
    }
    bool getVec3(const Prim& p, const std::string& n, float out[3]) const {
        const Value* v = get(p, n);
        if (!v) return false;
        std::vector<double> d = asNumbers(*v, p.path + "." + n);
        if (d.size() < 3) invalid(p.path + "." + n + ": expected 3 components");
        for (int i = 0; i < 3; ++i) out[i] = (float)d[i];
        return true;
    }
    std::string getToken(const Prim& p, const std::string& n) const {
        const Value* v = get(p, n);
        return v && (v->kind == Value::Token || v->kind == Value::String) ? v->str : std::string();
    }
    std::vector<double> getFlat(const Prim& p, const std::string& n, size_t arity) const {
        std::vector<double> r;
        const Value* v = get(p, n);
        if (!v) return r;
        if (v->kind != Value::Array) invalid(p.path + "." + n + ": expected an array");
        r.reserve(v->items.size() * arity);
        for (const Value& it : v->items) {
            std::vector<double> d = asNumbers(it, p.path + "." + n);
            if (d.size() != arity) invalid(p.path + "." + n + ": wrong element arity");
            r.insert(r.end(), d.begin(), d.end());
        }
        return r;
    }
    std::string metaString(const Value* v) const {
        return v && (v->kind == Value::Token || v->kind == Value::String) ? v->str : std::string();
    }

    Prim* at(const std::string& path) const {
        auto it = L.byPath.find(path);
        return it == L.byPath.end() ? nullptr : it->second;
    }
    // Resolve a target path written on `owner` (absolute, or relative with ./ and ../) into
    // (prim path, property name).
    void resolveTarget(const Prim& owner, const std::string& t, std::string& primPath, std::string& propName) const {
        std::string full = t;
        if (full.empty() || full[0] != '/') {
            std::vector<std::string> parts;
            std::stringstream a(owner.path);
            std::string seg;
            while (std::getline(a, seg, '/'))
                if (!seg.empty()) parts.push_back(seg);
            std::stringstream b(t);
            while (std::getline(b, seg, '/')) {
                if (seg.empty() || seg == ".") continue;
                if (seg == "..") {
                    if (!parts.empty()) parts.pop_back();
                } else {
                    parts.push_back(seg);
                }
            }
            full.clear();
            for (const auto& s : parts) full += "/" + s;
            if (full.empty()) full = "/";
        }
        const size_t slash = full.rfind('/');
        const size_t dot = full.find('.', slash == std::string::npos ? 0 : slash);
        if (dot == std::string::npos) {
            primPath = full;
            propName.clear();
        } else {
            primPath = full.substr(0, dot);
            propName = full.substr(dot + 1);
        }
    }

    // --- transforms (UsdGeomXformable / UsdGeomXformCache semantics) ---
    // Local transform at time t; `atDefault` reads the ops' default values instead.
    M4 localXform(const Prim& p, double t, bool atDefault, bool* reset) const {
        *reset = false;
        M4 M = M4::identity();
        const Property* order = prop(p, "xformOpOrder");
        if (!order || !order->hasValue) return M;
        if (!order->samples.empty()) unsupported("animated xformOpOrder on " + p.path);
        for (const Value& opv : order->value.items) {
            std::string op = opv.str;
            if (op == "!resetXformStack!") {
                *reset = true;
                M = M4::identity();
                continue;
            }
            bool invert = false;
            if (op.rfind("!invert!", 0) == 0) {
                invert = true;
                op = op.substr(8);
            }
            if (op.rfind("xformOp:", 0) != 0) invalid("bad xformOp name '" + op + "' on " + p.path);
            std::string kind = op.substr(8);
            const size_t colon = kind.find(':');
            if (colon != std::string::npos) kind = kind.substr(0, colon);
            Value tmp;
            const Value* v = nullptr;
            if (atDefault) {
                const Property* a = prop(p, op);
                if (a && !a->samples.empty() && !a->hasValue)
                    unsupported("camera transform given only by time samples on " + p.path);
                v = a && a->hasValue ? &a->value : nullptr;
            } else {
                v = getAt(p, op, t, tmp);
            }
            if (!v) invalid("xformOp " + op + " on " + p.path + " has no value");
            // float- and half-typed ops hold single-precision values, widened to double
            const Property* opProp = prop(p, op);
            const bool single = opProp && (opProp->type.rfind("float", 0) == 0 || opProp->type.rfind("half", 0) == 0 ||
                                           opProp->type == "quatf" || opProp->type == "quath");
            auto nums = [&](const Value& x) {
                std::vector<double> d = asNumbers(x, op);
                if (single)
                    for (double& e : d) e = (double)(float)e;
                return d;
            };
            M4 O = M4::identity();
            if (kind == "translate") {
                std::vector<double> d = nums(*v);
                if (d.size() != 3) invalid(op + ": expected 3 components");
                O.m[3][0] = d[0]; O.m[3][1] = d[1]; O.m[3][2] = d[2];
            } else if (kind == "scale") {
                std::vector<double> d = nums(*v);
                if (d.size() != 3) invalid(op + ": expected 3 components");
                O.m[0][0] = d[0]; O.m[1][1] = d[1]; O.m[2][2] = d[2];
            } else if (kind == "rotateX" || kind == "rotateY" || kind == "rotateZ") {
                std::vector<double> d = nums(*v);
                if (d.size() != 1) invalid(op + ": expected one angle");
                O = axisRotation(kind[6] - 'X', d[0]);
            } else if (kind.size() == 9 && kind.rfind("rotate", 0) == 0) {
                // rotateABC: about A first, then B, then C (p * R_A * R_B * R_C)
                std::vector<double> d = nums(*v);
                if (d.size() != 3) invalid(op + ": expected 3 angles");
                for (int k = 0; k < 3; ++k) {
                    const int ax = kind[6 + k] - 'X';
                    if (ax < 0 || ax > 2) invalid("bad rotation op " + op);
                    O = O * axisRotation(ax, d[(size_t)ax]);
                }
            } else if (kind == "orient") {
                std::vector<double> d = nums(*v);
                if (d.size() != 4) invalid(op + ": expected a quaternion (real, i, j, k)");
                O = quatMatrix(d[0], d[1], d[2], d[3]);
            } else if (kind == "transform") {
                if (v->kind != Value::Tuple || v->items.size() != 4) invalid(op + ": expected a 4x4 matrix");
                for (int i = 0; i < 4; ++i) {
                    std::vector<double> row = nums(v->items[(size_t)i]);
                    if (row.size() != 4) invalid(op + ": expected a 4x4 matrix");
                    for (int j = 0; j < 4; ++j) O.m[i][j] = row[(size_t)j];
                }
            } else {
                unsupported("xformOp kind '" + kind + "' on " + p.path);
            }
            if (invert) O = inverse(O);
            M = O * M;  // later ops in xformOpOrder apply to points first
        }
        return M;
    }
    M4 localToWorld(const Prim& p, double t) const {
        if (!p.parent) return M4::identity();
        bool reset = false;
        const M4 local = localXform(p, t, false, &reset);
        if (reset) return local;
        return local * localToWorld(*p.parent, t);
    }
};

// ----------------------------------------------------------------------------------------------
// scene under construction
// ----------------------------------------------------------------------------------------------
struct MeshData {
    std::string name;
    std::vector<V3> pos, nrm;
    std::vector<float> uv;  // 2 per vertex
    size_t tris() const { return pos.size() / 3; }
};

struct MaterialOut {
    ahip_material m;
    bool isStandard = false;
};

ahip_material lambertMat(float r, float g, float b) {
    ahip_material m;
    std::memset(&m, 0, sizeof m);
    m.type = AHIP_MAT_LAMBERT;
    m.albedo[0] = r; m.albedo[1] = g; m.albedo[2] = b;
    m.roughness_texture = -1;
    return m;
}

// StandardSurfaceMaterial::Params defaults (src/shading/StandardSurface.h:117-148)
ahip_material standardDefaults() {
    ahip_material m;
    std::memset(&m, 0, sizeof m);
    m.type = AHIP_MAT_STANDARD;
    m.albedo[0] = m.albedo[1] = m.albedo[2] = 0.8f;
    m.base = 1.0f;
    m.metalness = 0.0f;
    m.roughness = 0.5f;
    m.specular = 1.0f;
    m.specular_color[0] = m.specular_color[1] = m.specular_color[2] = 1.0f;
    m.specular_ior = 1.5f;
    m.coat = 0.0f;
    m.coat_roughness = 0.1f;
    m.transmission = 0.0f;
    m.emission_weight = 0.0f;
    m.roughness_texture = -1;
    return m;
}

// USDLoader.cpp:399-405 isGlassName
bool isGlassName(const std::string& name) {
    std::string l = name;
    for (char& c : l) c = (char)std::tolower((unsigned char)c);
    return l.find("glass") != std::string::npos || l.find("window") != std::string::npos ||
           l.find("glazing") != std::string::npos;
}

class Loader {
public:
    Loader(Stage& st, uint32_t w, uint32_t h) : S(st), W(w), H(h) {}

    Stage& S;
    uint32_t W, H;
    std::vector<MeshData> meshes;
    std::vector<int32_t> meshMat;
    std::vector<MaterialOut> mats;
    std::vector<ahip_light> lights;
    std::unordered_map<std::string, uint32_t> matByPath;
    std::unordered_map<uint32_t, uint32_t> matByDisplayColor;
    std::vector<const Prim*> cameras;
    bool hasCamera = false;
    float camFrom[3] = {0, 0, 0}, camAt[3] = {0, 0, 1}, camUp[3] = {0, 1, 0}, vfov = 0.f;
    std::string cameraPath;

    void run(const std::string& cameraOverride) {
        const Layer& L = S.L;
        for (const char* k : {"subLayers", "relocates"})
            if (L.meta.count(k)) unsupported(std::string("layer composition (") + k + ")");
        S.zUp = S.metaString(find(L.meta, "upAxis")) == "Z";
        const Value* st = find(L.meta, "startTimeCode");
        const Value* et = find(L.meta, "endTimeCode");
        if (st && et) {  // UsdStage::HasAuthoredTimeCodeRange
            asNumber(*st, S.startTime);
            asNumber(*et, S.endTime);
        }
        mats.push_back({lambertMat(0.5f, 0.5f, 0.5f), false});  // default material, index 0
        for (const auto& c : L.root.children) visit(*c);
        selectCamera(cameraOverride);
        autoEmissiveLights();
    }

private:
    static const Value* find(const std::map<std::string, Value>& m, const char* k) {
        auto it = m.find(k);
        return it == m.end() ? nullptr : &it->second;
    }

    void visit(const Prim& p) {
        if (p.spec != "def") return;  // over / class: not defined or abstract
        const Value* active = find(p.meta, "active");
        double a = 1.0;
        if (active && asNumber(*active, a) && a == 0.0) return;
        for (const char* k : {"references", "payload", "inherits", "specializes", "variants", "variantSets"})
            if (p.meta.count(k)) unsupported(std::string("composition arc '") + k + "' on " + p.path);
        if (p.type == "Mesh") loadMesh(p);
        else if (p.type == "RectLight") rectLight(p);
        else if (p.type == "SphereLight") sphereLight(p);
        else if (p.type == "DiskLight") diskLight(p);
        else if (p.type == "DistantLight") distantLight(p);
        else if (p.type == "DomeLight")
            unsupported(p.type + " " + p.path + " (an HDRI dome comes from the scene description only)");
        else if (p.type == "Camera") cameras.push_back(&p);
        for (const auto& c : p.children) visit(*c);
    }

    // --- materials ----------------------------------------------------------------------------
    // UsdShadeMaterialBindingAPI::ComputeBoundMaterial (all-purpose, direct bindings): the
    // nearest `material:binding` on the prim or an ancestor, unless an ancestor's binding is
    // marked strongerThanDescendants (the outermost such wins).
    const Prim* boundMaterial(const Prim& p) const {
        const Prim* found = nullptr;
        bool haveNearest = false;
        for (const Prim* q = &p; q && q->parent; q = q->parent) {
            const Property* r = S.prop(*q, "material:binding");
            if (!r || !r->isRel || r->targets.empty()) continue;
            std::string pp, pn;
            S.resolveTarget(*q, r->targets[0], pp, pn);
            const Prim* m = S.at(pp);
            if (m && !(m->type == "Material" && m->spec == "def")) m = nullptr;
            const bool stronger = S.metaString(find(r->meta, "bindMaterialAs")) == "strongerThanDescendants";
            if (!haveNearest || stronger) found = m;
            haveNearest = true;
        }
        return found;
    }

    // Follow a connection source through NodeGraph outputs to a Shader
    // (USDLoader.cpp:251-279 resolveShaderThroughNodeGraph), depth-limited.
    const Prim* shaderBehind(const Prim& owner, const std::string& target, int depth = 0) const {
        if (depth > 8) return nullptr;
        std::string pp, pn;
        S.resolveTarget(owner, target, pp, pn);
        const Prim* q = S.at(pp);
        if (!q) return nullptr;
        if (q->type != "NodeGraph" && q->type != "Material") return q->type == "Shader" ? q : nullptr;
        const Property* out = pn.empty() ? nullptr : S.prop(*q, pn);
        if (!out || !out->hasConnect || out->targets.empty()) {
            for (const auto& kv : q->props)
                if (kv.first.rfind("outputs:", 0) == 0 && kv.second.hasConnect && !kv.second.targets.empty()) {
                    out = &kv.second;
                    break;
                }
        }
        if (!out || !out->hasConnect || out->targets.empty()) return nullptr;
        return shaderBehind(*q, out->targets[0], depth + 1);
    }

    // A connected input that resolves to a texture with a file is what the reference turns
    // into a textured parameter (USDLoader.cpp:136-246, 281-316, 344-384); the device path
    // has no image textures for these inputs.
    void refuseTexture(const Prim& shader, const std::string& input) const {
        const Property* in = S.prop(shader, input);
        if (!in || !in->hasConnect || in->targets.empty()) return;
        const Prim* src = shaderBehind(shader, in->targets[0]);
        if (!src) return;
        const std::string id = S.getToken(*src, "info:id");
        if (id == "UsdUVTexture" || id.rfind("ND_image_", 0) == 0) {
            const Value* f = S.get(*src, "inputs:file");
            if (f && f->kind == Value::Asset && !f->str.empty())
                unsupported("texture-driven input " + shader.path + "." + input + " ('" + f->str + "')");
        }
    }
    float floatInput(const Prim& sh, const char* name, float def) const {
        refuseTexture(sh, name);
        return S.getFloat(sh, name, def);
    }
    void colorInput(const Prim& sh, const char* name, const float def[3], float out[3]) const {
        refuseTexture(sh, name);
        if (!S.getVec3(sh, name, out))
            for (int i = 0; i < 3; ++i) out[i] = def[i];
    }

    // USDLoader.cpp:526-779 resolveMaterial (OSL off, no OpenPBR network)
    MaterialOut resolveMaterial(const Prim& mat) const {
        for (const Prim* q : descendants(mat))
            if (q->type == "Shader" && S.getToken(*q, "info:id") == "ND_open_pbr_surface_surfaceshader")
                unsupported("OpenPBR (MaterialX) network in " + mat.path);
        const Prim* surface = nullptr;
        const Property* so = S.prop(mat, "outputs:surface");
        if (so && so->hasConnect && !so->targets.empty()) surface = shaderBehind(mat, so->targets[0]);
        if (!surface) {
            std::string blenderName = S.getToken(mat, "userProperties:blender:data_name");
            if (isGlassName(mat.name) || isGlassName(blenderName)) {  // USDLoader.cpp:386-397
                ahip_material g = standardDefaults();
                g.albedo[0] = g.albedo[1] = g.albedo[2] = 1.f;
                g.roughness = 0.f;
                g.metalness = 0.f;
                g.transmission = 1.f;
                g.specular_ior = 1.5f;
                g.specular = 1.f;
                return {g, true};
            }
            return {lambertMat(0.5f, 0.5f, 0.5f), false};
        }
        if (S.getToken(*surface, "info:id") != "UsdPreviewSurface") return {lambertMat(0.5f, 0.5f, 0.5f), false};
        const Prim& sh = *surface;
        ahip_material p = standardDefaults();
        const float half[3] = {0.5f, 0.5f, 0.5f}, black[3] = {0.f, 0.f, 0.f};
        colorInput(sh, "inputs:diffuseColor", half, p.albedo);
        p.roughness = floatInput(sh, "inputs:roughness", 1.0f);
        p.metalness = floatInput(sh, "inputs:metallic", 0.0f);
        const float opacity = floatInput(sh, "inputs:opacity", 1.0f);
        p.specular_ior = floatInput(sh, "inputs:ior", 1.5f);
        const float explicitT = floatInput(sh, "inputs:transmission", 0.0f);
        p.transmission = std::max(explicitT, 1.f - opacity);
        const float specDefault = (p.transmission > 0.001f && p.metalness < 0.001f) ? 1.0f
                                  : (p.metalness > 0.01f ? 0.f : 0.5f);
        p.specular = floatInput(sh, "inputs:specular", specDefault);
        p.coat = floatInput(sh, "inputs:clearcoat", 0.f);
        p.coat_roughness = floatInput(sh, "inputs:clearcoatRoughness", 0.1f);
        colorInput(sh, "inputs:emissiveColor", black, p.emission);
        const bool emits = p.emission[0] != 0.f || p.emission[1] != 0.f || p.emission[2] != 0.f;
        p.emission_weight = emits ? 1.f : 0.f;
        refuseTexture(sh, "inputs:normal");
        return {p, true};
    }
    static std::vector<const Prim*> descendants(const Prim& p) {
        std::vector<const Prim*> r;
        for (const auto& c : p.children) {
            r.push_back(c.get());
            auto d = descendants(*c);
            r.insert(r.end(), d.begin(), d.end());
        }
        return r;
    }
    uint32_t materialIndex(const Prim& mat) {
        auto it = matByPath.find(mat.path);
        if (it != matByPath.end()) return it->second;
        const uint32_t idx = (uint32_t)mats.size();
        mats.push_back(resolveMaterial(mat));
        matByPath[mat.path] = idx;
        return idx;
    }

    // --- meshes (USDLoader.cpp:886-1033, 1292-1430) --------------------------------------------
    void loadMesh(const Prim& p) {
        const M4 x0 = S.localToWorld(p, S.startTime);
        const M4 x1 = S.localToWorld(p, S.endTime);
        if (!(x0 == x1)) unsupported("animated transform of mesh " + p.path + " (motion blur)");
        std::vector<double> points = S.getFlat(p, "points", 3);
        std::vector<double> counts = S.getFlat(p, "faceVertexCounts", 1);
        std::vector<double> fvi = S.getFlat(p, "faceVertexIndices", 1);
        if (points.empty() || counts.empty() || fvi.empty()) return;  // skipped: no geometry
        const size_t np = points.size() / 3;
        std::vector<double> normals = S.getFlat(p, "normals", 3);
        const Property* na = S.prop(p, "normals");
        const bool nFaceVarying = na && S.metaString(find(na->meta, "interpolation")) == "faceVarying";
        std::vector<double> uvs;
        bool uvFaceVarying = false;
        for (const char* pv : {"primvars:st", "primvars:UVMap"}) {
            const Property* ua = S.prop(p, pv);
            if (!ua) continue;
            uvs = S.getFlat(p, pv, 2);
            uvFaceVarying = S.metaString(find(ua->meta, "interpolation")) == "faceVarying";
            std::vector<double> ui = S.getFlat(p, std::string(pv) + ":indices", 1);
            if (!ui.empty()) {  // UsdGeomPrimvar::ComputeFlattened
                std::vector<double> flat;
                flat.reserve(ui.size() * 2);
                for (double k : ui) {
                    if (k < 0 || (size_t)k * 2 + 1 >= uvs.size()) invalid("primvar index out of range on " + p.path);
                    flat.push_back(uvs[(size_t)k * 2]);
                    flat.push_back(uvs[(size_t)k * 2 + 1]);
                }
                uvs.swap(flat);
            }
            break;
        }
        const M4 invT = transposeOf(inverse(x0));

        MeshData md;
        md.name = p.name;
        std::vector<uint32_t> faceTriStart, faceTriCount;
        size_t faceStart = 0;
        auto pointAt = [&](size_t i) {
            if (i >= np) invalid("face vertex index out of range on " + p.path);
            return transformPoint(x0, (double)(float)points[i * 3], (double)(float)points[i * 3 + 1],
                                  (double)(float)points[i * 3 + 2], S.zUp);
        };
        for (double cnt : counts) {
            const long nv = (long)cnt;
            if (nv < 0 || faceStart + (size_t)std::max(nv, 0L) > fvi.size())
                invalid("faceVertexCounts exceed faceVertexIndices on " + p.path);
            faceTriStart.push_back((uint32_t)md.tris());
            for (long tri = 0; tri < nv - 2; ++tri) {
                const size_t f[3] = {faceStart, faceStart + (size_t)tri + 1, faceStart + (size_t)tri + 2};
                long vi[3];
                for (int k = 0; k < 3; ++k) vi[k] = (long)fvi[f[k]];
                V3 P[3];
                for (int k = 0; k < 3; ++k) {
                    if (vi[k] < 0) invalid("negative face vertex index on " + p.path);
                    P[k] = pointAt((size_t)vi[k]);
                    md.pos.push_back(P[k]);
                }
                for (int k = 0; k < 3; ++k) {
                    const size_t ni = nFaceVarying ? f[k] : (size_t)vi[k];
                    if (!normals.empty() && ni < normals.size() / 3) {
                        md.nrm.push_back(transformNormal(invT, (float)normals[ni * 3], (float)normals[ni * 3 + 1],
                                                         (float)normals[ni * 3 + 2], S.zUp));
                    } else {
                        md.nrm.push_back(safeNormalize(cross(P[1] - P[0], P[2] - P[0])));
                    }
                }
                for (int k = 0; k < 3; ++k) {
                    const size_t ui = uvFaceVarying ? f[k] : (size_t)vi[k];
                    if (!uvs.empty() && ui < uvs.size() / 2) {
                        md.uv.push_back((float)uvs[ui * 2]);
                        md.uv.push_back((float)uvs[ui * 2 + 1]);
                    } else {
                        md.uv.push_back(0.f);
                        md.uv.push_back(0.f);
                    }
                }
            }
            faceTriCount.push_back((uint32_t)md.tris() - faceTriStart.back());
            faceStart += (size_t)std::max(nv, 0L);
        }
        if (md.pos.empty()) return;
        const uint32_t meshID = (uint32_t)meshes.size();
        meshes.push_back(std::move(md));
        meshMat.push_back(-1);

        // GeomSubsets with their own bound material become meshes of their own
        std::vector<char> covered(faceTriStart.size(), 0);
        for (const auto& c : p.children) {
            const Prim& sub = *c;
            if (sub.type != "GeomSubset" || sub.spec != "def") continue;
            const std::string et = S.getToken(sub, "elementType");
            if (!et.empty() && et != "face") continue;
            std::vector<double> faces = S.getFlat(sub, "indices", 1);
            if (faces.empty()) continue;
            const Prim* sm = boundMaterial(sub);
            if (!sm) continue;
            const uint32_t smIdx = materialIndex(*sm);
            MeshData part;
            part.name = meshes[meshID].name;
            for (double fd : faces) {
                const long fi = (long)fd;
                if (fi < 0 || fi >= (long)faceTriStart.size()) continue;
                appendTris(meshes[meshID], faceTriStart[(size_t)fi], faceTriCount[(size_t)fi], part);
            }
            if (part.pos.empty()) continue;
            meshes.push_back(std::move(part));
            meshMat.push_back((int32_t)smIdx);
            for (double fd : faces) {
                const long fi = (long)fd;
                if (fi >= 0 && fi < (long)covered.size()) covered[(size_t)fi] = 1;
            }
        }

        uint32_t matIdx = 0;
        if (const Prim* bm = boundMaterial(p)) {
            matIdx = materialIndex(*bm);
        } else {
            std::vector<double> dc = S.getFlat(p, "primvars:displayColor", 3);
            if (!dc.empty()) {
                const float c[3] = {(float)dc[0], (float)dc[1], (float)dc[2]};
                const uint32_t key = ((uint32_t)(c[0] * 255.f + 0.5f) << 16) | ((uint32_t)(c[1] * 255.f + 0.5f) << 8) |
                                     (uint32_t)(c[2] * 255.f + 0.5f);
                auto it = matByDisplayColor.find(key);
                if (it != matByDisplayColor.end()) {
                    matIdx = it->second;
                } else {
                    matIdx = (uint32_t)mats.size();
                    mats.push_back({lambertMat(c[0], c[1], c[2]), false});
                    matByDisplayColor[key] = matIdx;
                }
            }
        }
        meshMat[meshID] = (int32_t)matIdx;

        if (std::find(covered.begin(), covered.end(), 1) != covered.end()) {
            MeshData residual;
            residual.name = meshes[meshID].name;
            for (size_t fi = 0; fi < covered.size(); ++fi)
                if (!covered[fi]) appendTris(meshes[meshID], faceTriStart[fi], faceTriCount[fi], residual);
            meshes[meshID] = std::move(residual);
        }
    }
    static void appendTris(const MeshData& src, uint32_t first, uint32_t count, MeshData& dst) {
        for (uint32_t t = 0; t < count; ++t) {
            const size_t b = (size_t)(first + t) * 3;
            if (b + 2 >= src.pos.size()) continue;
            for (size_t k = 0; k < 3; ++k) {
                dst.pos.push_back(src.pos[b + k]);
                dst.nrm.push_back(src.nrm[b + k]);
                dst.uv.push_back(src.uv[(b + k) * 2]);
                dst.uv.push_back(src.uv[(b + k) * 2 + 1]);
            }
        }
    }

    // --- lights (USDLoader.cpp:1433-1545) -----------------------------------------------------
    void pushArea(V3 c, V3 u, V3 v, const float Le[3]) {
        ahip_light l;
        std::memset(&l, 0, sizeof l);
        l.type = AHIP_LIGHT_AREA;
        const V3 a[3] = {c, u, v};
        float* dst[3] = {l.center, l.u_axis, l.v_axis};
        for (int i = 0; i < 3; ++i) {
            dst[i][0] = a[i].x;
            dst[i][1] = a[i].y;
            dst[i][2] = a[i].z;
        }
        for (int i = 0; i < 3; ++i) l.Le[i] = Le[i];
        l.intensity = 1.f;
        l.rot[0] = l.rot[4] = l.rot[8] = 1.f;
        lights.push_back(l);
    }
    void lightColor(const Prim& p, float intensity, float Le[3]) const {
        float color[3] = {1.f, 1.f, 1.f};
        S.getVec3(p, "inputs:color", color);
        const bool normalize = S.getBool(p, "inputs:normalize", false);
        const float leScale = normalize ? (1.f / 3.14159265f) : 1.f;
        for (int i = 0; i < 3; ++i) Le[i] = color[i] * intensity * leScale;
    }
    float exposedIntensity(const Prim& p) const {  // USDLoader.cpp:319-324
        return S.getFloat(p, "inputs:intensity", 1.f) * std::pow(2.f, S.getFloat(p, "inputs:exposure", 0.f));
    }
    void rectLight(const Prim& p) {
        const M4 x = S.localToWorld(p, S.startTime);
        const float w = S.getFloat(p, "inputs:width", 1.f), h = S.getFloat(p, "inputs:height", 1.f);
        float Le[3];
        lightColor(p, S.getFloat(p, "inputs:intensity", 1.f), Le);
        const V3 c = transformPoint(x, 0, 0, 0, S.zUp);
        const V3 u = transformPoint(x, w * 0.5, 0, 0, S.zUp) - c;
        const V3 v = transformPoint(x, 0, h * 0.5, 0, S.zUp) - c;
        pushArea(c, u, v, Le);
    }
    // USDLoader.cpp:1553-1575: the light shines along its local -Z, so the direction toward it is the
    // local +Z in world space; radiance colour x intensity x 2^exposure; a hard (delta) direction.
    void distantLight(const Prim& p) {
        const M4 x = S.localToWorld(p, S.startTime);
        const V3 o = transformPoint(x, 0, 0, 0, S.zUp), z = transformPoint(x, 0, 0, 1, S.zUp);
        const V3 d = z - o;
        const float len = d.length();
        float color[3] = {1.f, 1.f, 1.f};
        S.getVec3(p, "inputs:color", color);
        const float intensity = exposedIntensity(p);
        ahip_light l;
        std::memset(&l, 0, sizeof l);
        l.type = AHIP_LIGHT_DIRECTIONAL;
        const V3 n = len > 1e-8f ? d * (1.f / len) : V3{0.f, 0.f, 1.f};
        l.direction[0] = n.x;
        l.direction[1] = n.y;
        l.direction[2] = n.z;
        for (int i = 0; i < 3; ++i) l.Le[i] = color[i] * intensity;
        l.intensity = 1.f;
        l.rot[0] = l.rot[4] = l.rot[8] = 1.f;
        l.half_angle = 0.f;
        lights.push_back(l);
    }
    void sphereLight(const Prim& p) {
        const M4 x = S.localToWorld(p, S.startTime);
        const float r = S.getFloat(p, "inputs:radius", 0.5f);
        float Le[3];
        lightColor(p, S.getFloat(p, "inputs:intensity", 1.f), Le);
        pushArea(transformPoint(x, 0, 0, 0, S.zUp), {r, 0.f, 0.f}, {0.f, 0.f, r}, Le);
    }
    void diskLight(const Prim& p) {
        const M4 x = S.localToWorld(p, S.startTime);
        const float r = S.getFloat(p, "inputs:radius", 0.5f);
        float Le[3];
        lightColor(p, exposedIntensity(p), Le);
        const V3 c = transformPoint(x, 0, 0, 0, S.zUp);
        const V3 xe = transformPoint(x, r, 0, 0, S.zUp) - c;
        const V3 ye = transformPoint(x, 0, r, 0, S.zUp) - c;
        const float half = r * 0.8862f;  // square of the disk's area: r * sqrt(pi) / 2
        auto lsq = [](V3 a) { return a.x * a.x + a.y * a.y + a.z * a.z; };
        const V3 u = lsq(xe) > 1e-12f ? xe * (half / xe.length()) : V3{half, 0.f, 0.f};
        const V3 v = lsq(ye) > 1e-12f ? ye * (half / ye.length()) : V3{0.f, 0.f, half};
        pushArea(c, u, v, Le);
    }

    // USDLoader.cpp:1762-1826: every standard_surface mesh with emission becomes an area light
    // spanning its bounding box, the normal along the box's thinnest axis.
    void autoEmissiveLights() {
        for (size_t mid = 0; mid < meshes.size(); ++mid) {
            const int32_t mi = meshMat[mid];
            if (mi < 0 || !mats[(size_t)mi].isStandard) continue;
            const ahip_material& m = mats[(size_t)mi].m;
            if (m.emission_weight <= 0.f) continue;
            const float Le[3] = {m.emission[0] * m.emission_weight, m.emission[1] * m.emission_weight,
                                 m.emission[2] * m.emission_weight};
            if (Le[0] == 0.f && Le[1] == 0.f && Le[2] == 0.f) continue;
            V3 lo{std::numeric_limits<float>::infinity(), std::numeric_limits<float>::infinity(),
                  std::numeric_limits<float>::infinity()};
            V3 hi{-lo.x, -lo.y, -lo.z};
            for (const V3& q : meshes[mid].pos) {
                lo = {std::min(lo.x, q.x), std::min(lo.y, q.y), std::min(lo.z, q.z)};
                hi = {std::max(hi.x, q.x), std::max(hi.y, q.y), std::max(hi.z, q.z)};
            }
            if (!(lo.x <= hi.x && lo.y <= hi.y && lo.z <= hi.z)) continue;
            const V3 d = hi - lo;
            const V3 c = (lo + hi) * 0.5f;
            V3 u, v;
            if (d.x <= d.y && d.x <= d.z) {
                u = {0.f, d.y * 0.5f, 0.f};
                v = {0.f, 0.f, d.z * 0.5f};
            } else if (d.y <= d.x && d.y <= d.z) {
                u = {d.x * 0.5f, 0.f, 0.f};
                v = {0.f, 0.f, d.z * 0.5f};
            } else {
                u = {d.x * 0.5f, 0.f, 0.f};
                v = {0.f, d.y * 0.5f, 0.f};
            }
            pushArea(c, u, v, Le);
        }
    }

    // --- camera (USDLoader.cpp:1084-1161, 1635-1722) -------------------------------------------
    void selectCamera(const std::string& overridePath) {
        const Prim* cam = nullptr;
        if (!overridePath.empty()) {
            const Prim* q = S.at(overridePath);
            if (q && q->type == "Camera" && q->spec == "def") cam = q;
        }
        if (!cam) {
            std::vector<const Prim*> all;
            for (const auto& c : S.L.root.children) collectDefined(*c, all);
            for (const Prim* q : all) {
                if (q->type != "RenderSettings") continue;
                const Property* r = S.prop(*q, "camera");
                if (!r || !r->isRel || r->targets.empty()) continue;
                std::string pp, pn;
                S.resolveTarget(*q, r->targets[0], pp, pn);
                const Prim* c = S.at(pp);
                if (c && c->type == "Camera" && c->spec == "def") {
                    cam = c;
                    break;
                }
            }
        }
        if (!cam && !cameras.empty()) cam = cameras.front();
        if (!cam) return;
        buildCamera(*cam);
    }
    void collectDefined(const Prim& p, std::vector<const Prim*>& out) const {
        if (p.spec != "def") return;
        const Value* active = find(p.meta, "active");
        double a = 1.0;
        if (active && asNumber(*active, a) && a == 0.0) return;
        out.push_back(&p);
        for (const auto& c : p.children) collectDefined(*c, out);
    }
    void buildCamera(const Prim& p) {
        bool reset = false;
        const M4 local = S.localXform(p, 0.0, true, &reset);  // the camera's own ops at the default time
        const M4 full = local * S.localToWorld(*p.parent, S.startTime);
        const V3 origin = transformPoint(full, 0, 0, 0, S.zUp);
        const V3 target = transformPoint(full, 0, 0, -1, S.zUp);
        const V3 upWS = transformPoint(full, 0, 1, 0, S.zUp);
        const V3 up = safeNormalize(upWS - origin);
        const float focalLen = S.getFloat(p, "focalLength", 50.f);
        const float hAperture = S.getFloat(p, "horizontalAperture", 20.955f);
        const float fStop = S.getFloat(p, "fStop", 0.f);
        const float focusDistance = S.getFloat(p, "focusDistance", 0.f);
        if (fStop > 0.f && focusDistance > 0.f)
            unsupported("thin-lens camera " + p.path + " (fStop and focusDistance set; the device camera is a pinhole)");
        const float aspect = (float)W / (float)H;
        const float vAperture = hAperture / aspect;
        const float vfovRad = 2.f * std::atan(vAperture * 0.5f / focalLen);
        vfov = vfovRad * 180.f / 3.14159265f;
        camFrom[0] = origin.x; camFrom[1] = origin.y; camFrom[2] = origin.z;
        camAt[0] = target.x; camAt[1] = target.y; camAt[2] = target.z;
        camUp[0] = up.x; camUp[1] = up.y; camUp[2] = up.z;
        hasCamera = true;
        cameraPath = p.path;
    }
};

}  // namespace

// ------------------------------------------------------------------------------------------------
// C ABI
// ------------------------------------------------------------------------------------------------
struct ahip_usd_scene {
    std::vector<float> positions, normals, uvs;
    std::vector<uint32_t> indices, vo, io, nc, uc;
    std::vector<int32_t> mm;
    std::vector<ahip_material> materials;
    std::vector<ahip_light> lights;
    std::vector<std::string> meshNames;
    bool hasCamera = false;
    float from[3], at[3], up[3], vfov = 0.f;
    uint32_t width = 0, height = 0;
    std::string cameraPath;
    ahip_scene_desc desc;
};

namespace {
void setErr(char* err, size_t n, const std::string& m) {
    if (!err || n == 0) return;
    std::snprintf(err, n, "%s", m.c_str());
}

ahip_usd_scene* build(const std::string& text, uint32_t w, uint32_t h, const std::string& cam) {
    if (text.rfind("#usda", 0) != 0) invalid("not a text USD layer (no '#usda' header)");
    Stage st;
    Parser(text).parseLayer(st.L);
    Loader ld(st, w, h);
    ld.run(cam);

    auto s = std::make_unique<ahip_usd_scene>();
    s->vo.push_back(0);
    s->io.push_back(0);
    for (size_t i = 0; i < ld.meshes.size(); ++i) {
        const MeshData& m = ld.meshes[i];
        const uint32_t nv = (uint32_t)m.pos.size();
        for (uint32_t k = 0; k < nv; ++k) {
            s->positions.insert(s->positions.end(), {m.pos[k].x, m.pos[k].y, m.pos[k].z});
            s->normals.insert(s->normals.end(), {m.nrm[k].x, m.nrm[k].y, m.nrm[k].z});
            s->uvs.insert(s->uvs.end(), {m.uv[2 * k], m.uv[2 * k + 1]});
            s->indices.push_back(k);
        }
        s->vo.push_back(s->vo.back() + nv);
        s->io.push_back(s->io.back() + nv);
        s->nc.push_back(nv);
        s->uc.push_back(nv);
        s->mm.push_back(ld.meshMat[i]);
        s->meshNames.push_back(m.name);
    }
    for (const MaterialOut& m : ld.mats) s->materials.push_back(m.m);
    s->lights = ld.lights;
    // the lights' scene sphere once every mesh is loaded (USDLoader.cpp:1722-1744): the bounds'
    // centroid and 1.5 x the half diagonal
    if (!s->positions.empty()) {
        float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
        for (size_t i = 0; i < s->positions.size(); i += 3)
            for (int k = 0; k < 3; ++k) {
                lo[k] = std::min(lo[k], s->positions[i + k]);
                hi[k] = std::max(hi[k], s->positions[i + k]);
            }
        const V3 diag{hi[0] - lo[0], hi[1] - lo[1], hi[2] - lo[2]};
        const float radius = diag.length() * 0.5f * 1.5f;
        for (ahip_light& l : s->lights)
            if (l.type == AHIP_LIGHT_DIRECTIONAL || l.type == AHIP_LIGHT_DOME) {
                for (int k = 0; k < 3; ++k) l.scene_center[k] = (lo[k] + hi[k]) * 0.5f;
                l.scene_radius = radius;
            }
    }
    s->hasCamera = ld.hasCamera;
    std::memcpy(s->from, ld.camFrom, sizeof s->from);
    std::memcpy(s->at, ld.camAt, sizeof s->at);
    std::memcpy(s->up, ld.camUp, sizeof s->up);
    s->vfov = ld.vfov;
    s->width = w;
    s->height = h;
    s->cameraPath = ld.cameraPath;

    ahip_scene_desc& d = s->desc;
    std::memset(&d, 0, sizeof d);
    d.num_meshes = (uint32_t)ld.meshes.size();
    d.mesh_vertex_offset = s->vo.data();
    d.mesh_index_offset = s->io.data();
    d.mesh_normal_count = s->nc.data();
    d.mesh_uv_count = s->uc.data();
    d.mesh_material = s->mm.data();
    d.positions = s->positions.data();
    d.normals = s->normals.data();
    d.uvs = s->uvs.data();
    d.indices = s->indices.data();
    d.num_vertices = (uint32_t)(s->positions.size() / 3);
    d.num_indices = (uint32_t)s->indices.size();
    d.num_materials = (uint32_t)s->materials.size();
    d.materials = s->materials.data();
    d.num_lights = (uint32_t)s->lights.size();
    d.lights = s->lights.data();
    d.env_light = -1;  // USDLoader.cpp:1726: black environment
    return s.release();
}
}  // namespace

extern "C" {

int ahip_usd_load_text(const char* text, size_t len, uint32_t film_width, uint32_t film_height,
                       const char* camera_path, ahip_usd_scene** out, char* err, size_t err_len) {
    if (out) *out = nullptr;
    if (!text || !out || film_width == 0 || film_height == 0) {
        setErr(err, err_len, "ahip_usd_load_text: null argument or empty film");
        return AHIP_E_INVALID_ARG;
    }
    try {
        *out = build(std::string(text, len), film_width, film_height, camera_path ? camera_path : "");
        setErr(err, err_len, "");
        return AHIP_OK;
    } catch (const LoadError& e) {
        setErr(err, err_len, e.what());
        return e.status;
    } catch (const std::bad_alloc&) {
        setErr(err, err_len, "out of host memory while loading USD");
        return AHIP_E_OUT_OF_MEMORY;
    } catch (const std::exception& e) {
        setErr(err, err_len, e.what());
        return AHIP_E_INVALID_ARG;
    }
}

int ahip_usd_load(const char* path, uint32_t film_width, uint32_t film_height, const char* camera_path,
                  ahip_usd_scene** out, char* err, size_t err_len) {
    if (out) *out = nullptr;
    if (!path) {
        setErr(err, err_len, "ahip_usd_load: null path");
        return AHIP_E_INVALID_ARG;
    }
    std::ifstream f(path, std::ios::binary);
    if (!f) {
        setErr(err, err_len, std::string("cannot open '") + path + "'");
        return AHIP_E_INVALID_ARG;
    }
    std::stringstream ss;
    ss << f.rdbuf();
    const std::string text = ss.str();
    return ahip_usd_load_text(text.data(), text.size(), film_width, film_height, camera_path, out, err, err_len);
}

void ahip_usd_free(ahip_usd_scene* s) { delete s; }

const ahip_scene_desc* ahip_usd_scene_desc(const ahip_usd_scene* s) { return s ? &s->desc : nullptr; }

int ahip_usd_camera(const ahip_usd_scene* s, float from[3], float at[3], float up[3], float* vfov_degrees,
                    ahip_camera* cam) {
    if (!s) return AHIP_E_INVALID_ARG;
    if (!s->hasCamera) return AHIP_E_NOT_PREPARED;
    for (int i = 0; i < 3; ++i) {
        if (from) from[i] = s->from[i];
        if (at) at[i] = s->at[i];
        if (up) up[i] = s->up[i];
    }
    if (vfov_degrees) *vfov_degrees = s->vfov;
    if (cam) return ahip_camera_pinhole(s->from, s->at, s->up, s->vfov, s->width, s->height, cam);
    return AHIP_OK;
}

const char* ahip_usd_mesh_name(const ahip_usd_scene* s, uint32_t mesh) {
    if (!s || mesh >= s->meshNames.size()) return nullptr;
    return s->meshNames[mesh].c_str();
}

}  // extern "C"
