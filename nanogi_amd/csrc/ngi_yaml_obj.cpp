// YAML-subset and Wavefront-OBJ readers for the scene loader.
//
// The reference reads scenes with yaml-cpp (include/nanogi/rt.hpp:1525) and
// meshes with Assimp (rt.hpp:1640-1727).  Neither library is available, so
// both are replaced by small readers covering what schema.yml and the
// reference's OBJ inputs use: block maps and sequences, flow sequences and
// maps, plain and quoted scalars, comments; OBJ v/vt/vn/f with polygon fan
// triangulation, first object and first material run only.
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <sstream>
#include <tuple>

#include "ngi_internal.h"

namespace ngi {

// ---------------------------------------------------------------------------
// Error reporting
// ---------------------------------------------------------------------------
static thread_local std::string g_last_error;

int fail(int status, const char* fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_last_error = buf;
    return status;
}

void clear_error() { g_last_error.clear(); }

// ---------------------------------------------------------------------------
// YAML subset
// ---------------------------------------------------------------------------
const YNode* YNode::get(const char* key) const {
    if (kind != Map) return nullptr;
    for (const auto& kv : map)
        if (kv.first == key) return &kv.second;
    return nullptr;
}

namespace {

struct YLine {
    int indent;
    std::string text;
    int lineno;
};

std::string rtrim(const std::string& s) {
    size_t e = s.find_last_not_of(" \t\r\n");
    return e == std::string::npos ? std::string() : s.substr(0, e + 1);
}

std::string trim(const std::string& s) {
    size_t b = s.find_first_not_of(" \t\r\n");
    if (b == std::string::npos) return std::string();
    size_t e = s.find_last_not_of(" \t\r\n");
    return s.substr(b, e - b + 1);
}

// Remove a trailing comment: '#' at line start or after whitespace, outside quotes.
std::string strip_comment(const std::string& s) {
    char q = 0;
    for (size_t i = 0; i < s.size(); ++i) {
        char c = s[i];
        if (q) {
            if (c == q) q = 0;
            continue;
        }
        if (c == '\'' || c == '"') {
            q = c;
            continue;
        }
        if (c == '#' && (i == 0 || s[i - 1] == ' ' || s[i - 1] == '\t')) return s.substr(0, i);
    }
    return s;
}

struct YParser {
    std::vector<YLine> lines;
    std::string err;
    int errline = 0;

    bool error(int line, const char* msg) {
        if (err.empty()) {
            err = msg;
            errline = line;
        }
        return false;
    }

    static std::string unquote(const std::string& s) {
        if (s.size() >= 2 && ((s[0] == '\'' && s.back() == '\'') || (s[0] == '"' && s.back() == '"'))) {
            std::string r;
            const char q = s[0];
            for (size_t i = 1; i + 1 < s.size(); ++i) {
                if (q == '\'' && s[i] == '\'' && i + 2 < s.size() && s[i + 1] == '\'') {
                    r += '\'';
                    ++i;
                } else if (q == '"' && s[i] == '\\' && i + 2 < s.size()) {
                    char n = s[++i];
                    r += (n == 'n') ? '\n' : (n == 't') ? '\t' : n;
                } else {
                    r += s[i];
                }
            }
            return r;
        }
        return s;
    }

    // Find the ':' that separates a mapping key (outside quotes/brackets),
    // followed by space or end of line.
    static long find_key_colon(const std::string& s) {
        char q = 0;
        int depth = 0;
        for (size_t i = 0; i < s.size(); ++i) {
            char c = s[i];
            if (q) {
                if (c == q) q = 0;
                continue;
            }
            if (c == '\'' || c == '"') q = c;
            else if (c == '[' || c == '{') ++depth;
            else if (c == ']' || c == '}') --depth;
            else if (c == ':' && depth == 0 && (i + 1 == s.size() || s[i + 1] == ' ' || s[i + 1] == '\t'))
                return (long)i;
        }
        return -1;
    }

    // Flow content: [a, b], {k: v}, or a scalar.
    bool parse_flow(const std::string& s, size_t& i, YNode& out, int line) {
        while (i < s.size() && (s[i] == ' ' || s[i] == '\t')) ++i;
        if (i >= s.size()) {
            out.kind = YNode::Null;
            return true;
        }
        if (s[i] == '[') {
            out.kind = YNode::Seq;
            ++i;
            for (;;) {
                while (i < s.size() && (s[i] == ' ' || s[i] == '\t')) ++i;
                if (i >= s.size()) return error(line, "unterminated flow sequence");
                if (s[i] == ']') {
                    ++i;
                    return true;
                }
                YNode item;
                if (!parse_flow_item(s, i, item, line, "],")) return false;
                out.seq.push_back(std::move(item));
                while (i < s.size() && (s[i] == ' ' || s[i] == '\t')) ++i;
                if (i < s.size() && s[i] == ',') ++i;
                else if (i < s.size() && s[i] == ']') continue;
                else return error(line, "expected ',' or ']' in flow sequence");
            }
        }
        if (s[i] == '{') {
            out.kind = YNode::Map;
            ++i;
            for (;;) {
                while (i < s.size() && (s[i] == ' ' || s[i] == '\t')) ++i;
                if (i >= s.size()) return error(line, "unterminated flow mapping");
                if (s[i] == '}') {
                    ++i;
                    return true;
                }
                size_t c = s.find(':', i);
                if (c == std::string::npos) return error(line, "expected ':' in flow mapping");
                std::string key = unquote(trim(s.substr(i, c - i)));
                i = c + 1;
                YNode val;
                if (!parse_flow_item(s, i, val, line, "},")) return false;
                out.map.emplace_back(key, std::move(val));
                while (i < s.size() && (s[i] == ' ' || s[i] == '\t')) ++i;
                if (i < s.size() && s[i] == ',') ++i;
                else if (i < s.size() && s[i] == '}') continue;
                else return error(line, "expected ',' or '}' in flow mapping");
            }
        }
        // Plain or quoted scalar to end of string.
        out.kind = YNode::Scalar;
        out.scalar = unquote(trim(s.substr(i)));
        i = s.size();
        return true;
    }

    bool parse_flow_item(const std::string& s, size_t& i, YNode& out, int line, const char* stops) {
        while (i < s.size() && (s[i] == ' ' || s[i] == '\t')) ++i;
        if (i < s.size() && (s[i] == '[' || s[i] == '{')) return parse_flow(s, i, out, line);
        size_t b = i;
        char q = 0;
        if (i < s.size() && (s[i] == '\'' || s[i] == '"')) {
            q = s[i++];
            while (i < s.size() && s[i] != q) ++i;
            if (i >= s.size()) return error(line, "unterminated quoted scalar");
            ++i;
        } else {
            while (i < s.size() && !strchr(stops, s[i])) ++i;
        }
        out.kind = YNode::Scalar;
        out.scalar = unquote(trim(s.substr(b, i - b)));
        return true;
    }

    bool is_seq_item(const std::string& t) { return t == "-" || (t.size() >= 2 && t[0] == '-' && (t[1] == ' ' || t[1] == '\t')); }

    // Parse the block starting at lines[i] whose indent defines the block.
    bool parse_block(size_t& i, YNode& out) {
        if (i >= lines.size()) {
            out.kind = YNode::Null;
            return true;
        }
        const int indent = lines[i].indent;
        if (is_seq_item(lines[i].text)) {
            out.kind = YNode::Seq;
            while (i < lines.size() && lines[i].indent == indent && is_seq_item(lines[i].text)) {
                std::string rest = lines[i].text.substr(1);
                size_t off = 1;
                while (off < lines[i].text.size() && (lines[i].text[off] == ' ' || lines[i].text[off] == '\t')) ++off;
                rest = lines[i].text.substr(off);
                YNode item;
                if (rest.empty()) {
                    ++i;
                    if (i < lines.size() && lines[i].indent > indent) {
                        if (!parse_block(i, item)) return false;
                    }
                } else {
                    // Re-read the rest of the line as a block at the column it starts.
                    lines[i].indent = indent + (int)off;
                    lines[i].text = rest;
                    if (!parse_block(i, item)) return false;
                }
                out.seq.push_back(std::move(item));
            }
            if (i < lines.size() && lines[i].indent > indent)
                return error(lines[i].lineno, "bad indentation after sequence item");
            return true;
        }
        // Mapping, or a lone scalar/flow node.
        long c = find_key_colon(lines[i].text);
        if (c < 0) {
            size_t p = 0;
            int line = lines[i].lineno;
            std::string t = lines[i].text;
            ++i;
            if (i < lines.size() && lines[i].indent > indent) return error(lines[i].lineno, "unexpected indentation");
            return parse_flow(t, p, out, line);
        }
        out.kind = YNode::Map;
        while (i < lines.size() && lines[i].indent == indent) {
            const std::string& t = lines[i].text;
            if (is_seq_item(t)) return error(lines[i].lineno, "sequence item inside mapping");
            long k = find_key_colon(t);
            if (k < 0) return error(lines[i].lineno, "expected 'key: value'");
            std::string key = unquote(trim(t.substr(0, (size_t)k)));
            std::string rest = trim(t.substr((size_t)k + 1));
            int line = lines[i].lineno;
            ++i;
            YNode val;
            if (!rest.empty()) {
                size_t p = 0;
                if (!parse_flow(rest, p, val, line)) return false;
            } else if (i < lines.size() && lines[i].indent > indent) {
                if (!parse_block(i, val)) return false;
            } else if (i < lines.size() && lines[i].indent == indent && is_seq_item(lines[i].text)) {
                // "key:\n- a\n- b" (sequence at the key's indentation)
                if (!parse_block(i, val)) return false;
            }
            for (const auto& kv : out.map)
                if (kv.first == key) return error(line, "duplicate mapping key");
            out.map.emplace_back(key, std::move(val));
        }
        if (i < lines.size() && lines[i].indent > indent) return error(lines[i].lineno, "bad indentation");
        return true;
    }
};

}  // namespace

int yaml_parse(const std::string& text, YNode& out) {
    YParser p;
    std::istringstream in(text);
    std::string raw;
    int lineno = 0;
    while (std::getline(in, raw)) {
        ++lineno;
        for (char& ch : raw)
            if (ch == '\t') ch = ' ';
        std::string s = rtrim(strip_comment(raw));
        if (trim(s).empty()) continue;
        if (s == "---" || s == "...") continue;
        size_t ind = s.find_first_not_of(' ');
        p.lines.push_back({(int)ind, s.substr(ind), lineno});
    }
    out = YNode();
    if (p.lines.empty()) return NGI_OK;
    size_t i = 0;
    if (p.lines[0].indent != 0) return fail(NGI_ERR_PARSE, "YAML: line %d: document must start at column 0", p.lines[0].lineno);
    if (!p.parse_block(i, out) || i != p.lines.size()) {
        if (p.err.empty() && i < p.lines.size()) return fail(NGI_ERR_PARSE, "YAML: line %d: unexpected content", p.lines[i].lineno);
        return fail(NGI_ERR_PARSE, "YAML: line %d: %s", p.errline, p.err.c_str());
    }
    return NGI_OK;
}

// ---------------------------------------------------------------------------
// OBJ
// ---------------------------------------------------------------------------
int obj_load(const std::string& path, bool smooth_normals, ObjMesh& out) {
    std::ifstream f(path);
    if (!f) return fail(NGI_ERR_IO, "cannot open mesh '%s'", path.c_str());
    std::vector<double> v, vt, vn;
    std::vector<std::tuple<long, long, long>> corners;  // per triangle corner: (v, vt, vn) 0-based, -1 absent
    bool seen_faces = false, stop = false;
    std::string cur_mtl;
    bool mtl_set = false;
    std::string line;
    int lineno = 0;
    while (!stop && std::getline(f, line)) {
        ++lineno;
        size_t h = line.find('#');
        if (h != std::string::npos) line.resize(h);
        std::istringstream ls(line);
        std::string tag;
        if (!(ls >> tag)) continue;
        if (tag == "v") {
            double x, y, z;
            if (!(ls >> x >> y >> z)) return fail(NGI_ERR_PARSE, "OBJ %s:%d: bad vertex", path.c_str(), lineno);
            v.insert(v.end(), {x, y, z});
        } else if (tag == "vt") {
            double a = 0, b = 0;
            if (!(ls >> a)) return fail(NGI_ERR_PARSE, "OBJ %s:%d: bad texcoord", path.c_str(), lineno);
            ls >> b;
            vt.insert(vt.end(), {a, b});
        } else if (tag == "vn") {
            double x, y, z;
            if (!(ls >> x >> y >> z)) return fail(NGI_ERR_PARSE, "OBJ %s:%d: bad normal", path.c_str(), lineno);
            vn.insert(vn.end(), {x, y, z});
        } else if (tag == "o") {
            if (seen_faces) stop = true;
        } else if (tag == "usemtl") {
            std::string name;
            ls >> name;
            if (seen_faces && mtl_set && name != cur_mtl) stop = true;
            cur_mtl = name;
            mtl_set = true;
        } else if (tag == "f") {
            std::vector<std::tuple<long, long, long>> poly;
            std::string tok;
            while (ls >> tok) {
                long idx[3] = {0, 0, 0};
                bool has[3] = {false, false, false};
                size_t p = 0;
                for (int k = 0; k < 3 && p <= tok.size(); ++k) {
                    size_t e = tok.find('/', p);
                    std::string part = tok.substr(p, e == std::string::npos ? std::string::npos : e - p);
                    if (!part.empty()) {
                        char* endp = nullptr;
                        idx[k] = strtol(part.c_str(), &endp, 10);
                        if (*endp) return fail(NGI_ERR_PARSE, "OBJ %s:%d: bad face index", path.c_str(), lineno);
                        has[k] = true;
                    }
                    if (e == std::string::npos) break;
                    p = e + 1;
                }
                if (!has[0]) return fail(NGI_ERR_PARSE, "OBJ %s:%d: face without vertex", path.c_str(), lineno);
                const long nv = (long)v.size() / 3, nt = (long)vt.size() / 2, nn = (long)vn.size() / 3;
                long a = idx[0] > 0 ? idx[0] - 1 : nv + idx[0];
                long b = has[1] ? (idx[1] > 0 ? idx[1] - 1 : nt + idx[1]) : -1;
                long c = has[2] ? (idx[2] > 0 ? idx[2] - 1 : nn + idx[2]) : -1;
                if (a < 0 || a >= nv || (has[1] && (b < 0 || b >= nt)) || (has[2] && (c < 0 || c >= nn)))
                    return fail(NGI_ERR_PARSE, "OBJ %s:%d: face index out of range", path.c_str(), lineno);
                poly.emplace_back(a, b, c);
            }
            if (poly.size() < 3) return fail(NGI_ERR_PARSE, "OBJ %s:%d: face with fewer than 3 vertices", path.c_str(), lineno);
            for (size_t k = 1; k + 1 < poly.size(); ++k) {
                corners.push_back(poly[0]);
                corners.push_back(poly[k]);
                corners.push_back(poly[k + 1]);
            }
            seen_faces = true;
        }
    }
    if (corners.empty()) return fail(NGI_ERR_INVALID_SCENE, "no mesh is found in '%s'", path.c_str());

    bool all_normals = true, any_tex = false;
    for (auto& c : corners) {
        if (std::get<2>(c) < 0) all_normals = false;
        if (std::get<1>(c) >= 0) any_tex = true;
    }
    out = ObjMesh();
    const size_t ntri = corners.size() / 3;
    if (all_normals) {
        std::map<std::tuple<long, long, long>, uint32_t> remap;
        for (auto& c : corners) {
            auto it = remap.find(c);
            uint32_t id;
            if (it == remap.end()) {
                id = (uint32_t)(out.positions.size() / 3);
                remap[c] = id;
                long a = std::get<0>(c), b = std::get<1>(c), n = std::get<2>(c);
                out.positions.insert(out.positions.end(), {v[3 * a], v[3 * a + 1], v[3 * a + 2]});
                out.normals.insert(out.normals.end(), {vn[3 * n], vn[3 * n + 1], vn[3 * n + 2]});
                if (any_tex) {
                    out.texcoords.push_back(b >= 0 ? vt[2 * b] : 0.0);
                    out.texcoords.push_back(b >= 0 ? vt[2 * b + 1] : 0.0);
                }
            } else {
                id = it->second;
            }
            out.faces.push_back(id);
        }
        return NGI_OK;
    }
    // Generated normals: flat (one vertex per corner) or smooth (area-weighted
    // face normals accumulated per position index).
    std::vector<double> acc;
    if (smooth_normals) {
        acc.assign(v.size(), 0.0);
        for (size_t t = 0; t < ntri; ++t) {
            long ia = std::get<0>(corners[3 * t]), ib = std::get<0>(corners[3 * t + 1]), ic = std::get<0>(corners[3 * t + 2]);
            d3 a(v[3 * ia], v[3 * ia + 1], v[3 * ia + 2]), b(v[3 * ib], v[3 * ib + 1], v[3 * ib + 2]), c(v[3 * ic], v[3 * ic + 1], v[3 * ic + 2]);
            d3 n = cross(b - a, c - a);
            for (long id : {ia, ib, ic}) {
                acc[3 * id] += n.x;
                acc[3 * id + 1] += n.y;
                acc[3 * id + 2] += n.z;
            }
        }
    }
    for (size_t t = 0; t < ntri; ++t) {
        long ia = std::get<0>(corners[3 * t]), ib = std::get<0>(corners[3 * t + 1]), ic = std::get<0>(corners[3 * t + 2]);
        d3 a(v[3 * ia], v[3 * ia + 1], v[3 * ia + 2]), b(v[3 * ib], v[3 * ib + 1], v[3 * ib + 2]), c(v[3 * ic], v[3 * ic + 1], v[3 * ic + 2]);
        d3 fn = cross(b - a, c - a);
        double l = length(fn);
        fn = l > 0 ? fn / l : d3(0, 0, 1);
        for (int k = 0; k < 3; ++k) {
            auto& cn = corners[3 * t + k];
            long id = std::get<0>(cn);
            out.faces.push_back((uint32_t)(out.positions.size() / 3));
            out.positions.insert(out.positions.end(), {v[3 * id], v[3 * id + 1], v[3 * id + 2]});
            d3 n = fn;
            if (std::get<2>(cn) >= 0) {
                long ni = std::get<2>(cn);
                n = d3(vn[3 * ni], vn[3 * ni + 1], vn[3 * ni + 2]);
            } else if (smooth_normals) {
                d3 s(acc[3 * id], acc[3 * id + 1], acc[3 * id + 2]);
                double sl = length(s);
                if (sl > 0) n = s / sl;
            }
            out.normals.insert(out.normals.end(), {n.x, n.y, n.z});
            if (any_tex) {
                long ti = std::get<1>(cn);
                out.texcoords.push_back(ti >= 0 ? vt[2 * ti] : 0.0);
                out.texcoords.push_back(ti >= 0 ? vt[2 * ti + 1] : 0.0);
            }
        }
    }
    return NGI_OK;
}

}  // namespace ngi

extern "C" const char* ngi_last_error(void) { return ngi::g_last_error.c_str(); }
