// PLY reader for the mesh builder: the subset of PLY the reference accepts
// (src/third-party/plyfile.c with PlyLoader's property mapping,
// src/PlyLoader.cpp:81-224):
//  - ".ply" is appended when the path lacks it (plyfile.c:798-850);
//  - ascii and binary_little_endian; old type names only (char short int
//    uchar ushort uint float double, plyfile.c:41-44);
//  - vertex properties x y z nx ny nz red green blue are read as float,
//    absent normals = 0, absent colours = 255; other properties/elements skipped;
//  - face property vertex_indices (a list) is read as int.
// Big-endian files are rejected: the reference converts before swapping bytes
// (PlyLoader.cpp:104-120,152), so it has no meaningful output for them.
#include <cerrno>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "svo_internal.hpp"

namespace svo {
namespace {

PType parse_type(const std::string &s) {
    static const char *names[] = {"invalid", "char", "short", "int", "uchar", "ushort", "uint", "float", "double"};
    for (int i = 1; i <= 8; ++i)
        if (s == names[i]) return PType(i);
    return P_INVALID;
}

int type_size(PType t) { return ply_type_size(t); }

// Reads one binary scalar as double (the value plyfile.c's get_binary_item
// yields before store_item converts it to the internal type).
bool read_bin(const uint8_t *&p, const uint8_t *end, PType t, double &v) {
    int n = type_size(t);
    if (end - p < n) return false;
    switch (t) {
        case P_CHAR: { int8_t x; std::memcpy(&x, p, 1); v = x; break; }
        case P_UCHAR: { uint8_t x; std::memcpy(&x, p, 1); v = x; break; }
        case P_SHORT: { int16_t x; std::memcpy(&x, p, 2); v = x; break; }
        case P_USHORT: { uint16_t x; std::memcpy(&x, p, 2); v = x; break; }
        case P_INT: { int32_t x; std::memcpy(&x, p, 4); v = x; break; }
        case P_UINT: { uint32_t x; std::memcpy(&x, p, 4); v = x; break; }
        case P_FLOAT: { float x; std::memcpy(&x, p, 4); v = x; break; }
        case P_DOUBLE: { double x; std::memcpy(&x, p, 8); v = x; break; }
        default: return false;
    }
    p += n;
    return true;
}

bool is_int_type(PType t) { return t != P_FLOAT && t != P_DOUBLE; }

}  // namespace

int ply_open(const char *path_in, PlyLayout &L) {
    std::string path(path_in);
    if (path.size() < 4 || path.compare(path.size() - 4, 4, ".ply") != 0) path += ".ply";
    if (!L.buf.open(path.c_str())) return fail(SVO_ERR_IO, "cannot open PLY " + path);
    const FileView &buf = L.buf;
    // ---- header ----
    size_t pos = 0;
    auto next_line = [&](std::string &line) -> bool {
        if (pos >= buf.size()) return false;
        size_t e = pos;
        while (e < buf.size() && buf[e] != '\n') e++;
        line.assign(reinterpret_cast<const char *>(buf.data()) + pos, e - pos);
        if (!line.empty() && line.back() == '\r') line.pop_back();
        pos = e < buf.size() ? e + 1 : e;
        return true;
    };
    auto words = [](const std::string &s) {
        std::vector<std::string> w;
        size_t i = 0;
        while (i < s.size()) {
            while (i < s.size() && (s[i] == ' ' || s[i] == '\t')) i++;
            size_t j = i;
            while (j < s.size() && s[j] != ' ' && s[j] != '\t') j++;
            if (j > i) w.push_back(s.substr(i, j - i));
            i = j;
        }
        return w;
    };
    std::string line;
    if (!next_line(line) || words(line) != std::vector<std::string>{"ply"}) return fail(SVO_ERR_FORMAT, "not a PLY file");
    int format = -1;  // 0 ascii, 1 binary LE
    std::vector<PlyElem> &elems = L.elems;
    elems.clear();
    bool ended = false;
    while (next_line(line)) {
        std::vector<std::string> w = words(line);
        if (w.empty()) continue;
        if (w[0] == "format" && w.size() >= 2) {
            if (w[1] == "ascii") format = 0;
            else if (w[1] == "binary_little_endian") format = 1;
            else if (w[1] == "binary_big_endian") return fail(SVO_ERR_UNSUPPORTED, "big-endian PLY is not supported");
            else return fail(SVO_ERR_FORMAT, "unknown PLY format " + w[1]);
        } else if (w[0] == "element" && w.size() >= 3) {
            PlyElem e;
            e.name = w[1];
            e.count = std::strtoll(w[2].c_str(), nullptr, 10);
            if (e.count < 0) return fail(SVO_ERR_FORMAT, "negative element count");
            elems.push_back(e);
        } else if (w[0] == "property") {
            if (elems.empty()) return fail(SVO_ERR_FORMAT, "property before element");
            PlyProp p;
            if (w.size() >= 5 && w[1] == "list") {
                p.list = true;
                p.count_type = parse_type(w[2]);
                p.type = parse_type(w[3]);
                p.name = w[4];
                if (!p.count_type || !p.type) return fail(SVO_ERR_FORMAT, "unknown PLY list type in: " + line);
            } else if (w.size() >= 3) {
                p.type = parse_type(w[1]);
                p.name = w[2];
                if (!p.type) return fail(SVO_ERR_FORMAT, "unknown PLY property type in: " + line);
            } else {
                return fail(SVO_ERR_FORMAT, "bad property line: " + line);
            }
            elems.back().props.push_back(p);
        } else if (w[0] == "end_header") {
            ended = true;
            break;
        }
    }
    if (!ended || format < 0) return fail(SVO_ERR_FORMAT, "PLY header incomplete");
    // Each record takes at least: the sizes of its scalars and list counts
    // (binary), or one character and one separator per property (ascii; the
    // file's last token may end without a separator). Counts the body cannot
    // hold are refused here, before anything is sized or looped by them;
    // elements without properties are capped at INT32_MAX records.
    const uint64_t bodyBytes = buf.size() - pos;
    for (const PlyElem &e : elems) {
        uint64_t rmin = 0;
        for (const PlyProp &p : e.props) rmin += format == 1 ? uint64_t(type_size(p.list ? p.count_type : p.type)) : 2;
        const uint64_t cap = rmin ? (bodyBytes + 1) / rmin : uint64_t(INT32_MAX);
        if (uint64_t(e.count) > cap) return fail(SVO_ERR_FORMAT, "PLY element " + e.name + ": count exceeds the data");
    }
    L.binary = format == 1;
    L.body = pos;
    L.vertex = L.face = -1;
    for (size_t i = 0; i < elems.size(); ++i) {
        if (elems[i].name == "vertex") L.vertex = int(i);
        else if (elems[i].name == "face") L.face = int(i);
    }
    if (L.vertex < 0 || L.face < 0) return fail(SVO_ERR_FORMAT, "PLY file has to have triangles and vertices");
    const PlyElem *ve = &elems[size_t(L.vertex)], *fe = &elems[size_t(L.face)];
    bool avail[9] = {false};
    for (const PlyProp &p : ve->props)
        for (int t = 0; t < 9; ++t)
            if (!p.list && p.name == kPlyVertexNames[t]) avail[t] = true;
    L.has_normals = avail[3] && avail[4] && avail[5];
    bool faceIdx = false;
    for (const PlyProp &p : fe->props)
        if (p.list && p.name == "vertex_indices") faceIdx = true;
    if (!faceIdx) return fail(SVO_ERR_FORMAT, "No face information found");
    return SVO_OK;
}

int read_ply(const char *path_in, std::vector<float> &verts, int &has_normals, std::vector<int32_t> &face_counts,
             std::vector<int32_t> &face_idx) {
    PlyLayout L;
    int r = ply_open(path_in, L);
    if (r) return r;
    return ply_parse_host(L, verts, has_normals, face_counts, face_idx);
}

int ply_parse_host(const PlyLayout &L, std::vector<float> &verts, int &has_normals, std::vector<int32_t> &face_counts,
                   std::vector<int32_t> &face_idx) {
    const FileView &buf = L.buf;
    const std::vector<PlyElem> &elems = L.elems;
    const PlyElem *ve = &elems[size_t(L.vertex)], *fe = &elems[size_t(L.face)];
    const int format = L.binary ? 1 : 0;
    const size_t pos = L.body;
    const char *const *vnames = kPlyVertexNames;
    has_normals = L.has_normals;
    const float defaults[9] = {0, 0, 0, 0, 0, 0, 255.0f, 255.0f, 255.0f};
    verts.assign(size_t(ve->count) * 9, 0.0f);
    face_counts.clear();
    face_idx.clear();

    // ---- body ----
    const uint8_t *p = buf.data() + pos, *end = buf.data() + buf.size();
    const char *tp = reinterpret_cast<const char *>(p);
    const char *tend = reinterpret_cast<const char *>(end);
    auto next_token = [&](std::string &tok) -> bool {
        while (tp < tend && (*tp == ' ' || *tp == '\t' || *tp == '\n' || *tp == '\r')) tp++;
        const char *s = tp;
        while (tp < tend && !(*tp == ' ' || *tp == '\t' || *tp == '\n' || *tp == '\r')) tp++;
        if (tp == s) return false;
        tok.assign(s, size_t(tp - s));
        return true;
    };
    // value of one item as plyfile.c sees it (ascii: atoi for integer types, atof otherwise)
    auto read_item = [&](PType t, double &v) -> bool {
        if (format == 1) return read_bin(p, end, t, v);
        std::string tok;
        if (!next_token(tok)) return false;
        v = is_int_type(t) ? double(std::atoi(tok.c_str())) : std::atof(tok.c_str());
        return true;
    };
    for (const PlyElem &e : elems) {
        const bool isV = &e == ve, isF = &e == fe;
        for (int64_t i = 0; i < e.count; ++i) {
            float row[9];
            if (isV) std::memcpy(row, defaults, sizeof(row));
            for (const PlyProp &pr : e.props) {
                if (pr.list) {
                    double cv;
                    if (!read_item(pr.count_type, cv)) return fail(SVO_ERR_FORMAT, "PLY data truncated");
                    if (!(cv >= 0.0)) return fail(SVO_ERR_FORMAT, "negative list length");
                    if (cv > 4294967295.0) return fail(SVO_ERR_FORMAT, "PLY data truncated");
                    const int64_t cnt = int64_t(cv);
                    const bool take = isF && pr.name == "vertex_indices";
                    if (take) face_counts.push_back(int32_t(cnt));
                    for (int64_t k = 0; k < cnt; ++k) {
                        double v;
                        if (!read_item(pr.type, v)) return fail(SVO_ERR_FORMAT, "PLY data truncated");
                        if (take) {
                            // int conversion as plyfile.c's store_item; out of
                            // int range (or NaN) the index is missing anyway
                            if (!(v > -2147483649.0 && v < 2147483648.0))
                                return fail(SVO_ERR_FORMAT, "face references a missing vertex");
                            face_idx.push_back(int32_t(v));
                        }
                    }
                } else {
                    double v;
                    if (!read_item(pr.type, v)) return fail(SVO_ERR_FORMAT, "PLY data truncated");
                    if (isV)
                        for (int t = 0; t < 9; ++t)
                            if (pr.name == vnames[t]) row[t] = float(v);
                }
            }
            if (isV) std::memcpy(&verts[size_t(i) * 9], row, sizeof(row));
        }
    }
    for (int32_t v : face_idx)
        if (v < 0 || int64_t(v) >= ve->count) return fail(SVO_ERR_FORMAT, "face references a missing vertex");
    return SVO_OK;
}

int ply_type_size(PType t) {
    static const int sz[] = {0, 1, 2, 4, 1, 2, 4, 4, 8};
    return sz[t];
}

const char *const kPlyVertexNames[9] = {"x", "y", "z", "nx", "ny", "nz", "red", "green", "blue"};

}  // namespace svo
