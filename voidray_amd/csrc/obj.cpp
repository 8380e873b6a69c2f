// Wavefront OBJ loading for Mesh::from_file (voidray_renderer/src/core/mesh.rs:46-74),
// which reads the file with obj-rs 0.7.0 as Obj<TexturedVertex, u32>.  The
// behaviour restated here is that crate's published one:
//   * statements: `v x y z [w]`, `vt u [v [w]]` (missing components 0),
//     `vn x y z`, `f` with vertex references `p`, `p/t`, `p//n` or `p/t/n`;
//     everything after `#` is a comment; other statements (o, g, s, usemtl,
//     mtllib, ...) carry nothing a TexturedVertex mesh uses and are skipped;
//   * references are 1-based; a negative one counts back from the elements
//     read so far (-1 = the last);
//   * a TexturedVertex mesh needs every face to be a triangle with position,
//     texture and normal references; anything else is an error (the crate
//     refuses untriangulated models and models without normals or uvs);
//   * vertices are the distinct (position, texture, normal) triples in order
//     of first use, and the index buffer refers to them.
// mesh.rs then keeps position, normal and (u, v) of each vertex.  The
// reference unwraps every error (a panic); here the call returns
// VR_ERR_INVALID with the file, line and reason.
#include <algorithm>
#include <cerrno>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <unordered_map>
#include <vector>

#include "internal.h"

namespace vr {
namespace {

struct ObjError {
    std::string msg;
};

struct Key {
    uint32_t p, t, n;
    bool operator==(const Key& o) const { return p == o.p && t == o.t && n == o.n; }
};
struct KeyHash {
    size_t operator()(const Key& k) const {
        uint64_t h = (uint64_t)k.p * 0x9E3779B97F4A7C15ull ^ ((uint64_t)k.t << 21) * 0xC2B2AE3D27D4EB4Full ^
                     (uint64_t)k.n * 0x165667B19E3779F9ull;
        return (size_t)(h ^ (h >> 29));
    }
};

// Rust's f32::from_str: decimal or exponent notation, optional sign, inf/nan
// words; no hexadecimal floats (strtof would take them).
float parse_f32(const char* s, size_t len, long line) {
    std::string tok(s, len);
    for (char c : tok)
        if (c == 'x' || c == 'X') throw ObjError{"line " + std::to_string(line) + ": bad number '" + tok + "'"};
    errno = 0;
    char* end = nullptr;
    float v = std::strtof(tok.c_str(), &end);
    if (end != tok.c_str() + tok.size() || tok.empty())
        throw ObjError{"line " + std::to_string(line) + ": bad number '" + tok + "'"};
    return v;  // out-of-range values round to inf or 0 as Rust's parser does
}

long parse_index(const char* s, size_t len, long line) {
    std::string tok(s, len);
    char* end = nullptr;
    errno = 0;
    long v = std::strtol(tok.c_str(), &end, 10);
    if (tok.empty() || end != tok.c_str() + tok.size() || errno)
        throw ObjError{"line " + std::to_string(line) + ": bad index '" + tok + "'"};
    return v;
}

uint32_t resolve(long idx, size_t count, long line, const char* what) {
    long long r = idx > 0 ? (long long)idx - 1 : (long long)count + idx;
    if (idx == 0 || r < 0 || (size_t)r >= count)
        throw ObjError{"line " + std::to_string(line) + ": " + what + " index " + std::to_string(idx) +
                       " out of range (" + std::to_string(count) + " defined)"};
    return (uint32_t)r;
}

struct Tokens {
    std::vector<std::pair<const char*, size_t>> t;
    void split(const char* b, const char* e) {
        t.clear();
        while (b < e) {
            while (b < e && (*b == ' ' || *b == '\t' || *b == '\r')) ++b;
            const char* s = b;
            while (b < e && !(*b == ' ' || *b == '\t' || *b == '\r')) ++b;
            if (b > s) t.emplace_back(s, (size_t)(b - s));
        }
    }
};

void load(const char* path, std::vector<float>& P, std::vector<float>& Nn, std::vector<float>& UV,
          std::vector<uint32_t>& I) {
    FILE* f = std::fopen(path, "rb");
    if (!f) throw ObjError{std::string("cannot open: ") + std::strerror(errno)};
    std::string text;
    char buf[1 << 16];
    size_t got;
    while ((got = std::fread(buf, 1, sizeof buf, f)) > 0) text.append(buf, got);
    std::fclose(f);

    std::vector<float> pos, tex, nrm;  // 3 per element
    std::unordered_map<Key, uint32_t, KeyHash> cache;
    Tokens tk;
    long line = 0;
    const char* p = text.data();
    const char* end = p + text.size();
    while (p < end) {
        const char* eol = static_cast<const char*>(std::memchr(p, '\n', (size_t)(end - p)));
        if (!eol) eol = end;
        ++line;
        const char* hash = static_cast<const char*>(std::memchr(p, '#', (size_t)(eol - p)));
        tk.split(p, hash ? hash : eol);
        p = eol + 1;
        if (tk.t.empty()) continue;
        const std::string kw(tk.t[0].first, tk.t[0].second);
        const size_t nargs = tk.t.size() - 1;
        auto num = [&](size_t k) { return parse_f32(tk.t[k].first, tk.t[k].second, line); };
        if (kw == "v") {
            if (nargs != 3 && nargs != 4) throw ObjError{"line " + std::to_string(line) + ": v needs 3 or 4 numbers"};
            for (size_t k = 1; k <= 3; ++k) pos.push_back(num(k));
            if (nargs == 4) (void)num(4);
        } else if (kw == "vt") {
            if (nargs < 1 || nargs > 3) throw ObjError{"line " + std::to_string(line) + ": vt needs 1 to 3 numbers"};
            for (size_t k = 1; k <= 3; ++k) tex.push_back(k <= nargs ? num(k) : 0.0f);
        } else if (kw == "vn") {
            if (nargs != 3) throw ObjError{"line " + std::to_string(line) + ": vn needs 3 numbers"};
            for (size_t k = 1; k <= 3; ++k) nrm.push_back(num(k));
        } else if (kw == "f") {
            if (nargs < 3) throw ObjError{"line " + std::to_string(line) + ": a face needs at least 3 vertices"};
            if (nargs != 3)
                throw ObjError{"line " + std::to_string(line) +
                               ": model is not triangulated (a TexturedVertex mesh takes triangles only)"};
            for (size_t k = 1; k <= 3; ++k) {
                const char* s = tk.t[k].first;
                const size_t n = tk.t[k].second;
                const char* s1 = static_cast<const char*>(std::memchr(s, '/', n));
                const char* s2 = s1 ? static_cast<const char*>(std::memchr(s1 + 1, '/', (size_t)(s + n - s1 - 1)))
                                    : nullptr;
                if (!s1 || !s2 || s2 == s1 + 1)
                    throw ObjError{"line " + std::to_string(line) +
                                   ": a TexturedVertex mesh needs position/texture/normal references"};
                const long pi = parse_index(s, (size_t)(s1 - s), line);
                const long ti = parse_index(s1 + 1, (size_t)(s2 - s1 - 1), line);
                const long ni = parse_index(s2 + 1, (size_t)(s + n - s2 - 1), line);
                const Key key{resolve(pi, pos.size() / 3, line, "position"),
                              resolve(ti, tex.size() / 3, line, "texture"),
                              resolve(ni, nrm.size() / 3, line, "normal")};
                auto it = cache.find(key);
                uint32_t v;
                if (it == cache.end()) {
                    v = (uint32_t)(P.size() / 3);
                    cache.emplace(key, v);
                    for (int c = 0; c < 3; ++c) P.push_back(pos[3 * key.p + c]);
                    for (int c = 0; c < 3; ++c) Nn.push_back(nrm[3 * key.n + c]);
                    UV.push_back(tex[3 * key.t + 0]);
                    UV.push_back(tex[3 * key.t + 1]);
                } else {
                    v = it->second;
                }
                I.push_back(v);
            }
        }
        // any other statement: nothing a TexturedVertex mesh keeps
    }
}

template <class T> T* copy_out(const std::vector<T>& v) {
    T* p = static_cast<T*>(std::malloc(std::max<size_t>(1, v.size()) * sizeof(T)));
    if (!p) throw std::bad_alloc();
    if (!v.empty()) std::memcpy(p, v.data(), v.size() * sizeof(T));
    return p;
}

}  // namespace
}  // namespace vr

using namespace vr;

extern "C" {

int vr_mesh_load_obj(const char* path, vr_mesh* out) {
    if (!path || !out) return fail(VR_ERR_INVALID, "vr_mesh_load_obj: null argument");
    std::memset(out, 0, sizeof *out);
    try {
        std::vector<float> P, Nn, UV;
        std::vector<uint32_t> I;
        load(path, P, Nn, UV, I);
        if (P.size() / 3 > 0xFFFFFFFFull || I.size() / 3 > 0xFFFFFFFFull)
            return fail(VR_ERR_INVALID, std::string(path) + ": too many vertices or faces");
        float* pp = copy_out(P);
        float* pn = copy_out(Nn);
        float* pu = copy_out(UV);
        uint32_t* pi = copy_out(I);
        out->num_vertices = (uint32_t)(P.size() / 3);
        out->num_triangles = (uint32_t)(I.size() / 3);
        out->positions = pp;
        out->normals = pn;
        out->uvs = pu;
        out->indices = pi;
        return 0;
    } catch (const ObjError& e) {
        return fail(VR_ERR_INVALID, std::string(path) + ": " + e.msg);
    } catch (const std::bad_alloc&) {
        return fail(VR_ERR_OOM, "host allocation failed");
    } catch (...) {
        return fail(VR_ERR_INVALID, std::string(path) + ": unknown error");
    }
}

int vr_mesh_release(vr_mesh* m) {
    if (!m) return 0;
    std::free(const_cast<float*>(m->positions));
    std::free(const_cast<float*>(m->normals));
    std::free(const_cast<float*>(m->uvs));
    std::free(const_cast<uint32_t*>(m->indices));
    std::memset(m, 0, sizeof *m);
    return 0;
}

}  // extern "C"
