// ingest.cpp -- scene ingestion: Wavefront OBJ files and Beamcast GX triangle page stores into
// the host scene mirror, ready for the device build (bc_build_bvh) and upload.
//
// OBJ: restates LoadOBJ (renderer.go:2996-3069) with the same results -- the same triangles in
// the same order (fan triangulation of each face, degenerate triangles dropped by AddTriangle),
// the same error for the first offending line, the same camera -- but parses the file in
// parallel: the file is cut into chunks at line starts; each chunk is tokenised by one thread
// (vertices, faces with their raw indices, the chunk's first syntax error); a prefix over the
// chunks' vertex counts resolves relative (negative) indices and range checks; the primitives of
// each chunk are made in parallel and concatenated in file order.  The error reported is the one
// of the lowest line number, which is the one LoadOBJ stops at.
//
// Page store: the BGXPAGE1 format of PagedTriangleStore (gx_stream.cpp:9-55): a 16-byte header
// (magic, version 1, page count), a page table of (file offset, triangle count, reserved) and
// 80-byte PackedTriangle records (v0, e1, e2, normal as 16-byte float vectors, material).  Pages
// are appended as append_page_to_scene does (triangle v0, v0 + e1, v0 + e2 with float sums),
// through the Go AddTriangle.
#include <algorithm>
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "../../include/beamcast_host.h"
#include "host_scene.hpp"

namespace {

using bcx::V;
using bcx::Vec3;

// unicode.IsSpace over UTF-8 input: ASCII \t \n \v \f \r and space, and the multi-byte spaces
// U+0085, U+00A0, U+1680, U+2000-U+200A, U+2028, U+2029, U+202F, U+205F, U+3000.  Returns the
// length of the space at p (0 if none).
size_t space_len(const char* p, const char* end) {
    const unsigned char c = (unsigned char)*p;
    if (c == ' ' || (c >= '\t' && c <= '\r')) return 1;
    if (c == 0xC2 && p + 1 < end) {
        const unsigned char d = (unsigned char)p[1];
        return d == 0x85 || d == 0xA0 ? 2 : 0;
    }
    if (c == 0xE1 && p + 2 < end) return (unsigned char)p[1] == 0x9A && (unsigned char)p[2] == 0x80 ? 3 : 0;
    if (c == 0xE2 && p + 2 < end) {
        const unsigned char d = (unsigned char)p[1], e = (unsigned char)p[2];
        if (d == 0x80 && ((e >= 0x80 && e <= 0x8A) || e == 0xA8 || e == 0xA9 || e == 0xAF)) return 3;
        if (d == 0x81 && e == 0x9F) return 3;
        return 0;
    }
    if (c == 0xE3 && p + 2 < end) return (unsigned char)p[1] == 0x80 && (unsigned char)p[2] == 0x80 ? 3 : 0;
    return 0;
}

// strings.Fields
void fields(const char* p, const char* end, std::vector<std::pair<const char*, const char*>>& out) {
    out.clear();
    while (p < end) {
        size_t s;
        while (p < end && (s = space_len(p, end)) > 0) p += s;
        if (p >= end) break;
        const char* b = p;
        while (p < end && space_len(p, end) == 0) p++;
        out.emplace_back(b, p);
    }
}

// strconv.ParseFloat(tok, 64): decimal and hexadecimal (with a p exponent) forms, Inf / Infinity
// / NaN in any case with an optional sign; overflow to +-Inf is an error, underflow is not.
bool parse_float(const char* b, const char* e, double& out) {
    const size_t n = (size_t)(e - b);
    if (n == 0) return false;
    char small[128];
    std::string big;
    char* buf = small;
    if (n >= sizeof small) {
        big.assign(b, n);
        buf = &big[0];
    } else {
        std::memcpy(buf, b, n);
        buf[n] = 0;
    }
    for (size_t i = 0; i < n; i++)
        if (buf[i] == '(' || buf[i] == '_') return false;  // nan(...) and digit separators are not Go syntax here
    const char* q = buf + ((buf[0] == '+' || buf[0] == '-') ? 1 : 0);
    const bool hex = q[0] == '0' && (q[1] == 'x' || q[1] == 'X');
    if (hex && !std::strpbrk(q, "pP")) return false;  // Go requires the p exponent of a hex mantissa
    errno = 0;
    char* stop = nullptr;
    const double v = std::strtod(buf, &stop);
    if (stop != buf + n) return false;
    if (errno == ERANGE && std::isinf(v)) return false;
    const bool special = std::isinf(v) || std::isnan(v);
    if (special) {  // only the spellings Inf, Infinity, NaN (any case, signed)
        std::string w(q);
        for (auto& ch : w) ch = (char)std::tolower((unsigned char)ch);
        if (w != "inf" && w != "infinity" && w != "nan") return false;
        if (w == "nan" && q != buf) return false;  // Go rejects a signed NaN
    }
    out = v;
    return true;
}

// strconv.Atoi: optional sign, decimal digits, fits int64
bool parse_int(const char* b, const char* e, long long& out) {
    if (b >= e) return false;
    bool neg = false;
    if (*b == '+' || *b == '-') { neg = *b == '-'; b++; }
    if (b >= e) return false;
    unsigned long long v = 0;
    for (const char* p = b; p < e; p++) {
        if (*p < '0' || *p > '9') return false;
        const unsigned d = (unsigned)(*p - '0');
        if (v > (9223372036854775808ull - d) / 10) return false;
        v = v * 10 + d;
    }
    if (!neg && v > 9223372036854775807ull) return false;
    out = neg ? (long long)(0 - v) : (long long)v;
    return true;
}

constexpr size_t kMaxLine = 65536;  // bufio.MaxScanTokenSize: a longer line ends the scan

struct Face {
    long long line;       // chunk-local line number (0-based)
    long long vbefore;    // chunk-local vertices read before this line
    size_t first, count;  // raw indices in Chunk::idx
};

struct Chunk {
    const char* b = nullptr;
    const char* e = nullptr;
    long long lines = 0;            // lines in the chunk
    std::vector<Vec3> verts;
    std::vector<Face> faces;
    std::vector<long long> idx;     // raw face indices
    long long err_line = -1;        // first error (chunk-local line), -1 if none
    std::string err;
    bool err_plain = false;         // error text without the "OBJ line N: " prefix
    double mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
    bcx::Scene prims;               // this chunk's triangles
};

void set_err(Chunk& c, long long line, const char* msg, bool plain = false) {
    if (c.err_line < 0 || line < c.err_line) {
        c.err_line = line;
        c.err = msg;
        c.err_plain = plain;
    }
}

void tokenise(Chunk& c) {
    std::vector<std::pair<const char*, const char*>> parts;
    const char* p = c.b;
    long long line = 0;
    while (p < c.e) {
        const char* nl = (const char*)std::memchr(p, '\n', (size_t)(c.e - p));
        const char* le = nl ? nl : c.e;
        const long long ln = line++;
        const char* next = nl ? nl + 1 : c.e;
        if ((size_t)(le - p) + 1 > kMaxLine) { set_err(c, ln, "bufio.Scanner: token too long", true); break; }
        fields(p, le, parts);
        p = next;
        if (parts.empty() || *parts[0].first == '#') continue;  // blank or comment (after TrimSpace)
        const size_t n0 = (size_t)(parts[0].second - parts[0].first);
        if (n0 == 1 && *parts[0].first == 'v') {
            if (parts.size() < 4) { set_err(c, ln, "malformed vertex"); break; }
            double x, y, z;
            const bool ok1 = parse_float(parts[1].first, parts[1].second, x);
            const bool ok2 = parse_float(parts[2].first, parts[2].second, y);
            const bool ok3 = parse_float(parts[3].first, parts[3].second, z);
            if (!ok1 || !ok2 || !ok3) { set_err(c, ln, "invalid vertex"); break; }
            c.verts.push_back(V(x, y, z));
            c.mn[0] = bcx::GoMin(c.mn[0], x); c.mn[1] = bcx::GoMin(c.mn[1], y); c.mn[2] = bcx::GoMin(c.mn[2], z);
            c.mx[0] = bcx::GoMax(c.mx[0], x); c.mx[1] = bcx::GoMax(c.mx[1], y); c.mx[2] = bcx::GoMax(c.mx[2], z);
        } else if (n0 == 1 && *parts[0].first == 'f') {
            if (parts.size() < 4) { set_err(c, ln, "face needs 3 vertices"); break; }
            Face f{ln, (long long)c.verts.size(), c.idx.size(), parts.size() - 1};
            bool bad = false;
            for (size_t k = 1; k < parts.size(); k++) {
                const char* slash = (const char*)std::memchr(parts[k].first, '/', (size_t)(parts[k].second - parts[k].first));
                long long v = 0;
                if (!parse_int(parts[k].first, slash ? slash : parts[k].second, v) || v == 0) { bad = true; break; }
                c.idx.push_back(v);
            }
            if (bad) { set_err(c, ln, "invalid index"); break; }
            c.faces.push_back(f);
        }
    }
    c.lines = line;
}

// Resolves a chunk's faces against the global vertex array and makes its triangles.
void make_triangles(Chunk& c, const std::vector<Vec3>& verts, long long vbase, int mat, long long stop_line) {
    std::vector<long long> idx;
    for (const Face& f : c.faces) {
        if (stop_line >= 0 && f.line >= stop_line) break;
        const long long nv = vbase + f.vbefore;  // len(verts) when LoadOBJ reads this line
        idx.clear();
        bool bad = false;
        for (size_t k = 0; k < f.count; k++) {
            long long n = c.idx[f.first + k];
            if (n < 0) n = nv + n + 1;
            n--;
            if (n < 0 || n >= nv) { bad = true; break; }
            idx.push_back(n);
        }
        if (bad) { set_err(c, f.line, "index out of range"); return; }
        for (size_t i = 1; i + 1 < idx.size(); i++)
            c.prims.AddTriangle(verts[(size_t)idx[0]], verts[(size_t)idx[i]], verts[(size_t)idx[i + 1]], mat);
    }
}

// LoadOBJ (renderer.go:2996-3069) into s (primitives and one material; not built).  cam10:
// position, target, up, fov.  Returns false with the LoadOBJ error text.
bool load_obj(const char* path, bcx::Scene& s, double* cam10, std::string& err, int threads) {
    FILE* f = std::fopen(path, "rb");
    if (!f) { err = std::string("open ") + path + ": " + std::strerror(errno); return false; }
    std::string data;
    {
        std::fseek(f, 0, SEEK_END);
        const long sz = std::ftell(f);
        std::fseek(f, 0, SEEK_SET);
        if (sz < 0) { std::fclose(f); err = "cannot size " + std::string(path); return false; }
        data.resize((size_t)sz);
        if (sz > 0 && std::fread(&data[0], 1, (size_t)sz, f) != (size_t)sz) { std::fclose(f); err = "read error"; return false; }
        std::fclose(f);
    }
    s = bcx::Scene();
    s.Name = path;
    bcx::Material m;
    m.Kind = bcx::MatDiffuse;
    m.Color = V(.72, .76, .82);
    const int mat = s.AddMaterial(m);
    // chunks cut at line starts
    const size_t n = data.size();
    const int T = std::max(1, std::min(threads, (int)std::max<size_t>(1, n / (1 << 20))));
    std::vector<Chunk> ch((size_t)T);
    const char* base = data.data();
    size_t pos = 0;
    for (int t = 0; t < T; t++) {
        size_t end = t + 1 == T ? n : std::max(pos, n * (size_t)(t + 1) / (size_t)T);
        while (end < n && end > 0 && base[end - 1] != '\n') end++;
        ch[(size_t)t].b = base + pos;
        ch[(size_t)t].e = base + end;
        pos = end;
    }
    // chunk t on thread t; a chunk whose thread cannot start runs on the calling thread, so no
    // exception leaves the C entry points
    auto run = [&](auto fn) {
        std::vector<std::thread> th;
        int t = 1;
        try {
            th.reserve((size_t)T);
            for (; t < T; t++) th.emplace_back(fn, t);
        } catch (...) {
        }
        for (int u = t; u < T; u++) fn(u);
        fn(0);
        for (auto& x : th) x.join();
    };
    run([&](int t) { tokenise(ch[(size_t)t]); });
    // global vertex array, line offsets
    std::vector<long long> vbase((size_t)T + 1, 0), lbase((size_t)T + 1, 0);
    for (int t = 0; t < T; t++) {
        vbase[(size_t)t + 1] = vbase[(size_t)t] + (long long)ch[(size_t)t].verts.size();
        lbase[(size_t)t + 1] = lbase[(size_t)t] + ch[(size_t)t].lines;
    }
    // the first syntax error bounds what LoadOBJ reads
    long long stop = -1;
    for (int t = 0; t < T && stop < 0; t++)
        if (ch[(size_t)t].err_line >= 0) stop = lbase[(size_t)t] + ch[(size_t)t].err_line;
    std::vector<Vec3> verts((size_t)vbase[(size_t)T]);
    run([&](int t) {
        std::copy(ch[(size_t)t].verts.begin(), ch[(size_t)t].verts.end(), verts.begin() + vbase[(size_t)t]);
    });
    // faces before the first syntax error (all of them without one)
    run([&](int t) {
        make_triangles(ch[(size_t)t], verts, vbase[(size_t)t], mat, stop < 0 ? -1 : std::max(0LL, stop - lbase[(size_t)t]));
    });
    long long eline = -1;
    std::string emsg;
    bool eplain = false;
    for (int t = 0; t < T; t++)
        if (ch[(size_t)t].err_line >= 0) {
            const long long g = lbase[(size_t)t] + ch[(size_t)t].err_line;
            if (eline < 0 || g < eline) { eline = g; emsg = ch[(size_t)t].err; eplain = ch[(size_t)t].err_plain; }
        }
    if (eline >= 0) { err = eplain ? emsg : "OBJ line " + std::to_string(eline + 1) + ": " + emsg; return false; }
    size_t total = 0;
    for (auto& c : ch) total += c.prims.Primitives.size();
    s.Primitives.reserve(total);
    for (auto& c : ch) s.Primitives.insert(s.Primitives.end(), c.prims.Primitives.begin(), c.prims.Primitives.end());
    if (s.Primitives.empty()) { err = "OBJ contains no renderable faces"; return false; }
    double mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (auto& c : ch)
        for (int a = 0; a < 3; a++) { mn[a] = bcx::GoMin(mn[a], c.mn[a]); mx[a] = bcx::GoMax(mx[a], c.mx[a]); }
    const Vec3 lo = V(mn[0], mn[1], mn[2]), hi = V(mx[0], mx[1], mx[2]);
    const Vec3 center = (lo + hi) * .5;
    const Vec3 d = hi - lo;
    const double rad = std::max(.5, std::sqrt(bcx::Dot(d, d)) * .5);
    const Vec3 pos3 = center + V(rad * 1.4, rad * .9, rad * 2.6);
    if (cam10) {
        const double cv[10] = {pos3.x, pos3.y, pos3.z, center.x, center.y, center.z, 0, 1, 0, 42};
        std::memcpy(cam10, cv, sizeof cv);
    }
    return true;
}

// ---------------------------------------------------------------------------------------------
// BGXPAGE1 page store

constexpr uint64_t kPageMagic = 0x3145474150584742ull;  // "BGXPAGE1"
struct PageHeader {
    uint64_t magic;
    uint32_t version, page_count;
};
struct PageEntry {
    uint64_t offset;
    uint32_t count, reserved;
};
static_assert(sizeof(PageHeader) == 16 && sizeof(PageEntry) == 16, "page store layout");
static_assert(sizeof(bch_packed_triangle) == 80, "PackedTriangle is 80 bytes");

bool read_table(FILE* f, std::vector<PageEntry>& pages, std::string& err) {
    PageHeader h{};
    if (std::fread(&h, sizeof h, 1, f) != 1 || h.magic != kPageMagic || h.version != 1) {
        err = "invalid Beamcast GX page store";
        return false;
    }
    pages.resize(h.page_count);
    if (h.page_count && std::fread(pages.data(), sizeof(PageEntry), h.page_count, f) != h.page_count) {
        err = "truncated Beamcast GX page table";
        return false;
    }
    return true;
}

}  // namespace

struct bch_pages {
    std::string path;
    std::vector<PageEntry> pages;
};

extern "C" {

bch_scene* bch_load_obj(const char* path, double* camera_out, char* err, int err_cap, int threads) {
    if (!path) return nullptr;
    bch_scene* s = new bch_scene;
    std::string e;
    if (!load_obj(path, s->s, camera_out, e, threads > 0 ? threads : (int)std::max(1u, std::thread::hardware_concurrency()))) {
        if (err && err_cap > 0) std::snprintf(err, (size_t)err_cap, "%s", e.c_str());
        delete s;
        return nullptr;
    }
    return s;
}

int bch_pages_write(const char* path, const bch_packed_triangle* tris, int64_t n, uint32_t per_page) {
    if (!path || per_page == 0 || n < 0 || (n > 0 && !tris)) return -1;
    const uint64_t pages = ((uint64_t)n + per_page - 1) / per_page;
    if (pages > 0xFFFFFFFFull) return -1;
    FILE* f = std::fopen(path, "wb");
    if (!f) return -2;
    PageHeader h{kPageMagic, 1, (uint32_t)pages};
    bool ok = std::fwrite(&h, sizeof h, 1, f) == 1;
    uint64_t off = sizeof(PageHeader) + pages * sizeof(PageEntry);
    for (uint64_t p = 0; ok && p < pages; p++) {
        const uint32_t cnt = (uint32_t)std::min<uint64_t>(per_page, (uint64_t)n - p * per_page);
        PageEntry d{off, cnt, 0};
        ok = std::fwrite(&d, sizeof d, 1, f) == 1;
        off += (uint64_t)cnt * sizeof(bch_packed_triangle);
    }
    if (ok && n > 0) ok = std::fwrite(tris, sizeof(bch_packed_triangle), (size_t)n, f) == (size_t)n;
    ok = (std::fclose(f) == 0) && ok;
    return ok ? 0 : -2;
}

bch_pages* bch_pages_open(const char* path, char* err, int err_cap) {
    std::string e = "cannot open page store";
    FILE* f = path ? std::fopen(path, "rb") : nullptr;
    bch_pages* p = nullptr;
    if (f) {
        p = new bch_pages;
        p->path = path;
        if (!read_table(f, p->pages, e)) { delete p; p = nullptr; }
        std::fclose(f);
    }
    if (!p && err && err_cap > 0) std::snprintf(err, (size_t)err_cap, "%s", e.c_str());
    return p;
}

void bch_pages_close(bch_pages* p) { delete p; }

int64_t bch_pages_count(const bch_pages* p) { return p ? (int64_t)p->pages.size() : 0; }

int bch_pages_info(const bch_pages* p, int64_t page, uint64_t* offset, uint32_t* count) {
    if (!p || page < 0 || page >= (int64_t)p->pages.size()) return -1;
    if (offset) *offset = p->pages[(size_t)page].offset;
    if (count) *count = p->pages[(size_t)page].count;
    return 0;
}

int bch_pages_load(const bch_pages* p, int64_t page, bch_packed_triangle* out) {
    if (!p || page < 0 || page >= (int64_t)p->pages.size() || !out) return -1;
    const PageEntry& d = p->pages[(size_t)page];
    FILE* f = std::fopen(p->path.c_str(), "rb");
    if (!f) return -2;
    bool ok = std::fseek(f, (long)d.offset, SEEK_SET) == 0;
    if (ok && d.count) ok = std::fread(out, sizeof(bch_packed_triangle), d.count, f) == d.count;
    std::fclose(f);
    return ok ? 0 : -3;
}

int bch_pages_append(bch_scene* s, const bch_pages* p, int64_t first, int64_t npages) {
    if (!s || !p || first < 0 || npages < 0 || first + npages > (int64_t)p->pages.size()) return -1;
    std::vector<bch_packed_triangle> buf;
    for (int64_t k = first; k < first + npages; k++) {
        buf.resize(p->pages[(size_t)k].count);
        const int rc = bch_pages_load(p, k, buf.data());
        if (rc != 0) return rc;
        for (const bch_packed_triangle& t : buf) {  // append_page_to_scene: v0, v0 + e1, v0 + e2 in float
            const float b[3] = {t.v0[0] + t.e1[0], t.v0[1] + t.e1[1], t.v0[2] + t.e1[2]};
            const float c[3] = {t.v0[0] + t.e2[0], t.v0[1] + t.e2[1], t.v0[2] + t.e2[2]};
            s->s.AddTriangle(V(t.v0[0], t.v0[1], t.v0[2]), V(b[0], b[1], b[2]), V(c[0], c[1], c[2]), (int)t.material);
        }
    }
    return 0;
}

}  // extern "C"

extern "C" int bch_write_obj(const bch_scene* s, const char* path) {
    if (!s || !path) return -1;
    FILE* f = std::fopen(path, "wb");
    if (!f) return -2;
    std::fprintf(f, "# %zu primitives\n", s->s.Primitives.size());
    long long nv = 0, k = 0;
    for (const bcx::Primitive& p : s->s.Primitives) {
        if (p.Kind != bcx::PrimTriangle) continue;  // OBJ holds no spheres
        for (const Vec3& q : {p.A, p.B, p.C}) std::fprintf(f, "v %.17g %.17g %.17g\n", q.x, q.y, q.z);
        nv += 3;
        if (k++ & 1) std::fprintf(f, "f -3 -2 -1\n");  // relative indices on every other face
        else std::fprintf(f, "f %lld %lld %lld\n", nv - 2, nv - 1, nv);
    }
    return std::fclose(f) == 0 ? 0 : -2;
}

extern "C" bch_scene* bch_ingest_obj(bc_ctx* ctx, const char* path, int bins, int leaf, int threads, double* camera_out,
                                     char* err, int err_cap) {
    bch_scene* s = bch_load_obj(path, camera_out, err, err_cap, threads);
    if (!s) return nullptr;
    s->s.BVHBins = bins;
    s->s.BVHLeafSize = leaf;
    std::string e;
    int rc = ctx ? s->s.BuildDevice(ctx, e) : BC_ERR_ARG;
    if (rc == BC_OK) {
        rc = bc_scene_upload(ctx, &s->s.View(false), s->s.GPUTopologyRevision, s->s.GPUGeometryRevision);
        if (rc != BC_OK) e = bc_last_error(ctx);
    } else if (!ctx) {
        e = "bch_ingest_obj needs a device context";
    }
    if (rc != BC_OK) {
        if (err && err_cap > 0) std::snprintf(err, (size_t)err_cap, "%s", e.c_str());
        delete s;
        return nullptr;
    }
    return s;
}
