// Reader of Bullet's serialized scene files and the scene conversion of the
// reference's btBulletDataExtractor
// (opencl/gpu_rigidbody_pipeline2/BulletDataExtractor.cpp:594-1010) on top of
// the library's own registration calls.  Host code only: no device work.
#include "bullet_file.h"

#include <cmath>
#include <cstdio>
#include <cstring>
#include <functional>

#include "../../include/mi_rb.h"

namespace bullet {

namespace {

constexpr uint32_t code4(const char *s) {
    return (uint32_t)(uint8_t)s[0] | ((uint32_t)(uint8_t)s[1] << 8) |
           ((uint32_t)(uint8_t)s[2] << 16) | ((uint32_t)(uint8_t)s[3] << 24);
}
constexpr uint32_t CODE_ENDB = code4("ENDB"), CODE_DNA1 = code4("DNA1"), CODE_SHAP = code4("SHAP"),
                   CODE_RBDY = code4("RBDY");

// "*m_name[3][2]" -> base name, pointer flag, element count
bool parse_name(const std::string &s, std::string &base, bool &pointer, int &count) {
    size_t i = 0;
    pointer = false;
    if (s.compare(0, 2, "(*") == 0) {  // function pointer
        pointer = true;
        i = 2;
    }
    while (i < s.size() && s[i] == '*') {
        pointer = true;
        i++;
    }
    size_t j = i;
    while (j < s.size() && s[j] != '[' && s[j] != ')') j++;
    base = s.substr(i, j - i);
    count = 1;
    while (j < s.size()) {
        if (s[j] == '[') {
            size_t k = s.find(']', j);
            if (k == std::string::npos) return false;
            long v = std::strtol(s.c_str() + j + 1, nullptr, 10);
            if (v <= 0 || v > (1 << 20)) return false;
            count *= (int)v;
            if (count > (1 << 24)) return false;
            j = k + 1;
        } else {
            j++;
        }
    }
    return !base.empty();
}

}  // namespace

uint32_t File::u32(size_t at) const {
    uint32_t v;
    std::memcpy(&v, bytes.data() + at, 4);
    return swap ? __builtin_bswap32(v) : v;
}
uint16_t File::u16(size_t at) const {
    uint16_t v;
    std::memcpy(&v, bytes.data() + at, 2);
    return swap ? __builtin_bswap16(v) : v;
}
uint64_t File::u64(size_t at) const {
    uint64_t v;
    std::memcpy(&v, bytes.data() + at, 8);
    return swap ? __builtin_bswap64(v) : v;
}

bool File::parse(std::vector<uint8_t> data, std::string &err) {
    bytes = std::move(data);
    const size_t n = bytes.size();
    if (n < 12 || std::memcmp(bytes.data(), "BULLET", 6) != 0) {
        err = "not a .bullet file (no BULLET header)";
        return false;
    }
    const char fp = (char)bytes[6], ps = (char)bytes[7], en = (char)bytes[8];
    if ((fp != 'f' && fp != 'd') || (ps != '_' && ps != '-') || (en != 'v' && en != 'V')) {
        err = "unknown header flags";
        return false;
    }
    doubles = fp == 'd';
    ptr_size = ps == '-' ? 8 : 4;
    swap = en == 'V';  // the readers below assume a little-endian host
    version = 0;
    for (int i = 9; i < 12; i++) {
        if (bytes[i] < '0' || bytes[i] > '9') {
            err = "bad version digits";
            return false;
        }
        version = version * 10 + (bytes[i] - '0');
    }
    // chunks: code, len, old pointer, struct index, count
    const size_t hdr = 16 + (size_t)ptr_size;
    size_t at = 12;
    int dna_chunk = -1;
    chunks.clear();
    while (at < n) {  // up to ENDB or the end of the file
        if (at + hdr > n) {
            err = "truncated chunk header";
            return false;
        }
        Chunk c;
        std::memcpy(&c.code, bytes.data() + at, 4);
        c.len = (int32_t)u32(at + 4);
        c.old_ptr = ptr_size == 8 ? u64(at + 8) : u32(at + 8);
        c.sdna = (int32_t)u32(at + 8 + ptr_size);
        c.nr = (int32_t)u32(at + 12 + ptr_size);
        c.data = at + hdr;
        if (c.len < 0 || c.data + (size_t)c.len > n) {
            err = "chunk data past the end of the file";
            return false;
        }
        if (c.code == CODE_ENDB) break;
        if (c.code == CODE_DNA1) dna_chunk = (int)chunks.size();
        chunks.push_back(c);
        at = c.data + (size_t)c.len;
        if (chunks.size() > (1u << 24)) {
            err = "too many chunks";
            return false;
        }
    }
    if (dna_chunk < 0) {
        err = "no DNA1 chunk";
        return false;
    }
    // DNA: SDNA, NAME, TYPE, TLEN, STRC; sections aligned to 4 bytes from
    // the start of the chunk
    const Chunk &dc = chunks[dna_chunk];
    const size_t d0 = dc.data, d1 = dc.data + (size_t)dc.len;
    size_t p = d0;
    auto tag = [&](const char *t) {
        if (p + 4 > d1 || std::memcmp(bytes.data() + p, t, 4) != 0) return false;
        p += 4;
        return true;
    };
    auto align = [&]() { p = d0 + ((p - d0 + 3) & ~(size_t)3); };
    auto strings = [&](std::vector<std::string> &out) {
        if (p + 4 > d1) return false;
        const uint32_t cnt = u32(p);
        p += 4;
        if (cnt > (1u << 20)) return false;
        out.clear();
        for (uint32_t i = 0; i < cnt; i++) {
            size_t e = p;
            while (e < d1 && bytes[e] != 0) e++;
            if (e >= d1) return false;
            out.emplace_back((const char *)bytes.data() + p, e - p);
            p = e + 1;
        }
        align();
        return true;
    };
    if (!tag("SDNA") || !tag("NAME") || !strings(names) || !tag("TYPE") || !strings(types) ||
        !tag("TLEN")) {
        err = "malformed DNA (names or types)";
        return false;
    }
    if (p + 2 * types.size() > d1) {
        err = "malformed DNA (type lengths)";
        return false;
    }
    tlen.resize(types.size());
    for (size_t i = 0; i < types.size(); i++) tlen[i] = u16(p + 2 * i);
    p += 2 * types.size();
    align();
    if (!tag("STRC") || p + 4 > d1) {
        err = "malformed DNA (structs)";
        return false;
    }
    const uint32_t ns = u32(p);
    p += 4;
    if (ns > (1u << 16)) {
        err = "malformed DNA (struct count)";
        return false;
    }
    structs.clear();
    struct_of_type.clear();
    for (uint32_t i = 0; i < ns; i++) {
        if (p + 4 > d1) {
            err = "malformed DNA (struct header)";
            return false;
        }
        Struct s;
        s.type = u16(p);
        const int nf = u16(p + 2);
        p += 4;
        if (s.type >= (int)types.size() || p + 4 * (size_t)nf > d1) {
            err = "malformed DNA (struct fields)";
            return false;
        }
        s.size = (size_t)tlen[s.type];
        size_t off = 0;
        for (int k = 0; k < nf; k++) {
            Field f;
            f.type = u16(p);
            const int nm = u16(p + 2);
            p += 4;
            if (f.type >= (int)types.size() || nm >= (int)names.size() ||
                !parse_name(names[nm], f.name, f.pointer, f.count)) {
                err = "malformed DNA (field)";
                return false;
            }
            const size_t elem = f.pointer ? (size_t)ptr_size : (size_t)tlen[f.type];
            f.offset = off;
            f.size = elem * (size_t)f.count;
            off += f.size;
            s.fields.push_back(f);
        }
        if (off != s.size) {
            err = "DNA struct " + types[s.type] + ": fields do not add up to its length";
            return false;
        }
        struct_of_type[types[s.type]] = (int)structs.size();
        structs.push_back(std::move(s));
    }
    chunk_of_ptr.clear();
    for (size_t i = 0; i < chunks.size(); i++) {
        const Chunk &c = chunks[i];
        if ((int)i == dna_chunk) continue;
        if (c.sdna < 0 || c.sdna >= (int)structs.size()) {
            err = "chunk with an unknown struct index";
            return false;
        }
        if (c.nr < 0 || (size_t)c.nr * structs[c.sdna].size > (size_t)c.len) {
            err = "chunk shorter than its elements";
            return false;
        }
        if (c.old_ptr) chunk_of_ptr[c.old_ptr] = i;
    }
    return true;
}

const Chunk *File::deref(uint64_t ptr) const {
    auto it = chunk_of_ptr.find(ptr);
    return it == chunk_of_ptr.end() ? nullptr : &chunks[it->second];
}

const char *File::type_of(const Chunk &c) const {
    if (c.sdna < 0 || c.sdna >= (int)structs.size()) return "";
    return types[structs[c.sdna].type].c_str();
}

bool File::locate(const Chunk &c, int elem, const char *path, int index, const Field **fout,
                  size_t &at) const {
    if (c.sdna < 0 || c.sdna >= (int)structs.size() || elem < 0 || elem >= c.nr) return false;
    int cur = c.sdna;
    size_t off = (size_t)elem * structs[cur].size;
    const Field *f = nullptr;
    const char *s = path;
    while (*s) {
        // component: name, optional [k]
        const char *e = s;
        while (*e && *e != '.' && *e != '[') e++;
        const std::string name(s, (size_t)(e - s));
        int sub = -1;
        if (*e == '[') {
            sub = std::atoi(e + 1);
            while (*e && *e != ']') e++;
            if (*e != ']') return false;
            e++;
        }
        // depth-first search through nested structs held by value
        struct Hit {
            const Field *f;
            size_t off;
        };
        Hit hit{nullptr, 0};
        int guard = 0;
        std::function<bool(int, size_t, int)> dfs = [&](int si, size_t base, int depth) -> bool {
            if (depth > 16 || ++guard > 4096) return false;
            for (const Field &fd : structs[si].fields) {
                if (fd.name == name) {
                    hit = {&fd, base + fd.offset};
                    return true;
                }
                if (!fd.pointer && fd.count == 1) {
                    auto it = struct_of_type.find(types[fd.type]);
                    if (it != struct_of_type.end() && dfs(it->second, base + fd.offset, depth + 1))
                        return true;
                }
            }
            return false;
        };
        if (!dfs(cur, 0, 0)) return false;
        f = hit.f;
        off += hit.off;
        const size_t esz = f->pointer ? (size_t)ptr_size : (size_t)tlen[f->type];
        if (sub >= 0) {
            if (sub >= f->count) return false;
            off += (size_t)sub * esz;
        }
        s = *e == '.' ? e + 1 : e;
        if (*s) {  // descend into the member's struct
            if (f->pointer || (f->count != 1 && sub < 0)) return false;
            auto it = struct_of_type.find(types[f->type]);
            if (it == struct_of_type.end()) return false;
            cur = it->second;
        } else if (index >= 0) {
            if (index >= f->count) return false;
            off += (size_t)index * esz;
        }
    }
    if (!f) return false;
    const size_t esz = f->pointer ? (size_t)ptr_size : (size_t)tlen[f->type];
    if (off + esz > (size_t)c.len) return false;
    *fout = f;
    at = c.data + off;
    return true;
}

bool File::read_float(const Chunk &c, int elem, const char *path, int index, float &out) const {
    const Field *f;
    size_t at;
    if (!locate(c, elem, path, index, &f, at) || f->pointer) return false;
    if (types[f->type] == "float") {
        uint32_t v = u32(at);
        std::memcpy(&out, &v, 4);
        return true;
    }
    if (types[f->type] == "double") {
        uint64_t v = u64(at);
        double d;
        std::memcpy(&d, &v, 8);
        out = (float)d;
        return true;
    }
    return false;
}

bool File::read_int(const Chunk &c, int elem, const char *path, int index, int32_t &out) const {
    const Field *f;
    size_t at;
    if (!locate(c, elem, path, index, &f, at) || f->pointer || types[f->type] != "int") return false;
    out = (int32_t)u32(at);
    return true;
}

bool File::read_ptr(const Chunk &c, int elem, const char *path, uint64_t &out) const {
    const Field *f;
    size_t at;
    if (!locate(c, elem, path, 0, &f, at) || !f->pointer) return false;
    out = ptr_size == 8 ? u64(at) : u32(at);
    return true;
}

bool extract(const File &f, Scene &out, std::string &err) {
    out = Scene();
    if (f.doubles) {
        err = "double-precision .bullet files are not supported";
        return false;
    }
    std::unordered_map<uint64_t, int> shape_of_ptr;
    for (const Chunk &c : f.chunks) {
        if (c.code != CODE_SHAP) continue;
        for (int e = 0; e < c.nr; e++) {
            ShapeRecord s;
            int32_t type = -1;
            if (!f.read_int(c, e, "m_shapeType", -1, type)) {
                err = "shape without m_shapeType";
                return false;
            }
            s.type = type;
            s.old_ptr = c.old_ptr + (uint64_t)e * f.structs[c.sdna].size;
            float v;
            for (int k = 0; k < 3; k++) {
                if (f.read_float(c, e, "m_localScaling.m_floats", k, v)) s.local_scaling[k] = v;
                if (f.read_float(c, e, "m_implicitShapeDimensions.m_floats", k, v))
                    s.implicit_dims[k] = v;
                if (f.read_float(c, e, "m_planeNormal.m_floats", k, v)) s.plane_normal[k] = v;
            }
            if (f.read_float(c, e, "m_collisionMargin", -1, v)) s.margin = v;
            if (f.read_float(c, e, "m_planeConstant", -1, v)) s.plane_constant = v;
            if (type == CONVEX_HULL) {
                uint64_t ptr = 0;
                int32_t np = 0;
                f.read_ptr(c, e, "m_unscaledPointsFloatPtr", ptr);
                f.read_int(c, e, "m_numUnscaledPoints", -1, np);
                const Chunk *pc = ptr ? f.deref(ptr) : nullptr;
                if (pc && np > 0) {
                    if (np > pc->nr) {
                        err = "convex hull with more points than its array";
                        return false;
                    }
                    for (int i = 0; i < np; i++)
                        for (int k = 0; k < 3; k++) {
                            float x;
                            if (!f.read_float(*pc, i, "m_floats", k, x)) {
                                err = "unreadable hull point";
                                return false;
                            }
                            s.points.push_back(x);
                        }
                }
            }
            shape_of_ptr[s.old_ptr] = (int)out.shapes.size();
            out.shapes.push_back(std::move(s));
        }
    }
    for (const Chunk &c : f.chunks) {
        if (c.code != CODE_RBDY) continue;
        for (int e = 0; e < c.nr; e++) {
            BodyRecord b;
            uint64_t sp = 0;
            if (f.read_ptr(c, e, "m_collisionShape", sp)) {
                auto it = shape_of_ptr.find(sp);
                if (it != shape_of_ptr.end()) b.shape = it->second;
            }
            if (!f.read_float(c, e, "m_inverseMass", -1, b.inverse_mass)) {
                err = "rigid body without m_inverseMass";
                return false;
            }
            char path[96];
            for (int r = 0; r < 3; r++)
                for (int k = 0; k < 3; k++) {
                    std::snprintf(path, sizeof(path), "m_worldTransform.m_basis.m_el[%d].m_floats", r);
                    if (!f.read_float(c, e, path, k, b.basis[3 * r + k])) {
                        err = "rigid body without a world transform";
                        return false;
                    }
                }
            for (int k = 0; k < 3; k++) {
                if (!f.read_float(c, e, "m_worldTransform.m_origin.m_floats", k, b.origin[k])) {
                    err = "rigid body without a world transform";
                    return false;
                }
                f.read_float(c, e, "m_linearVelocity.m_floats", k, b.linvel[k]);
                f.read_float(c, e, "m_angularVelocity.m_floats", k, b.angvel[k]);
            }
            f.read_float(c, e, "m_friction", -1, b.friction);
            f.read_float(c, e, "m_restitution", -1, b.restitution);
            out.bodies.push_back(b);
        }
    }
    return true;
}

void basis_to_quat(const float mf[9], float q[4]) {
    double m[9];
    for (int i = 0; i < 9; i++) m[i] = mf[i];
    const double tr = m[0] + m[4] + m[8];
    double x, y, z, w;
    if (tr > 0.0) {
        double s = std::sqrt(tr + 1.0);
        w = 0.5 * s;
        s = 0.5 / s;
        x = (m[7] - m[5]) * s;
        y = (m[2] - m[6]) * s;
        z = (m[3] - m[1]) * s;
    } else if (m[0] >= m[4] && m[0] >= m[8]) {
        double s = std::sqrt(m[0] - m[4] - m[8] + 1.0);
        x = 0.5 * s;
        s = 0.5 / s;
        y = (m[1] + m[3]) * s;
        z = (m[2] + m[6]) * s;
        w = (m[7] - m[5]) * s;
    } else if (m[4] >= m[8]) {
        double s = std::sqrt(m[4] - m[8] - m[0] + 1.0);
        y = 0.5 * s;
        s = 0.5 / s;
        x = (m[1] + m[3]) * s;
        z = (m[5] + m[7]) * s;
        w = (m[2] - m[6]) * s;
    } else {
        double s = std::sqrt(m[8] - m[0] - m[4] + 1.0);
        z = 0.5 * s;
        s = 0.5 / s;
        x = (m[2] + m[6]) * s;
        y = (m[5] + m[7]) * s;
        w = (m[3] - m[1]) * s;
    }
    q[0] = (float)x;
    q[1] = (float)y;
    q[2] = (float)z;
    q[3] = (float)w;
}

static bool read_path(const char *path, std::vector<uint8_t> &out) {
    FILE *fp = std::fopen(path, "rb");
    if (!fp) return false;
    out.clear();
    uint8_t buf[1 << 16];
    size_t k;
    while ((k = std::fread(buf, 1, sizeof(buf), fp)) > 0) {
        out.insert(out.end(), buf, buf + k);
        if (out.size() > ((size_t)1 << 31)) break;
    }
    std::fclose(fp);
    return true;
}

}  // namespace bullet

using namespace bullet;

namespace {
mi_rb_status open_scene(const char *path, File &f, Scene &sc) {
    if (!path) return MI_RB_INVALID_ARG;
    std::vector<uint8_t> data;
    if (!read_path(path, data)) return MI_RB_INVALID_ARG;
    std::string err;
    if (!f.parse(std::move(data), err) || !extract(f, sc, err)) return MI_RB_FORMAT;
    return MI_RB_OK;
}
}  // namespace

extern "C" mi_rb_status mi_rb_bullet_read(const char *path, mi_rb_bullet_shape *shapes,
                                          int max_shapes, mi_rb_bullet_body *bodies,
                                          int max_bodies, float *points, int max_points,
                                          int counts[4]) {
    if (!counts || max_shapes < 0 || max_bodies < 0 || max_points < 0) return MI_RB_INVALID_ARG;
    File f;
    Scene sc;
    mi_rb_status st = open_scene(path, f, sc);
    if (st != MI_RB_OK) return st;
    int npts = 0;
    for (size_t i = 0; i < sc.shapes.size(); i++) {
        const ShapeRecord &s = sc.shapes[i];
        const int np = (int)(s.points.size() / 3);
        if (shapes && (int)i < max_shapes) {
            mi_rb_bullet_shape &o = shapes[i];
            o.type = s.type;
            o.num_points = np;
            o.first_point = npts;
            for (int k = 0; k < 3; k++) {
                o.local_scaling[k] = s.local_scaling[k];
                o.implicit_dims[k] = s.implicit_dims[k];
                o.plane_normal[k] = s.plane_normal[k];
            }
            o.margin = s.margin;
            o.plane_constant = s.plane_constant;
        }
        for (int p = 0; p < np; p++, npts++)
            if (points && npts < max_points)
                for (int k = 0; k < 3; k++) points[3 * npts + k] = s.points[3 * p + k];
    }
    for (size_t i = 0; i < sc.bodies.size() && bodies && (int)i < max_bodies; i++) {
        const BodyRecord &b = sc.bodies[i];
        mi_rb_bullet_body &o = bodies[i];
        o.shape = b.shape;
        o.inverse_mass = b.inverse_mass;
        std::memcpy(o.basis, b.basis, sizeof(o.basis));
        std::memcpy(o.origin, b.origin, sizeof(o.origin));
        std::memcpy(o.linvel, b.linvel, sizeof(o.linvel));
        std::memcpy(o.angvel, b.angvel, sizeof(o.angvel));
        basis_to_quat(b.basis, o.quat);
        o.friction = b.friction;
        o.restitution = b.restitution;
    }
    counts[0] = (int)sc.shapes.size();
    counts[1] = (int)sc.bodies.size();
    counts[2] = npts;
    counts[3] = f.ptr_size | (f.swap ? 0x100 : 0);
    return MI_RB_OK;
}

extern "C" mi_rb_status mi_rb_load_bullet(void *world, const char *path, int keep_static,
                                          mi_rb_bullet_summary *out) {
    if (!world) return MI_RB_INVALID_ARG;
    File f;
    Scene sc;
    mi_rb_status st = open_scene(path, f, sc);
    if (st != MI_RB_OK) return st;
    mi_rb_bullet_summary sum{};
    sum.shapes = (int)sc.shapes.size();
    sum.bodies = (int)sc.bodies.size();
    sum.pointer_size = f.ptr_size;
    sum.big_endian = f.swap ? 1 : 0;
    // collision shapes in file order: boxes and convex hulls become hulls,
    // every other type is left out, as the reference's extractor does
    std::vector<int> hull(sc.shapes.size(), -1);
    for (size_t i = 0; i < sc.shapes.size(); i++) {
        const ShapeRecord &s = sc.shapes[i];
        std::vector<float> pts;
        float scale[3] = {1.f, 1.f, 1.f};
        if (s.type == BOX) {
            // the unit cube's corners scaled by half extent * scaling + margin
            for (int c = 0; c < 8; c++) {
                pts.push_back(c & 1 ? 1.f : -1.f);
                pts.push_back(c & 2 ? 1.f : -1.f);
                pts.push_back(c & 4 ? 1.f : -1.f);
            }
            for (int k = 0; k < 3; k++) scale[k] = s.implicit_dims[k] * s.local_scaling[k] + s.margin;
        } else if (s.type == CONVEX_HULL) {
            for (size_t p = 0; p + 2 < s.points.size(); p += 3)
                for (int k = 0; k < 3; k++) pts.push_back(s.points[p + k] * s.local_scaling[k]);
        } else {
            sum.unsupported_shapes++;
            continue;
        }
        int idx = -1;
        st = mi_rb_register_convex_hull(world, pts.data(), 12, (int)(pts.size() / 3), scale, &idx);
        if (st == MI_RB_INVALID_ARG) {  // degenerate point set: no hull
            sum.unsupported_shapes++;
            continue;
        }
        if (st != MI_RB_OK) return st;
        hull[i] = idx;
        sum.shapes_registered++;
    }
    // bodies grouped by shape, shapes in file order, bodies in file order
    // within a shape; static bodies only with keep_static
    for (size_t i = 0; i < sc.shapes.size(); i++) {
        if (hull[i] < 0) continue;
        for (const BodyRecord &b : sc.bodies) {
            if (b.shape != (int)i) continue;
            if (!keep_static && b.inverse_mass == 0.f) continue;
            const float mass = b.inverse_mass == 0.f ? 0.f : 1.f / b.inverse_mass;
            float q[4];
            basis_to_quat(b.basis, q);
            int bi = -1;
            st = mi_rb_register_body(world, mass, b.origin, q, hull[i], 0, &bi);
            if (st != MI_RB_OK) return st;
            sum.bodies_registered++;
        }
    }
    // without the file's static bodies the reference lays a 50 x 50 grid of
    // static square pyramids (base corners (+-1, 0, +-1), apex (0, -1, 0),
    // scaled by (2.5, 2, 2.5)) at 5.01 spacing from -120
    if (!keep_static) {
        const float pyr[15] = {-1, 0, -1, -1, 0, 1, 1, 0, 1, 1, 0, -1, 0, -1, 0};
        const float gs[3] = {2.5f, 2.f, 2.5f};
        int gshape = -1;
        st = mi_rb_register_convex_hull(world, pyr, 12, 5, gs, &gshape);
        if (st != MI_RB_OK) return st;
        const float q[4] = {0, 0, 0, 1};
        for (int i = 0; i < 50; i++)
            for (int j = 0; j < 50; j++) {
                const float pos[3] = {(float)(i * 5.01 - 120.0), 0.f, (float)(j * 5.01 - 120.0)};
                int bi = -1;
                st = mi_rb_register_body(world, 0.f, pos, q, gshape, 0, &bi);
                if (st != MI_RB_OK) return st;
                sum.ground_bodies++;
            }
    }
    if (out) *out = sum;
    return MI_RB_OK;
}
