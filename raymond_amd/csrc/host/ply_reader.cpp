// ASCII PLY reader for the meshes the Blender exporter writes
// (scripts/blender_exporter/raymond_blender/shapes.py:19-123), with the
// header checks, number parsing and bounds of the reference's reader
// (raymond/io/PLYReader.m:1-169, driven by host/scene/ShapeBuilder.swift:150-187
// and :260-283).
//
// The reference's reader is strict and so is this one: every header line must
// be exactly the exporter's, faces must be triangles, and numbers are plain
// decimals ([-]digits[.digits], no exponent), each ended by a space or a
// newline.  Where the reference asserts, this reader returns an error.
//
// Numbers: the reference accumulates all digits into an integer D and returns
// (float)D * powf(10, -k) for k digits after the point, in float arithmetic
// (PLYReader.m:59-87).  That is not always the correctly rounded value of the
// decimal: for "%.6f" text it is fl(fl(D) * fl(1e-6)).  This reader computes
// the same expression with the correctly rounded float of 10^-k.
//
// Bounds: the reference updates the maximum only when a coordinate is not a
// new minimum (`if (v < min) min = v; else if (v > max) max = v;`,
// PLYReader.m:104-108), so a coordinate that lowers the minimum never raises
// the maximum; the first vertex only sets the minimum.  The bounds feed the
// "Generated" texture coordinates (device/ShadingContext.metal:31), so this
// reader reproduces them as they are.
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../../include/rm_host.h"

namespace {

// correctly rounded float of 10^-k, k = 0..38 (glibc's strtof rounds
// correctly; the test suite checks every entry against exact rational
// rounding)
float pow10_neg(int k) {
    static const std::vector<float> table = [] {
        std::vector<float> t(39);
        for (int i = 0; i <= 38; i++) {
            char buf[16];
            snprintf(buf, sizeof buf, "1e-%d", i);
            t[i] = strtof(buf, nullptr);
        }
        return t;
    }();
    if (k < 0 || k > 38) throw std::runtime_error("number with more than 38 decimals");
    return table[k];
}

struct Cursor {
    const char* p;
    const char* end;
    size_t line = 1;

    std::string where() const { return "line " + std::to_string(line); }

    // one header line, without its newline
    std::string read_line() {
        const char* s = p;
        while (p < end && *p != '\n') p++;
        std::string l(s, p);
        if (p < end) p++;
        line++;
        return l;
    }
    void expect_line(const char* want) {
        const std::string got = read_line();
        if (got != want) throw std::runtime_error(where() + ": expected '" + want + "', got '" + got + "'");
    }
    // "element <name> <count>"
    uint32_t element_count(const char* name) {
        const std::string l = read_line();
        const std::string head = std::string("element ") + name + " ";
        if (l.compare(0, head.size(), head) != 0)
            throw std::runtime_error(where() + ": expected 'element " + name + " <count>'");
        const char* s = l.c_str() + head.size();
        char* e = nullptr;
        errno = 0;
        const long long n = strtoll(s, &e, 10);
        if (e == s || *e != '\0' || errno || n < 0 || n > 0xffffffffLL)
            throw std::runtime_error(where() + ": bad " + name + " count");
        return (uint32_t)n;
    }

    // PLYReader.m:59-87
    float number() {
        while (p < end && *p == ' ') p++;
        const bool neg = p < end && *p == '-';
        if (neg) p++;
        long long digits = 0;
        int ndigits = 0, frac = 0;
        bool point = false;
        while (true) {
            if (p >= end) throw std::runtime_error(where() + ": unexpected end of file in a number");
            const char c = *p;
            if (c >= '0' && c <= '9') {
                if (++ndigits > 18) throw std::runtime_error(where() + ": number with more than 18 digits");
                digits = digits * 10 + (c - '0');
                frac += point ? 1 : 0;
            } else if (c == '.') {
                if (point) throw std::runtime_error(where() + ": second decimal point");
                point = true;
            } else if (c == ' ' || c == '\n') {
                break;
            } else {
                throw std::runtime_error(where() + ": invalid character in a number");
            }
            p++;
        }
        const float d = (float)(neg ? -digits : digits);
        return d * pow10_neg(frac);
    }

    // strtol over leading whitespace (PLYReader.m:154-158)
    long integer() {
        while (p < end && (*p == ' ' || *p == '\n' || *p == '\t' || *p == '\r')) {
            if (*p == '\n') line++;
            p++;
        }
        const char* s = p;
        if (p < end && (*p == '-' || *p == '+')) p++;
        long v = 0;
        bool any = false;
        while (p < end && *p >= '0' && *p <= '9') {
            v = v * 10 + (*p - '0');
            if (v > 0xffffffffL) throw std::runtime_error(where() + ": integer out of range");
            p++;
            any = true;
        }
        if (!any) throw std::runtime_error(where() + ": expected an integer");
        return *s == '-' ? -v : v;
    }
};

void read_ply(const char* path, rmh_ply_mesh* out) {
    FILE* f = fopen(path, "rb");
    if (!f) throw std::runtime_error(std::string("could not open '") + path + "'");
    std::vector<char> text;
    char buf[1 << 16];
    size_t n;
    while ((n = fread(buf, 1, sizeof buf, f)) > 0) text.insert(text.end(), buf, buf + n);
    fclose(f);
    Cursor c{text.data(), text.data() + text.size()};

    // header (ShapeBuilder.swift:160-183)
    c.expect_line("ply");
    c.expect_line("format ascii 1.0");
    const std::string comment = c.read_line();
    if (comment.compare(0, 8, "comment ") != 0 && comment != "comment")
        throw std::runtime_error("line 3: expected a comment line");
    const uint32_t nv = c.element_count("vertex");
    for (const char* prop : {"x", "y", "z", "nx", "ny", "nz", "s", "t"})
        c.expect_line((std::string("property float ") + prop).c_str());
    const uint32_t nf = c.element_count("face");
    c.expect_line("property list uchar uint vertex_indices");
    c.expect_line("property uchar material_index");
    c.expect_line("end_header");

    out->num_vertices = nv;
    out->num_faces = nf;
    out->vertices = (float*)malloc(sizeof(float) * 3 * (size_t)nv + 1);
    out->normals = (float*)malloc(sizeof(float) * 3 * (size_t)nv + 1);
    out->texcoords = (float*)malloc(sizeof(float) * 2 * (size_t)nv + 1);
    out->indices = (uint32_t*)malloc(sizeof(uint32_t) * 3 * (size_t)nf + 1);
    out->face_palette = (uint8_t*)malloc((size_t)nf + 1);
    if (!out->vertices || !out->normals || !out->texcoords || !out->indices || !out->face_palette)
        throw std::runtime_error("out of host memory");

    // vertices (PLYReader.m:89-121)
    float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (uint32_t i = 0; i < nv; i++) {
        for (int j = 0; j < 3; j++) {
            const float v = c.number();
            if (v < lo[j]) lo[j] = v;
            else if (v > hi[j]) hi[j] = v;
            out->vertices[3 * (size_t)i + j] = v;
        }
        for (int j = 0; j < 3; j++) out->normals[3 * (size_t)i + j] = c.number();
        for (int j = 0; j < 2; j++) out->texcoords[2 * (size_t)i + j] = c.number();
        if (c.p >= c.end || *c.p != '\n') throw std::runtime_error(c.where() + ": expected the end of a vertex line");
        c.p++;
        c.line++;
    }
    for (int j = 0; j < 3; j++) {
        out->bounds_min[j] = lo[j];
        out->bounds_max[j] = hi[j];
    }

    // faces (PLYReader.m:123-163)
    for (uint32_t i = 0; i < nf; i++) {
        if (c.integer() != 3) throw std::runtime_error(c.where() + ": only triangles are supported");
        for (int j = 0; j < 3; j++) {
            const long idx = c.integer();
            if (idx < 0) throw std::runtime_error(c.where() + ": negative vertex index");
            out->indices[3 * (size_t)i + j] = (uint32_t)idx;
        }
        const long m = c.integer();
        if (m < 0 || m > 255) throw std::runtime_error(c.where() + ": material index is not a uchar");
        out->face_palette[i] = (uint8_t)m;
    }
}

}  // namespace

extern "C" {

int rmh_ply_read(const char* path, rmh_ply_mesh* out, char** out_error) {
    if (out_error) *out_error = nullptr;
    if (!path || !out) return -1;
    memset(out, 0, sizeof *out);
    try {
        read_ply(path, out);
        return 0;
    } catch (const std::exception& e) {
        rmh_ply_free(out);
        if (out_error) *out_error = strdup(e.what());
        return -1;
    }
}

void rmh_ply_free(rmh_ply_mesh* mesh) {
    if (!mesh) return;
    free(mesh->vertices);
    free(mesh->normals);
    free(mesh->texcoords);
    free(mesh->indices);
    free(mesh->face_palette);
    memset(mesh, 0, sizeof *mesh);
}

}  // extern "C"
