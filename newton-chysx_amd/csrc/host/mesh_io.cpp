#include "mesh_io.h"

#include <cstring>
#include <fstream>
#include <sstream>
#include <utility>

namespace chysx {
namespace io {

// ---------------------------------------------------------------------------
// OBJ (ChysX/src/io/obj_io.cpp:12-87)
// ---------------------------------------------------------------------------
bool load_obj(const std::string& path, ObjMesh& mesh) {
    std::ifstream in(path);
    if (!in) return false;
    mesh.positions.clear();
    mesh.triangles.clear();
    std::string line;
    while (std::getline(in, line)) {
        std::istringstream ls(line);
        std::string tag;
        if (!(ls >> tag)) continue;
        if (tag == "v") {
            float p[3] = {0.f, 0.f, 0.f};
            ls >> p[0] >> p[1] >> p[2];
            mesh.positions.insert(mesh.positions.end(), p, p + 3);
        } else if (tag == "f") {
            int idx[3];
            int k = 0;
            std::string entry;
            while (k < 3 && (ls >> entry)) idx[k++] = std::stoi(entry.substr(0, entry.find('/'))) - 1;
            if (k == 3) mesh.triangles.insert(mesh.triangles.end(), idx, idx + 3);
        }
    }
    return true;
}

bool save_obj(const std::string& path, const float* positions, int n_points, const int* triangles, int n_tris) {
    std::ofstream out(path);
    if (!out) return false;
    out << "# ChysX OBJ export  verts=" << n_points << " tris=" << n_tris << "\n";
    for (int i = 0; i < n_points; ++i)
        out << "v " << positions[3 * i] << " " << positions[3 * i + 1] << " " << positions[3 * i + 2] << "\n";
    for (int t = 0; t < n_tris; ++t)
        out << "f " << triangles[3 * t] + 1 << " " << triangles[3 * t + 1] + 1 << " " << triangles[3 * t + 2] + 1
            << "\n";
    return static_cast<bool>(out);
}

// ---------------------------------------------------------------------------
// Binary JSON encoder of the .bgeo format (token ids and length encoding of
// ChysX/src/io/bgeo_writer.h:55-65, bgeo_writer.cpp:18-85).  The document is
// built in memory and written at once.
// ---------------------------------------------------------------------------
namespace {

class Encoder {
public:
    enum : uint8_t {
        kArrayBegin = 0x5B, kArrayEnd = 0x5D, kMapBegin = 0x7B, kMapEnd = 0x7D, kString = 0x27,
        kInt32 = 0x13, kReal32 = 0x19, kMagic = 0x7F, kUniformArray = 0x40,
    };

    Encoder() {
        byte(kMagic);
        raw(uint32_t(0x624A534Eu));  // "NSJb" little-endian
    }
    const std::string& bytes() const noexcept { return buf_; }

    void open_array() { byte(kArrayBegin); }
    void close_array() { byte(kArrayEnd); }
    void open_map() { byte(kMapBegin); }
    void close_map() { byte(kMapEnd); }
    void str(const std::string& s) {
        byte(kString);
        length(s.size());
        buf_.append(s);
    }
    void i32(int32_t v) {
        byte(kInt32);
        raw(v);
    }
    void floats(const float* p, std::size_t count) { uniform(kReal32, p, count * sizeof(float), count); }
    void ints(const int32_t* p, std::size_t count) { uniform(kInt32, p, count * sizeof(int32_t), count); }
    // key / value pairs of the flat-array "maps" the geometry document uses
    void kv(const char* key, int32_t v) {
        str(key);
        i32(v);
    }
    void kv(const char* key, const char* v) {
        str(key);
        str(v);
    }

private:
    void byte(uint8_t b) { buf_.push_back(static_cast<char>(b)); }
    template <class T>
    void raw(T v) {
        char tmp[sizeof(T)];
        std::memcpy(tmp, &v, sizeof(T));
        buf_.append(tmp, sizeof(T));
    }
    // lengths: one byte below 0xF1, else a width tag 0xF2 / 0xF4 / 0xF8 and
    // a 16 / 32 / 64-bit count
    void length(uint64_t n) {
        if (n < 0xF1u) byte(static_cast<uint8_t>(n));
        else if (n < 0xFFFFu) { byte(0xF2); raw(static_cast<uint16_t>(n)); }
        else if (n < 0xFFFFFFFFull) { byte(0xF4); raw(static_cast<uint32_t>(n)); }
        else { byte(0xF8); raw(static_cast<int64_t>(n)); }
    }
    void uniform(uint8_t type, const void* p, std::size_t nbytes, std::size_t count) {
        byte(kUniformArray);
        byte(type);
        length(count);
        buf_.append(static_cast<const char*>(p), nbytes);
    }

    std::string buf_;
};

// One numeric attribute: [definition, data] with the values in pages of 1024.
void attribute(Encoder& e, const char* name, int tuple, std::size_t elements, const float* data, bool point_typed) {
    e.open_array();
    e.open_array();  // definition
    e.kv("scope", "public");
    e.kv("type", "numeric");
    e.kv("name", name);
    e.str("options");
    e.open_map();
    if (point_typed) {
        e.str("type");
        e.open_map();
        e.kv("type", "string");
        e.kv("value", "point");
        e.close_map();
    }
    e.close_map();
    e.close_array();
    e.open_array();  // data
    e.kv("size", tuple);
    e.kv("storage", "fpreal32");
    e.str("values");
    e.open_array();
    e.kv("size", tuple);
    e.kv("storage", "fpreal32");
    e.kv("pagesize", 1024);
    e.str("rawpagedata");
    e.floats(data, elements * static_cast<std::size_t>(tuple));
    e.close_array();
    e.close_array();
    e.close_array();
}

std::string geometry(const float* positions, int n_points, const int* triangles, int n_tris,
                     const float* point_colors, const float* face_colors) {
    Encoder e;
    const std::size_t nv = static_cast<std::size_t>(n_tris) * 3;
    e.open_array();
    e.kv("pointcount", n_points);
    e.kv("vertexcount", static_cast<int32_t>(nv));
    e.kv("primitivecount", n_tris);
    e.str("topology");
    e.open_array();
    e.str("pointref");
    e.open_array();
    e.str("indices");
    e.ints(triangles, nv);
    e.close_array();
    e.close_array();
    e.str("attributes");
    e.open_array();
    e.str("pointattributes");
    e.open_array();
    attribute(e, "P", 3, static_cast<std::size_t>(n_points), positions, true);
    if (point_colors) attribute(e, "Cd", 3, static_cast<std::size_t>(n_points), point_colors, false);
    e.close_array();
    if (face_colors) {
        e.str("primitiveattributes");
        e.open_array();
        attribute(e, "Cd", 3, static_cast<std::size_t>(n_tris), face_colors, false);
        e.close_array();
    }
    e.close_array();
    e.str("primitives");
    e.open_array();
    if (n_tris > 0) {  // one run of polygons: start vertex 0, n_tris of 3 vertices
        e.open_array();
        e.open_array();
        e.kv("type", "p_r");
        e.close_array();
        e.open_array();
        e.kv("s_v", 0);
        e.kv("n_p", n_tris);
        e.str("n_v");
        const std::vector<int32_t> three(static_cast<std::size_t>(n_tris), 3);
        e.ints(three.data(), three.size());
        e.close_array();
        e.close_array();
    }
    e.close_array();
    e.close_array();
    return e.bytes();
}

}  // namespace

bool BgeoWriter::write(const std::string& path, const float* positions, int n_points, const int* triangles,
                       int n_tris, const float* point_colors, const float* face_colors) {
    const std::string doc = geometry(positions, n_points, triangles, n_tris, point_colors, face_colors);
    std::ofstream out(path, std::ios::binary);
    if (!out) return false;
    out.write(doc.data(), static_cast<std::streamsize>(doc.size()));
    return static_cast<bool>(out);
}

bool BgeoWriter::write_multi(const std::string& path, const std::vector<BgeoMeshPiece>& pieces) {
    std::vector<float> pos, cd;
    std::vector<int> tri;
    int base = 0;
    for (const BgeoMeshPiece& p : pieces) {
        pos.insert(pos.end(), p.positions, p.positions + 3 * static_cast<std::size_t>(p.n_points));
        for (int k = 0; k < 3 * p.n_tris; ++k) tri.push_back(p.triangles[k] + base);
        for (int i = 0; i < p.n_points; ++i) cd.insert(cd.end(), {p.color_r, p.color_g, p.color_b});
        base += p.n_points;
    }
    return write(path, pos.data(), base, tri.data(), static_cast<int>(tri.size() / 3), cd.data(), nullptr);
}

void generate_box_mesh(float cx, float cy, float cz, float hx, float hy, float hz, std::vector<float>& positions,
                       std::vector<int>& triangles) {
    positions.resize(24);
    for (int i = 0; i < 8; ++i) {
        positions[3 * i] = cx + ((i & 1) ? hx : -hx);
        positions[3 * i + 1] = cy + ((i & 2) ? hy : -hy);
        positions[3 * i + 2] = cz + ((i & 4) ? hz : -hz);
    }
    // Each face: the corners with bit `axis` = `side`, walked so that the
    // triangles wind counter-clockwise seen from outside.
    triangles.clear();
    for (int axis = 2; axis >= 0; --axis) {
        const int u = 1 << ((axis + 1) % 3), v = 1 << ((axis + 2) % 3), a = 1 << axis;
        for (int side = 0; side < 2; ++side) {
            const int o = side ? a : 0;
            int q[4] = {o, o | u, o | u | v, o | v};  // counter-clockwise about +axis
            if (!side) std::swap(q[1], q[3]);          // outward = -axis
            triangles.insert(triangles.end(), {q[0], q[1], q[2], q[0], q[2], q[3]});
        }
    }
}

}  // namespace io
}  // namespace chysx
