#include "topology.h"

#include <algorithm>
#include <stdexcept>
#include <unordered_map>
#include <utility>

namespace chysx {
namespace topology {

std::vector<math::Vec2i> build_edges(const math::Vec3i* tris, int n_tris) {
    std::vector<std::pair<int, int>> e;
    e.reserve(static_cast<std::size_t>(n_tris) * 3);
    for (int t = 0; t < n_tris; ++t) {
        const math::Vec3i& f = tris[t];
        e.emplace_back(std::min(f.x, f.y), std::max(f.x, f.y));
        e.emplace_back(std::min(f.y, f.z), std::max(f.y, f.z));
        e.emplace_back(std::min(f.z, f.x), std::max(f.z, f.x));
    }
    std::sort(e.begin(), e.end());
    e.erase(std::unique(e.begin(), e.end()), e.end());
    std::vector<math::Vec2i> out;
    out.reserve(e.size());
    for (const auto& p : e) out.emplace_back(p.first, p.second);
    return out;
}

std::vector<math::Vec4i> detect_dihedrals(const math::Vec3i* tris, int n_tris) {
    const auto key = [](int from, int to) -> std::int64_t {
        return (static_cast<std::int64_t>(from) << 32) | static_cast<std::int64_t>(static_cast<std::uint32_t>(to));
    };
    std::unordered_map<std::int64_t, int> apex;
    apex.reserve(static_cast<std::size_t>(3 * n_tris));
    for (int t = 0; t < n_tris; ++t) {
        const math::Vec3i& f = tris[t];
        apex[key(f.x, f.y)] = f.z;
        apex[key(f.y, f.z)] = f.x;
        apex[key(f.z, f.x)] = f.y;
    }
    std::vector<math::Vec4i> out;
    out.reserve(static_cast<std::size_t>(2 * n_tris));
    for (const auto& kv : apex) {
        const int a = static_cast<int>(kv.first >> 32);
        const int b = static_cast<int>(kv.first & 0xffffffff);
        if (a >= b) continue;
        const auto back = apex.find(key(b, a));
        if (back == apex.end()) continue;
        out.emplace_back(a, b, kv.second, back->second);
    }
    return out;
}

Pattern build_pattern(int n, const int* rows, const int* cols, int num_pairs) {
    if (n < 0 || num_pairs < 0) throw std::invalid_argument("BlockCSR3::build_topology: negative size");
    Pattern p;
    p.n = n;
    std::vector<std::pair<int, int>> pr;
    pr.reserve(static_cast<std::size_t>(num_pairs));
    for (int k = 0; k < num_pairs; ++k) {
        const int r = rows[k], c = cols[k];
        if (r < 0 || r >= n || c < 0 || c >= n) throw std::out_of_range("BlockCSR3::build_topology: pair out of range");
        if (r != c) pr.emplace_back(r, c);
    }
    std::sort(pr.begin(), pr.end());
    pr.erase(std::unique(pr.begin(), pr.end()), pr.end());
    p.row_offsets.assign(static_cast<std::size_t>(n) + 1, 0);
    p.col_indices.resize(pr.size());
    for (std::size_t i = 0; i < pr.size(); ++i) {
        ++p.row_offsets[pr[i].first + 1];
        p.col_indices[i] = pr[i].second;
    }
    for (int i = 0; i < n; ++i) p.row_offsets[i + 1] += p.row_offsets[i];
    return p;
}

int resolve_slot(const Pattern& p, int r, int c) {
    if (r < 0 || r >= p.n || c < 0 || c >= p.n) throw std::out_of_range("BlockCSR3::resolve_slots: pair out of range");
    if (r == c) return -r - 1;
    const auto b = p.col_indices.begin() + p.row_offsets[r];
    const auto e = p.col_indices.begin() + p.row_offsets[r + 1];
    const auto it = std::lower_bound(b, e, c);
    if (it == e || *it != c)
        throw std::out_of_range("BlockCSR3::resolve_slots: off-diagonal (row, col) not in topology");
    return static_cast<int>(it - p.col_indices.begin());
}

void resolve_slots(const Pattern& p, const int* rows, const int* cols, int* out, int num_pairs) {
    if (num_pairs < 0) throw std::invalid_argument("BlockCSR3::resolve_slots: negative count");
    if (num_pairs > 0 && (!rows || !cols || !out)) throw std::invalid_argument("BlockCSR3::resolve_slots: null pointer");
    for (int k = 0; k < num_pairs; ++k) out[k] = resolve_slot(p, rows[k], cols[k]);
}

std::vector<int> slot_table3(const Pattern& p, const std::vector<math::Vec3i>& el) {
    std::vector<int> s(el.size() * 9);
    for (std::size_t e = 0; e < el.size(); ++e) {
        const int v[3] = {el[e].x, el[e].y, el[e].z};
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b) s[e * 9 + a * 3 + b] = resolve_slot(p, v[a], v[b]);
    }
    return s;
}

std::vector<int> slot_table4(const Pattern& p, const std::vector<math::Vec4i>& el) {
    std::vector<int> s(el.size() * 16);
    for (std::size_t e = 0; e < el.size(); ++e) {
        const int v[4] = {el[e].x, el[e].y, el[e].z, el[e].w};
        for (int a = 0; a < 4; ++a)
            for (int b = 0; b < 4; ++b) s[e * 16 + a * 4 + b] = resolve_slot(p, v[a], v[b]);
    }
    return s;
}

std::vector<int> slot_table_pins(const Pattern& p, const std::vector<int>& pins) {
    std::vector<int> s(pins.size());
    for (std::size_t c = 0; c < pins.size(); ++c) s[c] = resolve_slot(p, pins[c], pins[c]);
    return s;
}

void append_triangle_pairs(const std::vector<math::Vec3i>& tris, std::vector<int>& rows, std::vector<int>& cols) {
    rows.reserve(rows.size() + 6 * tris.size());
    cols.reserve(cols.size() + 6 * tris.size());
    for (const auto& t : tris) {
        const int v[3] = {t.x, t.y, t.z};
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b)
                if (a != b && v[a] != v[b]) {
                    rows.push_back(v[a]);
                    cols.push_back(v[b]);
                }
    }
}

void append_dihedral_pairs(const std::vector<math::Vec4i>& dih, std::vector<int>& rows, std::vector<int>& cols) {
    rows.reserve(rows.size() + 12 * dih.size());
    cols.reserve(cols.size() + 12 * dih.size());
    for (const auto& d : dih) {
        const int v[4] = {d.x, d.y, d.z, d.w};
        for (int a = 0; a < 4; ++a)
            for (int b = 0; b < 4; ++b)
                if (a != b && v[a] != v[b]) {
                    rows.push_back(v[a]);
                    cols.push_back(v[b]);
                }
    }
}

void vertex_triangles(const math::Vec3i* tris, int n_tris, int n_vertices, std::vector<int>& offsets,
                      std::vector<int>& list) {
    offsets.assign(static_cast<std::size_t>(n_vertices) + 1, 0);
    for (int t = 0; t < n_tris; ++t) {
        const int v[3] = {tris[t].x, tris[t].y, tris[t].z};
        for (int a = 0; a < 3; ++a) {
            if (v[a] < 0 || v[a] >= n_vertices)
                throw std::out_of_range("redistribute_mass_area_weighted: triangle index out of range");
            ++offsets[v[a] + 1];
        }
    }
    for (int i = 0; i < n_vertices; ++i) offsets[i + 1] += offsets[i];
    list.assign(offsets[n_vertices], 0);
    std::vector<int> fill(offsets.begin(), offsets.end() - 1);
    for (int t = 0; t < n_tris; ++t) {
        list[fill[tris[t].x]++] = t;
        list[fill[tris[t].y]++] = t;
        list[fill[tris[t].z]++] = t;
    }
}

}  // namespace topology
}  // namespace chysx
