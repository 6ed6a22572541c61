// Host topology of the cloth step — integer work whose results are
// bit-exact parity targets: unique edges, dihedral detection, the
// off-diagonal block-CSR pattern and the per-element slot tables.
#pragma once

#include <cstdint>
#include <vector>

#include "chysx_common.h"

namespace chysx {
namespace topology {

// Sorted unique undirected edges (min, max): ChysX/src/cloth/cloth_simulator.cu:330-349.
std::vector<math::Vec2i> build_edges(const math::Vec3i* tris, int n_tris);

// Interior edges shared by two consistently oriented triangles, as
// (v0, v1, v2, v3) with v0 < v1, v2 the apex of directed edge v0->v1 and v3
// the apex of v1->v0.  Emitted in the iteration order of a
// std::unordered_map<int64_t,int> reserved to 3T and filled in triangle
// order, which is the reference's order with libstdc++
// (cloth_simulator.cu:481-507).
std::vector<math::Vec4i> detect_dihedrals(const math::Vec3i* tris, int n_tris);

// Off-diagonal block-CSR pattern (row_offsets[n+1], ascending col_indices):
// ChysX/src/sparse/block_csr.cu:91-147.  Throws std::out_of_range for a pair
// outside [0, n).
struct Pattern {
    int n = 0;
    std::vector<int> row_offsets;
    std::vector<int> col_indices;
    int nnz() const { return static_cast<int>(col_indices.size()); }
};
Pattern build_pattern(int n, const int* rows, const int* cols, int num_pairs);

// Slot of block (r, c): -r-1 on the diagonal, else the index in values
// (block_csr.cu:168-204).  Throws std::out_of_range when absent.
int resolve_slot(const Pattern& p, int r, int c);
void resolve_slots(const Pattern& p, const int* rows, const int* cols, int* out, int num_pairs);

// Per-element slot tables slot[e*K*K + a*K + b] (constraint_n.h:115-153).
std::vector<int> slot_table3(const Pattern& p, const std::vector<math::Vec3i>& elems);
std::vector<int> slot_table4(const Pattern& p, const std::vector<math::Vec4i>& elems);
std::vector<int> slot_table_pins(const Pattern& p, const std::vector<int>& pins);

// Directed off-diagonal pairs of triangle / dihedral elements
// (cloth_simulator.cu:722-761): 6 per triangle, 12 per dihedral.
void append_triangle_pairs(const std::vector<math::Vec3i>& tris, std::vector<int>& rows, std::vector<int>& cols);
void append_dihedral_pairs(const std::vector<math::Vec4i>& dih, std::vector<int>& rows, std::vector<int>& cols);

// Vertex -> incident triangles (ascending), for the gather-form lumped mass.
void vertex_triangles(const math::Vec3i* tris, int n_tris, int n_vertices, std::vector<int>& offsets,
                      std::vector<int>& list);

}  // namespace topology
}  // namespace chysx
