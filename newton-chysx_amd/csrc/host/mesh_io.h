// chysx::io — frame writers of the headless exporter: Wavefront OBJ
// (ChysX/src/io/obj_io.h:13-28) and Houdini binary-JSON .bgeo triangle meshes
// (ChysX/src/io/bgeo_writer.h:20-99; used per frame by
// ChysX/src/render/batch_export.cu:38-69).
//
// The .bgeo document is a binary JSON stream: a magic token, then arrays,
// maps, strings, int32 scalars and uniform (typed, length-prefixed) arrays.
// The geometry document holds the point / vertex / primitive counts, the
// vertex -> point topology, point attributes P (fpreal32 x3, typed "point")
// and optional Cd, optional primitive Cd, and one polygon run of triangles.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace chysx {
namespace io {

struct ObjMesh {
    std::vector<float> positions;  // xyz per point
    std::vector<int> triangles;    // 0-based indices, 3 per triangle
};

// "v x y z" and "f a b c" lines (face entries "a", "a/t", "a/t/n", 1-based);
// other lines are skipped.  Returns false when the file cannot be opened.
bool load_obj(const std::string& path, ObjMesh& mesh);

// Writes "v" lines then 1-based "f" lines.  Returns false on I/O failure.
bool save_obj(const std::string& path, const float* positions, int n_points, const int* triangles, int n_tris);

struct BgeoMeshPiece {
    const float* positions;
    int n_points;
    const int* triangles;
    int n_tris;
    float color_r, color_g, color_b;
};

struct BgeoWriter {
    // One triangle mesh; optional per-point / per-triangle rgb colours.
    static bool write(const std::string& path, const float* positions, int n_points, const int* triangles,
                      int n_tris, const float* point_colors = nullptr, const float* face_colors = nullptr);
    // Several pieces merged into one mesh (indices offset per piece), each
    // piece's uniform colour as point Cd.
    static bool write_multi(const std::string& path, const std::vector<BgeoMeshPiece>& pieces);
};

// Axis-aligned box mesh: 8 corners (corner i has +x if bit 0, +y if bit 1,
// +z if bit 2), 12 outward-facing triangles.
void generate_box_mesh(float cx, float cy, float cz, float hx, float hy, float hz, std::vector<float>& positions,
                       std::vector<int>& triangles);

}  // namespace io
}  // namespace chysx
