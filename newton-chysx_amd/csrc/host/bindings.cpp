// pybind11 module `_chysx_native`: the reference's Python surface for the
// cloth step (ChysX/src/bindings.cpp:21-595,1068-1075) over the MI355X
// implementation.  Device buffers arrive as integer pointers (for example
// torch.Tensor.data_ptr() on ROCm); numpy inputs are force-cast to
// C-contiguous int32 / float32 like the reference's bindings.
#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

#include <stdexcept>

#include "cloth_batch.h"
#include "mesh_io.h"
#include "cloth_simulator.h"
#include "sparse_solver.h"
#include "topology.h"

namespace py = pybind11;
using chysx::cloth::ClothMaterial;
using chysx::cloth::ClothSimulator;

namespace {

using IArr = py::array_t<int, py::array::c_style | py::array::forcecast>;
using FArr = py::array_t<float, py::array::c_style | py::array::forcecast>;

const chysx::math::Vec3i* as_tris(const IArr& t, const char* who) {
    if (t.ndim() != 2 || t.shape(1) != 3)
        throw std::invalid_argument(std::string(who) + ": triangles must have shape (M, 3) and dtype int32");
    return reinterpret_cast<const chysx::math::Vec3i*>(t.data());
}

py::array_t<int> to_array(const std::vector<int>& v) {
    py::array_t<int> a(static_cast<py::ssize_t>(v.size()));
    std::copy(v.begin(), v.end(), a.mutable_data());
    return a;
}

}  // namespace


// Solver trace as an (E, 2) uint64 array {stamp, active}, cut after the last
// written entry.
py::array_t<unsigned long long> trace_array(const std::vector<unsigned long long>& t) {
    std::size_t e = 0;
    while (2 * e + 1 < t.size() && t[2 * e] != 0) ++e;
    py::array_t<unsigned long long> a({static_cast<py::ssize_t>(e), static_cast<py::ssize_t>(2)});
    if (e) std::memcpy(a.mutable_data(), t.data(), 2 * e * sizeof(unsigned long long));
    return a;
}

PYBIND11_MODULE(_chysx_native, m) {
    m.doc() = "ChysX cloth step on AMD Instinct MI355X (gfx950 HIP kernels behind a C ABI)";

    py::class_<ClothMaterial>(m, "ClothMaterial")
        .def(py::init<>())
        .def_readwrite("lame_mu", &ClothMaterial::lame_mu)
        .def_readwrite("lame_lambda", &ClothMaterial::lame_lambda)
        .def_readwrite("bending", &ClothMaterial::bending)
        .def_readwrite("density", &ClothMaterial::density)
        .def_readwrite("damping", &ClothMaterial::damping)
        .def_readwrite("gx", &ClothMaterial::gx)
        .def_readwrite("gy", &ClothMaterial::gy)
        .def_readwrite("gz", &ClothMaterial::gz);

    py::class_<ClothSimulator>(m, "ClothSimulator")
        .def(py::init<>())
        .def("set_material", &ClothSimulator::set_material, py::arg("material"))
        .def("set_external_buffers", &ClothSimulator::set_external_buffers, py::arg("pos_ptr"), py::arg("vel_ptr"),
             py::arg("particle_count"), py::arg("inv_mass_ptr") = 0, py::arg("force_ptr") = 0)
        .def("step", &ClothSimulator::step, py::arg("dt"), py::arg("cuda_stream") = 0, py::call_guard<py::gil_scoped_release>())
        .def("prepare", &ClothSimulator::prepare, py::arg("dt"), py::arg("cuda_stream") = 0,
             py::call_guard<py::gil_scoped_release>(),
             "Builds the pattern, scatter maps and device arrays ahead of the first step.")
        .def(
            "set_pins",
            [](ClothSimulator& self, IArr indices, FArr targets, float stiffness) {
                if (indices.ndim() != 1) throw std::invalid_argument("ClothSimulator.set_pins: indices must be 1-D");
                if (targets.ndim() != 2 || targets.shape(1) != 3)
                    throw std::invalid_argument("ClothSimulator.set_pins: targets must have shape (N, 3)");
                if (indices.shape(0) != targets.shape(0))
                    throw std::invalid_argument(
                        "ClothSimulator.set_pins: indices and targets must have the same length");
                self.set_pins(indices.data(), reinterpret_cast<const chysx::math::Vec3f*>(targets.data()),
                              static_cast<int>(indices.shape(0)), stiffness);
            },
            py::arg("indices"), py::arg("targets"), py::arg("stiffness") = 1.0e6f)
        .def("clear_pins", &ClothSimulator::clear_pins)
        // ---- static plane / box contact (reference bindings.cpp:594-700) ----
        .def(
            "add_static_plane",
            [](ClothSimulator& self, FArr n, float d) {
                if (n.ndim() != 1 || n.shape(0) != 3)
                    throw std::invalid_argument("ClothSimulator.add_static_plane: n must be a (3,) float32 vector");
                chysx::collision::PlaneShape p;
                p.n = chysx::math::Vec3f(n.at(0), n.at(1), n.at(2));
                p.d = d;
                self.add_static_plane(p);
            },
            py::arg("n"), py::arg("d"),
            "Adds a static plane dot(n, x) + d == 0; n points to the side the cloth stays on.")
        .def(
            "add_static_box",
            [](ClothSimulator& self, FArr center, FArr half_ext, FArr ex, FArr ey, FArr ez) {
                auto v3 = [](FArr& a, const char* name) {
                    if (a.ndim() != 1 || a.shape(0) != 3)
                        throw std::invalid_argument(std::string("ClothSimulator.add_static_box: ") + name +
                                                    " must be a (3,) float32 vector");
                    return chysx::math::Vec3f(a.at(0), a.at(1), a.at(2));
                };
                chysx::collision::BoxShape b;
                b.center = v3(center, "center");
                b.half_ext = v3(half_ext, "half_ext");
                b.ex = v3(ex, "ex");
                b.ey = v3(ey, "ey");
                b.ez = v3(ez, "ez");
                self.add_static_box(b);
            },
            py::arg("center"), py::arg("half_ext"), py::arg("ex"), py::arg("ey"), py::arg("ez"),
            "Adds a static oriented box (unit frame columns ex, ey, ez; half extents along them).")
        .def("clear_static_shapes", &ClothSimulator::clear_static_shapes)
        .def("set_static_contact_thickness", &ClothSimulator::set_static_contact_thickness, py::arg("thickness"))
        .def("static_contact_thickness", &ClothSimulator::static_contact_thickness)
        .def("set_static_contact_stiffness", &ClothSimulator::set_static_contact_stiffness, py::arg("stiffness"))
        .def("static_contact_stiffness", &ClothSimulator::static_contact_stiffness)
        .def("set_static_contact_friction", &ClothSimulator::set_static_contact_friction, py::arg("mu"))
        .def("static_contact_friction", &ClothSimulator::static_contact_friction)
        .def("set_static_contact_friction_epsilon", &ClothSimulator::set_static_contact_friction_epsilon,
             py::arg("eps_u"))
        .def("static_contact_friction_epsilon", &ClothSimulator::static_contact_friction_epsilon)
        .def("static_plane_count", &ClothSimulator::static_plane_count)
        .def("static_box_count", &ClothSimulator::static_box_count)
        .def("num_pins", [](const ClothSimulator& s) { return s.pins().size(); })
        .def(
            "update_pin_targets",
            [](ClothSimulator& self, FArr targets, std::uintptr_t stream) {
                if (targets.ndim() != 2 || targets.shape(1) != 3)
                    throw std::invalid_argument(
                        "ClothSimulator.update_pin_targets: targets must have shape (N, 3) and dtype float32");
                self.update_pin_targets(reinterpret_cast<const chysx::math::Vec3f*>(targets.data()),
                                        static_cast<int>(targets.shape(0)), stream);
            },
            py::arg("targets"), py::arg("cuda_stream") = 0)
        .def(
            "set_mesh",
            [](ClothSimulator& self, IArr tris) {
                const auto* t = as_tris(tris, "ClothSimulator.set_mesh");
                const int count = static_cast<int>(tris.shape(0));
                py::gil_scoped_release release;
                self.set_mesh(t, count);
            },
            py::arg("triangles"))
        .def("build_springs_from_current_positions", &ClothSimulator::build_springs_from_current_positions,
             py::arg("stiffness"), py::arg("cuda_stream") = 0)
        .def("num_springs", [](ClothSimulator& s) { return s.springs().size(); })
        .def("redistribute_mass_area_weighted", &ClothSimulator::redistribute_mass_area_weighted,
             py::arg("surface_density"), py::arg("inv_mass_ptr"), py::arg("particle_count"),
             py::arg("cuda_stream") = 0, py::call_guard<py::gil_scoped_release>())
        .def("build_fem_stretch_from_current_positions", &ClothSimulator::build_fem_stretch_from_current_positions,
             py::arg("stiffness"), py::arg("cuda_stream") = 0, py::call_guard<py::gil_scoped_release>())
        .def("num_fem_stretch_triangles", [](const ClothSimulator& s) { return s.fem_stretch().size(); })
        .def("build_fem_shear_from_current_positions", &ClothSimulator::build_fem_shear_from_current_positions,
             py::arg("stiffness"), py::arg("cuda_stream") = 0, py::call_guard<py::gil_scoped_release>())
        .def("num_fem_shear_triangles", [](const ClothSimulator& s) { return s.fem_shear().size(); })
        .def("build_fem_stvk_from_current_positions", &ClothSimulator::build_fem_stvk_from_current_positions,
             py::arg("mu"), py::arg("lam"), py::arg("thickness"), py::arg("cuda_stream") = 0,
             py::call_guard<py::gil_scoped_release>(),
             "StVK membrane (no reference counterpart): Lame mu, lambda [Pa], thickness [m].")
        .def("num_fem_stvk_triangles", [](const ClothSimulator& s) { return s.fem_stvk().size(); })
        .def("build_bending_from_current_positions", &ClothSimulator::build_bending_from_current_positions,
             py::arg("stiffness"), py::arg("cuda_stream") = 0, py::call_guard<py::gil_scoped_release>())
        .def("num_bending_dihedrals", [](const ClothSimulator& s) { return s.bending().size(); })
        .def(
            "set_solver_type",
            [](ClothSimulator& self, int t) {
                self.set_solver_type(t == 1 ? chysx::cloth::SolverType::VBD : chysx::cloth::SolverType::PCG);
            },
            py::arg("solver_type"))
        .def("set_pcg_iterations", &ClothSimulator::set_pcg_iterations, py::arg("max_iter"))
        .def("pcg_iterations", &ClothSimulator::pcg_iterations)
        .def("set_pcg_tolerance", &ClothSimulator::set_pcg_tolerance, py::arg("tolerance"))
        .def("pcg_tolerance", &ClothSimulator::pcg_tolerance)
        .def("set_lanes_per_row", &ClothSimulator::set_lanes_per_row, py::arg("lanes"))
        .def("set_tile_rows", &ClothSimulator::set_tile_rows, py::arg("rows"))
        .def("set_lds_staging", &ClothSimulator::set_lds_staging, py::arg("enabled"))
        .def("set_trace", &ClothSimulator::set_trace, py::arg("entries"))
        .def(
            "last_trace", [](ClothSimulator& s, std::uintptr_t stream) { return trace_array(s.last_trace(stream)); },
            py::arg("cuda_stream") = 0)
        .def("tile_rows", &ClothSimulator::tile_rows)
        .def("set_grid_blocks", &ClothSimulator::set_grid_blocks, py::arg("blocks"))
        .def("lanes_per_row", &ClothSimulator::lanes_per_row)
        .def("set_bf16_blocks", &ClothSimulator::set_bf16_blocks, py::arg("enabled"))
        .def("bf16_blocks", &ClothSimulator::bf16_blocks)
        .def("last_pcg_iterations", &ClothSimulator::last_pcg_iterations, py::arg("cuda_stream") = 0)
        .def("last_relative_residual", &ClothSimulator::last_relative_residual, py::arg("cuda_stream") = 0)
        .def("set_barrier_timeout_ticks", &ClothSimulator::set_barrier_timeout_ticks, py::arg("ticks"),
             "Bound of each grid-barrier wait of the solver in 100 MHz ticks (0: one second).")
        .def("set_pcg_form", &ClothSimulator::set_pcg_form, py::arg("form"),
             "0: one grid barrier per PCG iteration (default), 1: two")
        .def("pcg_form", &ClothSimulator::pcg_form)
        .def(
            "debug_move_roundtrip",
            [](ClothSimulator& s) {
                // Moves the simulator into a temporary and back (move
                // construction, then move assignment).
                ClothSimulator tmp(std::move(s));
                s = std::move(tmp);
            },
            "Moves this simulator out and back in (exercises the move constructor and assignment).")
        .def("num_particles", &ClothSimulator::num_particles)
        .def("num_off_diag_blocks", &ClothSimulator::num_off_diag_blocks)
        .def("debug_dump_last_solve",
             [](const ClothSimulator& s) {
                 const int n = s.num_particles();
                 const int nnz = s.num_off_diag_blocks();
                 py::array_t<float> diag({n, 3, 3});
                 py::array_t<int> row_offsets({n + 1});
                 py::array_t<int> col_indices({nnz});
                 py::array_t<float> values({nnz, 3, 3});
                 py::array_t<float> rhs({n, 3});
                 py::array_t<float> dx({n, 3});
                 if (n == 0) return py::make_tuple(diag, row_offsets, col_indices, values, rhs, dx);
                 s.debug_copy_hessian_diag(diag.mutable_data());
                 s.debug_copy_hessian_csr(row_offsets.mutable_data(), nnz > 0 ? col_indices.mutable_data() : nullptr,
                                          nnz > 0 ? values.mutable_data() : nullptr);
                 s.debug_copy_last_rhs(rhs.mutable_data());
                 s.debug_copy_last_dx(dx.mutable_data());
                 return py::make_tuple(diag, row_offsets, col_indices, values, rhs, dx);
             })
        .def("debug_bf16_matrix",
             [](const ClothSimulator& s) {
                 const int n = s.num_particles();
                 const int nnz = s.num_off_diag_blocks();
                 py::array_t<float> diag_c({n, 3, 3});
                 py::array_t<uint16_t> values16({nnz, 3, 3});
                 s.debug_copy_bf16_matrix(diag_c.mutable_data(), nnz > 0 ? values16.mutable_data() : nullptr);
                 return py::make_tuple(diag_c, values16);
             })
        .def(
            "debug_set_warm_start",
            [](ClothSimulator& s, FArr dx) {
                if (dx.ndim() != 2 || dx.shape(1) != 3 || dx.shape(0) != s.buffers().particle_count())
                    throw std::invalid_argument("ClothSimulator.debug_set_warm_start: dx must have shape (N, 3)");
                s.debug_set_warm_start(dx.data());
            },
            py::arg("dx"))
        .def("compute_energy", &ClothSimulator::compute_energy, py::arg("kind_mask") = 15)
        .def("device_bytes", &ClothSimulator::device_bytes)
        .def(
            "debug_slot_tables",
            [](const ClothSimulator& s) {
                py::dict d;
                d["pins"] = to_array(s.pins().hessian_slots());
                d["stretch"] = to_array(s.fem_stretch().hessian_slots());
                d["shear"] = to_array(s.fem_shear().hessian_slots());
                d["stvk"] = to_array(s.fem_stvk().hessian_slots());
                d["bending"] = to_array(s.bending().hessian_slots());
                return d;
            })
        .def("debug_dihedrals",
             [](const ClothSimulator& s) {
                 const auto& d = s.bending().indices();
                 py::array_t<int> a({static_cast<py::ssize_t>(d.size()), static_cast<py::ssize_t>(4)});
                 std::memcpy(a.mutable_data(), d.data(), d.size() * sizeof(chysx::math::Vec4i));
                 py::array_t<float> r(static_cast<py::ssize_t>(d.size()));
                 std::copy(s.bending().rest_angles().begin(), s.bending().rest_angles().end(), r.mutable_data());
                 return py::make_tuple(a, r);
             })
        .def("debug_stage_windows",
             [](const ClothSimulator& s) {
                 const auto& w = s.debug_stage_windows();
                 py::array_t<int> a({static_cast<py::ssize_t>(w.size() / 2), static_cast<py::ssize_t>(2)});
                 if (!w.empty()) std::memcpy(a.mutable_data(), w.data(), w.size() * sizeof(int));
                 return a;
             })
        .def("debug_entries16", &ClothSimulator::debug_entries16)
        .def("debug_edges",
             [](const ClothSimulator& s) {
                 const auto& e = s.buffers().edges;
                 py::array_t<int> a({static_cast<py::ssize_t>(e.size()), static_cast<py::ssize_t>(2)});
                 std::memcpy(a.mutable_data(), e.data(), e.size() * sizeof(chysx::math::Vec2i));
                 return a;
             })
        .def_property_readonly(
            "material", [](ClothSimulator& s) -> ClothMaterial& { return s.material(); },
            py::return_value_policy::reference_internal);

    // ---- batched form (many instances, one fused substep) -------------------
    py::class_<chysx::cloth::ClothBatch>(m, "ClothBatch")
        .def(py::init<>())
        .def(
            "add_instance",
            [](chysx::cloth::ClothBatch& b, ClothSimulator& sim) { return b.add_instance(sim); }, py::arg("simulator"),
            py::keep_alive<1, 2>())
        .def("num_instances", &chysx::cloth::ClothBatch::num_instances)
        .def("num_rows", &chysx::cloth::ClothBatch::num_rows)
        .def("device_bytes", &chysx::cloth::ClothBatch::device_bytes)
        .def("set_lanes_per_row", &chysx::cloth::ClothBatch::set_lanes_per_row, py::arg("lanes"))
        .def("set_grid_blocks", &chysx::cloth::ClothBatch::set_grid_blocks, py::arg("blocks"))
        .def("set_tile_rows", &chysx::cloth::ClothBatch::set_tile_rows, py::arg("rows"))
        .def("set_lds_staging", &chysx::cloth::ClothBatch::set_lds_staging, py::arg("enabled"))
        .def("lds_staging", &chysx::cloth::ClothBatch::lds_staging)
        .def("entries16", &chysx::cloth::ClothBatch::entries16)
        .def("set_graph", &chysx::cloth::ClothBatch::set_graph, py::arg("enabled"),
             "Replay substeps from a hipGraph (needs a non-null cuda_stream in step()).")
        .def("graph", &chysx::cloth::ClothBatch::graph)
        .def("graph_captures", &chysx::cloth::ClothBatch::graph_captures)
        .def("set_trace", &chysx::cloth::ClothBatch::set_trace, py::arg("entries"))
        .def(
            "last_trace",
            [](chysx::cloth::ClothBatch& b, std::uintptr_t stream) { return trace_array(b.last_trace(stream)); },
            py::arg("cuda_stream") = 0)
        .def("tile_rows", &chysx::cloth::ClothBatch::tile_rows)
        .def("set_bf16_blocks", &chysx::cloth::ClothBatch::set_bf16_blocks, py::arg("enabled"))
        .def("finalize", &chysx::cloth::ClothBatch::finalize, py::arg("dt") = 1.0f / 60.0f, py::arg("cuda_stream") = 0,
             py::call_guard<py::gil_scoped_release>())
        .def("remove_instance", &chysx::cloth::ClothBatch::remove_instance, py::arg("simulator"))
        .def("clear", &chysx::cloth::ClothBatch::clear)
        .def("lanes_per_row", &chysx::cloth::ClothBatch::lanes_per_row)
        .def("step", &chysx::cloth::ClothBatch::step, py::arg("dt"), py::arg("cuda_stream") = 0, py::call_guard<py::gil_scoped_release>())
        .def("assemble", &chysx::cloth::ClothBatch::assemble, py::arg("dt"), py::arg("cuda_stream") = 0, py::call_guard<py::gil_scoped_release>())
        .def("solve", &chysx::cloth::ClothBatch::solve, py::arg("cuda_stream") = 0, py::call_guard<py::gil_scoped_release>())
        .def("instance_sizes", &chysx::cloth::ClothBatch::instance_sizes)
        .def("last_pcg_iterations", &chysx::cloth::ClothBatch::last_pcg_iterations, py::arg("cuda_stream") = 0)
        .def("last_relative_residuals", &chysx::cloth::ClothBatch::last_relative_residuals,
             py::arg("cuda_stream") = 0)
        .def("set_barrier_timeout_ticks", &chysx::cloth::ClothBatch::set_barrier_timeout_ticks, py::arg("ticks"))
        .def("set_pcg_form", &chysx::cloth::ClothBatch::set_pcg_form, py::arg("form"),
             "0: one grid barrier per PCG iteration (default), 1: two")
        .def("pcg_form", &chysx::cloth::ClothBatch::pcg_form);

    // ---- standalone block-CSR matrix and PCG solver --------------------------
    // (ChysX/src/sparse/block_csr.h:59-135, ChysX/src/solver/pcg_solver.h:79-167)
    py::class_<chysx::sparse::BlockCSR3>(m, "BlockCSR3")
        .def(py::init<>())
        .def(
            "build_topology",
            [](chysx::sparse::BlockCSR3& A, int n, IArr rows, IArr cols) {
                if (rows.ndim() != 1 || cols.ndim() != 1 || rows.shape(0) != cols.shape(0))
                    throw std::invalid_argument("BlockCSR3.build_topology: rows/cols must be 1-D of equal length");
                A.build_topology(n, rows.data(), cols.data(), static_cast<int>(rows.shape(0)));
            },
            py::arg("num_block_rows"), py::arg("rows"), py::arg("cols"))
        .def(
            "upload",
            [](chysx::sparse::BlockCSR3& A, FArr diag, FArr values) {
                // Addition: host blocks (row-major 3x3) into the device arrays.
                const auto n = static_cast<py::ssize_t>(A.num_block_rows());
                const auto nnz = static_cast<py::ssize_t>(A.num_off_diag_blocks());
                if (diag.size() != 9 * n || values.size() != 9 * nnz)
                    throw std::invalid_argument("BlockCSR3.upload: diag (n, 3, 3) and values (nnz, 3, 3) expected");
                A.diag.host().assign(reinterpret_cast<const chysx::math::Mat3f*>(diag.data()),
                                     reinterpret_cast<const chysx::math::Mat3f*>(diag.data()) + n);
                A.values.host().assign(reinterpret_cast<const chysx::math::Mat3f*>(values.data()),
                                       reinterpret_cast<const chysx::math::Mat3f*>(values.data()) + nnz);
                A.diag.copy_to_device();
                A.values.copy_to_device();
            },
            py::arg("diag"), py::arg("values"))
        .def("num_block_rows", &chysx::sparse::BlockCSR3::num_block_rows)
        .def("num_off_diag_blocks", &chysx::sparse::BlockCSR3::num_off_diag_blocks)
        .def("row_offsets", [](const chysx::sparse::BlockCSR3& A) { return to_array(A.pattern().row_offsets); })
        .def("col_indices", [](const chysx::sparse::BlockCSR3& A) { return to_array(A.pattern().col_indices); });

    py::class_<chysx::solver::PCGSolver>(m, "PCGSolver")
        .def(py::init<>())
        .def(
            "solve",
            [](chysx::solver::PCGSolver& S, const chysx::sparse::BlockCSR3& A, std::uintptr_t b_ptr,
               std::uintptr_t x_ptr, int max_iterations, double tolerance, std::uintptr_t stream, int lanes) {
                const auto n = static_cast<std::size_t>(A.num_block_rows());
                chysx::solver::PCGParams prm;
                prm.max_iterations = max_iterations;
                prm.tolerance = tolerance;
                prm.lanes_per_row = lanes;
                py::gil_scoped_release release;
                return S.solve(A, chysx::DeviceSpan<chysx::math::Vec3f>::from_raw(b_ptr, n),
                               chysx::DeviceSpan<chysx::math::Vec3f>::from_raw(x_ptr, n), prm, stream);
            },
            py::arg("A"), py::arg("b_ptr"), py::arg("x_ptr"), py::arg("max_iterations") = 50,
            py::arg("tolerance") = 0.0, py::arg("cuda_stream") = 0, py::arg("lanes_per_row") = 8,
            "Asynchronous unless tolerance > 0; returns the iteration count (max_iterations without a tolerance).")
        .def("last_residual", &chysx::solver::PCGSolver::last_residual,
             "<r, z> of the last solve's final iteration (float, read on demand).")
        .def("last_iterations", &chysx::solver::PCGSolver::last_iterations)
        .def("last_relative_residual", &chysx::solver::PCGSolver::last_relative_residual);

    // ---- host topology (no device needed) ------------------------------------
    // ---- frame writers (ChysX/src/io/obj_io.h, bgeo_writer.h) -------------
    py::module_ io = m.def_submodule("io", "OBJ and Houdini .bgeo frame writers of the headless exporter");
    auto tri_check = [](const FArr& pos, const IArr& tri, const char* fn) {
        if (pos.ndim() != 2 || pos.shape(1) != 3) throw std::invalid_argument(std::string(fn) + ": positions must be (N, 3)");
        if (tri.ndim() != 2 || tri.shape(1) != 3) throw std::invalid_argument(std::string(fn) + ": triangles must be (T, 3)");
        const int n = static_cast<int>(pos.shape(0));
        const int* t = tri.data();
        for (py::ssize_t k = 0; k < tri.size(); ++k)
            if (t[k] < 0 || t[k] >= n) throw std::invalid_argument(std::string(fn) + ": triangle index out of range");
    };
    io.def(
        "load_obj",
        [](const std::string& path) {
            chysx::io::ObjMesh mesh;
            if (!chysx::io::load_obj(path, mesh)) throw std::runtime_error("load_obj: cannot open " + path);
            py::array_t<float> p({static_cast<py::ssize_t>(mesh.positions.size() / 3), static_cast<py::ssize_t>(3)});
            py::array_t<int> t({static_cast<py::ssize_t>(mesh.triangles.size() / 3), static_cast<py::ssize_t>(3)});
            if (!mesh.positions.empty()) std::memcpy(p.mutable_data(), mesh.positions.data(), mesh.positions.size() * 4);
            if (!mesh.triangles.empty()) std::memcpy(t.mutable_data(), mesh.triangles.data(), mesh.triangles.size() * 4);
            return py::make_tuple(p, t);
        },
        py::arg("path"), "(positions (N, 3) float32, triangles (T, 3) int32, 0-based)");
    io.def(
        "save_obj",
        [tri_check](const std::string& path, FArr positions, IArr triangles) {
            tri_check(positions, triangles, "save_obj");
            if (!chysx::io::save_obj(path, positions.data(), static_cast<int>(positions.shape(0)), triangles.data(),
                                     static_cast<int>(triangles.shape(0))))
                throw std::runtime_error("save_obj: cannot write " + path);
        },
        py::arg("path"), py::arg("positions"), py::arg("triangles"));
    io.def(
        "write_bgeo",
        [tri_check](const std::string& path, FArr positions, IArr triangles, py::object point_colors,
                    py::object face_colors) {
            tri_check(positions, triangles, "write_bgeo");
            FArr pc, fc;
            const float* pcp = nullptr;
            const float* fcp = nullptr;
            if (!point_colors.is_none()) {
                pc = point_colors.cast<FArr>();
                if (pc.ndim() != 2 || pc.shape(0) != positions.shape(0) || pc.shape(1) != 3)
                    throw std::invalid_argument("write_bgeo: point_colors must be (N, 3)");
                pcp = pc.data();
            }
            if (!face_colors.is_none()) {
                fc = face_colors.cast<FArr>();
                if (fc.ndim() != 2 || fc.shape(0) != triangles.shape(0) || fc.shape(1) != 3)
                    throw std::invalid_argument("write_bgeo: face_colors must be (T, 3)");
                fcp = fc.data();
            }
            if (!chysx::io::BgeoWriter::write(path, positions.data(), static_cast<int>(positions.shape(0)),
                                              triangles.data(), static_cast<int>(triangles.shape(0)), pcp, fcp))
                throw std::runtime_error("write_bgeo: cannot write " + path);
        },
        py::arg("path"), py::arg("positions"), py::arg("triangles"), py::arg("point_colors") = py::none(),
        py::arg("face_colors") = py::none());
    io.def(
        "write_bgeo_multi",
        [tri_check](const std::string& path, py::list pieces) {
            std::vector<FArr> keep_p;
            std::vector<IArr> keep_t;
            std::vector<chysx::io::BgeoMeshPiece> ps;
            for (py::handle h : pieces) {
                py::tuple t = h.cast<py::tuple>();
                if (t.size() != 3) throw std::invalid_argument("write_bgeo_multi: pieces are (positions, triangles, rgb)");
                keep_p.push_back(t[0].cast<FArr>());
                keep_t.push_back(t[1].cast<IArr>());
                tri_check(keep_p.back(), keep_t.back(), "write_bgeo_multi");
                auto rgb = t[2].cast<std::vector<float>>();
                if (rgb.size() != 3) throw std::invalid_argument("write_bgeo_multi: colour must have 3 components");
                ps.push_back({keep_p.back().data(), static_cast<int>(keep_p.back().shape(0)), keep_t.back().data(),
                              static_cast<int>(keep_t.back().shape(0)), rgb[0], rgb[1], rgb[2]});
            }
            if (!chysx::io::BgeoWriter::write_multi(path, ps))
                throw std::runtime_error("write_bgeo_multi: cannot write " + path);
        },
        py::arg("path"), py::arg("pieces"));
    io.def(
        "box_mesh",
        [](FArr center, FArr half_ext) {
            if (center.size() != 3 || half_ext.size() != 3) throw std::invalid_argument("box_mesh: (3,) vectors");
            std::vector<float> p;
            std::vector<int> t;
            chysx::io::generate_box_mesh(center.data()[0], center.data()[1], center.data()[2], half_ext.data()[0],
                                         half_ext.data()[1], half_ext.data()[2], p, t);
            py::array_t<float> pa({static_cast<py::ssize_t>(8), static_cast<py::ssize_t>(3)});
            py::array_t<int> ta({static_cast<py::ssize_t>(12), static_cast<py::ssize_t>(3)});
            std::memcpy(pa.mutable_data(), p.data(), p.size() * 4);
            std::memcpy(ta.mutable_data(), t.data(), t.size() * 4);
            return py::make_tuple(pa, ta);
        },
        py::arg("center"), py::arg("half_ext"));

    py::module_ topo = m.def_submodule("topology", "Host topology of the cloth step (bit-exact integer work)");
    topo.def(
        "build_edges",
        [](IArr tris) {
            const auto e = chysx::topology::build_edges(as_tris(tris, "build_edges"), static_cast<int>(tris.shape(0)));
            py::array_t<int> a({static_cast<py::ssize_t>(e.size()), static_cast<py::ssize_t>(2)});
            std::memcpy(a.mutable_data(), e.data(), e.size() * sizeof(chysx::math::Vec2i));
            return a;
        },
        py::arg("triangles"));
    topo.def(
        "detect_dihedrals",
        [](IArr tris) {
            const auto d =
                chysx::topology::detect_dihedrals(as_tris(tris, "detect_dihedrals"), static_cast<int>(tris.shape(0)));
            py::array_t<int> a({static_cast<py::ssize_t>(d.size()), static_cast<py::ssize_t>(4)});
            std::memcpy(a.mutable_data(), d.data(), d.size() * sizeof(chysx::math::Vec4i));
            return a;
        },
        py::arg("triangles"));
    topo.def(
        "build_pattern",
        [](int n, IArr rows, IArr cols) {
            if (rows.ndim() != 1 || cols.ndim() != 1 || rows.shape(0) != cols.shape(0))
                throw std::invalid_argument("build_pattern: rows/cols must be 1-D of equal length");
            const auto p = chysx::topology::build_pattern(n, rows.data(), cols.data(), static_cast<int>(rows.shape(0)));
            return py::make_tuple(to_array(p.row_offsets), to_array(p.col_indices));
        },
        py::arg("n"), py::arg("rows"), py::arg("cols"));
    topo.def(
        "cloth_system",
        [](int n, IArr tris, IArr pins, bool stretch, bool shear, bool bending) {
            // Pattern + slot tables of a cloth with the given element sets, as
            // ClothSimulator::ensure_hessian_topology builds them.
            const auto* t = as_tris(tris, "cloth_system");
            const int nt = static_cast<int>(tris.shape(0));
            std::vector<chysx::math::Vec3i> tv(t, t + nt);
            std::vector<chysx::math::Vec4i> dih;
            if (bending) dih = chysx::topology::detect_dihedrals(t, nt);
            std::vector<int> rows, cols;
            std::vector<chysx::math::Vec3i> none;
            chysx::topology::append_triangle_pairs(stretch ? tv : none, rows, cols);
            chysx::topology::append_triangle_pairs(shear ? tv : none, rows, cols);
            chysx::topology::append_dihedral_pairs(dih, rows, cols);
            const auto p = chysx::topology::build_pattern(n, rows.data(), cols.data(), static_cast<int>(rows.size()));
            std::vector<int> pv(pins.data(), pins.data() + pins.size());
            py::dict d;
            d["row_offsets"] = to_array(p.row_offsets);
            d["col_indices"] = to_array(p.col_indices);
            d["pin_slots"] = to_array(chysx::topology::slot_table_pins(p, pv));
            d["stretch_slots"] = to_array(stretch ? chysx::topology::slot_table3(p, tv) : std::vector<int>{});
            d["shear_slots"] = to_array(shear ? chysx::topology::slot_table3(p, tv) : std::vector<int>{});
            d["bending_slots"] = to_array(chysx::topology::slot_table4(p, dih));
            py::array_t<int> da({static_cast<py::ssize_t>(dih.size()), static_cast<py::ssize_t>(4)});
            std::memcpy(da.mutable_data(), dih.data(), dih.size() * sizeof(chysx::math::Vec4i));
            d["dihedrals"] = da;
            return d;
        },
        py::arg("n"), py::arg("triangles"), py::arg("pins"), py::arg("stretch") = true, py::arg("shear") = true,
        py::arg("bending") = true);
}
