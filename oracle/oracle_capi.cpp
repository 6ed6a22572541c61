// ctypes-facing C API of the CPU oracle — TEST INFRASTRUCTURE ONLY.
//
// Loaded by tests/, the smoke check and the CPU-baseline leg of bench.py;
// never by the product library.  All floating-point I/O is float64 (values
// that came from float32 round-trip exactly).  Functions return 0 on success
// and -1 on error; orc_last_error() gives the message.
#include <chrono>
#include <cstdio>
#include <memory>
#include <string>

#include "cloth_oracle.hpp"

using namespace chx_oracle;

namespace {
thread_local std::string g_err;

struct Handle {
    int prec = 32;
    std::unique_ptr<Sim<float>> f;
    std::unique_ptr<Sim<double>> d;
};

template <class F>
int guarded(F&& fn) {
    try {
        fn();
        return 0;
    } catch (const std::exception& e) {
        g_err = e.what();
        return -1;
    }
}

template <class T>
void put3(const std::vector<V3<T>>& v, double* out) {
    for (size_t i = 0; i < v.size(); ++i) {
        out[3 * i] = double(v[i].x);
        out[3 * i + 1] = double(v[i].y);
        out[3 * i + 2] = double(v[i].z);
    }
}
template <class T>
void stvk_eval(const double* x, const double* dinv, double muA, double lamA, int exact, double* E, double* g,
                      double* H) {
    V3<T> v[3];
    for (int a = 0; a < 3; ++a) v[a] = V3<T>(T(x[3 * a]), T(x[3 * a + 1]), T(x[3 * a + 2]));
    M2<T> D;
    for (int q = 0; q < 4; ++q) D.m[q] = T(dinv[q]);
    if (E) *E = double(stvk_energy(v[0], v[1], v[2], D, T(muA), T(lamA)));
    if (g) {
        V3<T> ge[3];
        stvk_gradient(v[0], v[1], v[2], D, T(muA), T(lamA), ge);
        for (int a = 0; a < 3; ++a) { g[3 * a] = double(ge[a].x); g[3 * a + 1] = double(ge[a].y); g[3 * a + 2] = double(ge[a].z); }
    }
    if (H) {
        M3<T> hb[9];
        stvk_hessian(v[0], v[1], v[2], D, T(muA), T(lamA), hb, exact != 0);
        for (int b = 0; b < 9; ++b)
            for (int q = 0; q < 9; ++q) H[9 * b + q] = double(hb[b].m[q]);
    }
}
}  // namespace

#define DISPATCH(h, expr)                      \
    do {                                       \
        if ((h)->prec == 64) { auto& S = *(h)->d; expr; } \
        else { auto& S = *(h)->f; expr; }      \
    } while (0)

extern "C" {

const char* orc_last_error() { return g_err.c_str(); }

int orc_num_threads() {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

void orc_set_num_threads(int n) {
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}

// ---- topology --------------------------------------------------------------
// out_edges must hold 2*3*T ints.
int orc_build_edges(const int* tris, int n_tris, int* out_edges, int* out_count) {
    return guarded([&] {
        auto e = build_edges(tris, n_tris);
        for (size_t i = 0; i < e.size(); ++i) { out_edges[2 * i] = e[i][0]; out_edges[2 * i + 1] = e[i][1]; }
        *out_count = static_cast<int>(e.size());
    });
}

// out must hold 4*3*T ints.
int orc_detect_dihedrals(const int* tris, int n_tris, int* out, int* out_count) {
    return guarded([&] {
        auto d = detect_dihedrals(tris, n_tris);
        for (size_t i = 0; i < d.size(); ++i)
            for (int j = 0; j < 4; ++j) out[4 * i + j] = d[i][j];
        *out_count = static_cast<int>(d.size());
    });
}

// ---- single StVK triangle (pinning tests) ---------------------------------------
// x: 3 vertices (9 doubles), dinv: Dm^-1 row-major (4), energy / gradient (9)
// / Hessian blocks H[a*3+b] row-major (81); exact != 0 skips the projection.
int orc_stvk_element(int precision, const double* x, const double* dinv, double muA, double lamA, int exact,
                     double* E, double* g, double* H) {
    return guarded([&] {
        if (precision == 64) stvk_eval<double>(x, dinv, muA, lamA, exact, E, g, H);
        else stvk_eval<float>(x, dinv, muA, lamA, exact, E, g, H);
    });
}

// ---- static contact primitives (pinning tests) --------------------------------
// box: centre, half extents, ex, ey, ez (15 doubles).  Returns the signed
// distance in *d and the outward normal in n[3].
int orc_box_sdf(const double* p, const double* box, double* d, double* n) {
    return guarded([&] {
        BoxShape<double> b;
        b.c = V3<double>(box[0], box[1], box[2]);
        b.h = V3<double>(box[3], box[4], box[5]);
        b.ex = V3<double>(box[6], box[7], box[8]);
        b.ey = V3<double>(box[9], box[10], box[11]);
        b.ez = V3<double>(box[12], box[13], box[14]);
        V3<double> nn;
        *d = box_signed_dist(V3<double>(p[0], p[1], p[2]), b, nn);
        n[0] = nn.x; n[1] = nn.y; n[2] = nn.z;
    });
}
double orc_ipc_f1_over_x(double u, double eps) { return ipc_f1_over_x(u, eps); }

// ---- contact SpMV sidecar -----------------------------------------------------
// pairs [m][4], weights [m][8], slips [m][4] or null; diag [n][9], x / y [n][3].
int orc_contact_bake(int count, const int* pairs, const double* weights, const double* slips, double k, double mu,
                     double eps, double alpha, double* diag) {
    return guarded([&] {
        ContactOp op{pairs, weights, slips, count, k, mu, eps};
        op.bake(diag, alpha);
    });
}
int orc_contact_apply(int count, const int* pairs, const double* weights, const double* slips, double k, double mu,
                      double eps, double alpha, const double* x, double* y) {
    return guarded([&] {
        ContactOp op{pairs, weights, slips, count, k, mu, eps};
        op.apply(x, y, alpha);
    });
}

// ---- simulator ---------------------------------------------------------------
void* orc_sim_create(int precision) {
    auto* h = new Handle;
    h->prec = (precision == 64) ? 64 : 32;
    if (h->prec == 64) h->d.reset(new Sim<double>); else h->f.reset(new Sim<float>);
    return h;
}
void orc_sim_destroy(void* hv) { delete static_cast<Handle*>(hv); }

int orc_sim_set_state(void* hv, int n, const double* pos, const double* vel, const double* inv_mass) {
    auto* h = static_cast<Handle*>(hv);
    return guarded([&] { DISPATCH(h, S.set_state(n, pos, vel, inv_mass)); });
}
int orc_sim_get_state(void* hv, double* pos, double* vel) {
    auto* h = static_cast<Handle*>(hv);
    return guarded([&] { DISPATCH(h, { put3(S.pos, pos); put3(S.vel, vel); }); });
}
int orc_sim_get_inv_mass(void* hv, double* out) {
    auto* h = static_cast<Handle*>(hv);
    return guarded([&] { DISPATCH(h, { for (int i = 0; i < S.n; ++i) out[i] = double(S.inv_mass[i]); }); });
}
int orc_sim_set_material(void* hv, double damping, double gx, double gy, double gz) {
    auto* h = static_cast<Handle*>(hv);
    return guarded([&] { DISPATCH(h, { using T = decltype(S.damping); S.damping = T(float(damping)); S.gx = T(float(gx)); S.gy = T(float(gy)); S.gz = T(float(gz)); }); });
}
int orc_sim_set_mesh(void* hv, const int* tris, int n_tris) {
    auto* h = static_cast<Handle*>(hv);
    return guarded([&] { DISPATCH(h, S.set_mesh(tris, n_tris)); });
}
int orc_sim_redistribute_mass(void* hv, double rho) {
    auto* h = static_cast<Handle*>(hv);
    return guarded([&] { DISPATCH(h, { using T = decltype(S.damping); S.redistribute_mass(T(float(rho))); }); });
}
int orc_sim_build_stretch(void* hv, double k) {
    auto* h = static_cast<Handle*>(hv);
    return guarded([&] { DISPATCH(h, { using T = decltype(S.damping); S.build_stretch(T(float(k))); }); });
}
int orc_sim_build_shear(void* hv, double k) {
    auto* h = static_cast<Handle*>(hv);
    return guarded([&] { DISPATCH(h, { using T = decltype(S.damping); S.build_shear(T(float(k))); }); });
}
int orc_sim_build_stvk(void* hv, double mu, double lam, double thickness) {
    auto* h = static_cast<Handle*>(hv);
    return guarded([&] { DISPATCH(h, { using T = decltype(S.damping); S.build_stvk(T(float(mu)), T(float(lam)), T(float(thickness))); }); });
}
int orc_sim_build_bending(void* hv, double k) {
    auto* h = static_cast<Handle*>(hv);
    return guarded([&] { DISPATCH(h, { using T = decltype(S.damping); S.build_bending(T(float(k))); }); });
}
int orc_sim_set_pins(void* hv, const int* idx, const double* targets, int n, double k) {
    auto* h = static_cast<Handle*>(hv);
    return guarded([&] { DISPATCH(h, { using T = decltype(S.damping); S.set_pins(idx, targets, n, T(float(k))); }); });
}
int orc_sim_update_pin_targets(void* hv, const double* targets, int n) {
    auto* h = static_cast<Handle*>(hv);
    return guarded([&] { DISPATCH(h, { S.update_pin_targets(targets, n); }); });
}
int orc_sim_add_plane(void* hv, const double* nrm, double d) {
    auto* h = static_cast<Handle*>(hv);
    return guarded([&] {
        DISPATCH(h, {
            using T = decltype(S.damping);
            PlaneShape<T> pl;
            pl.n = V3<T>(T(float(nrm[0])), T(float(nrm[1])), T(float(nrm[2])));
            pl.d = T(float(d));
            S.contact.planes.push_back(pl);
        });
    });
}
int orc_sim_add_box(void* hv, const double* box) {
    auto* h = static_cast<Handle*>(hv);
    return guarded([&] {
        DISPATCH(h, {
            using T = decltype(S.damping);
            auto v = [&](int o) { return V3<T>(T(float(box[o])), T(float(box[o + 1])), T(float(box[o + 2]))); };
            BoxShape<T> b;
            b.c = v(0); b.h = v(3); b.ex = v(6); b.ey = v(9); b.ez = v(12);
            S.contact.boxes.push_back(b);
        });
    });
}
int orc_sim_set_contact(void* hv, double thickness, double stiffness, double friction, double eps) {
    auto* h = static_cast<Handle*>(hv);
    return guarded([&] {
        DISPATCH(h, {
            using T = decltype(S.damping);
            S.contact.thickness = T(float(thickness));
            S.contact.stiffness = T(float(stiffness));
            S.contact.friction = T(float(friction));
            S.contact.eps = T(float(eps));
        });
    });
}
int orc_sim_set_pcg(void* hv, int max_iter, double tol) {
    auto* h = static_cast<Handle*>(hv);
    return guarded([&] { DISPATCH(h, { S.max_iter = max_iter > 0 ? max_iter : 1; S.tol = tol; }); });
}
// Returns the PCG iteration count, or -1 on error.
int orc_sim_step(void* hv, double dt) {
    auto* h = static_cast<Handle*>(hv);
    int it = -1;
    if (guarded([&] { DISPATCH(h, { using T = decltype(S.damping); it = S.step(T(float(dt))); }); }) != 0) return -1;
    return it;
}
// Runs `steps` substeps and reports wall seconds (CPU-baseline timing).
int orc_sim_run(void* hv, double dt, int steps, double* seconds, long long* total_iters) {
    auto* h = static_cast<Handle*>(hv);
    return guarded([&] {
        long long tot = 0;
        const auto t0 = std::chrono::steady_clock::now();
        for (int s = 0; s < steps; ++s) DISPATCH(h, { using T = decltype(S.damping); tot += S.step(T(float(dt))); });
        const auto t1 = std::chrono::steady_clock::now();
        *seconds = std::chrono::duration<double>(t1 - t0).count();
        *total_iters = tot;
    });
}
int orc_sim_last_stats(void* hv, int* iters, double* rr, double* bb) {
    auto* h = static_cast<Handle*>(hv);
    return guarded([&] { DISPATCH(h, { *iters = S.last_iters; *rr = S.last_rr; *bb = S.last_bb; }); });
}
int orc_sim_sizes(void* hv, int* n, int* nnz, int* n_edges, int* n_dih, int* n_tris) {
    auto* h = static_cast<Handle*>(hv);
    return guarded([&] { DISPATCH(h, {
        if (S.topology_dirty || S.topo_n != S.n) S.ensure_topology();
        *n = S.n; *nnz = static_cast<int>(S.pat.col_indices.size());
        *n_edges = static_cast<int>(S.edges.size()); *n_dih = static_cast<int>(S.dih.size());
        *n_tris = static_cast<int>(S.tris.size()); }); });
}
int orc_sim_get_pattern(void* hv, int* row_offsets, int* cols) {
    auto* h = static_cast<Handle*>(hv);
    return guarded([&] { DISPATCH(h, {
        if (S.topology_dirty || S.topo_n != S.n) S.ensure_topology();
        std::copy(S.pat.row_offsets.begin(), S.pat.row_offsets.end(), row_offsets);
        std::copy(S.pat.col_indices.begin(), S.pat.col_indices.end(), cols); }); });
}
// kind: 0 pins (P), 1 stretch (9T), 2 shear (9T), 3 bending (16D)
int orc_sim_get_slots(void* hv, int kind, int* out) {
    auto* h = static_cast<Handle*>(hv);
    return guarded([&] { DISPATCH(h, {
        if (S.topology_dirty || S.topo_n != S.n) S.ensure_topology();
        const std::vector<int>* v = kind == 0 ? &S.pin_slots : kind == 1 ? &S.stretch.slots : kind == 2 ? &S.shear.slots
                                  : kind == 4 ? &S.stvk.slots : &S.dih_slots;
        std::copy(v->begin(), v->end(), out); }); });
}
int orc_sim_get_dihedrals(void* hv, int* out, double* rest) {
    auto* h = static_cast<Handle*>(hv);
    return guarded([&] { DISPATCH(h, {
        for (size_t e = 0; e < S.dih.size(); ++e) { for (int j = 0; j < 4; ++j) out[4 * e + j] = S.dih[e][j]; if (rest) rest[e] = double(S.rest_angle[e]); } }); });
}
int orc_sim_get_edges(void* hv, int* out) {
    auto* h = static_cast<Handle*>(hv);
    return guarded([&] { DISPATCH(h, { for (size_t e = 0; e < S.edges.size(); ++e) { out[2 * e] = S.edges[e][0]; out[2 * e + 1] = S.edges[e][1]; } }); });
}
// kind 1 stretch / 2 shear: Dm_inv (4T) and area (T)
int orc_sim_get_rest_shape(void* hv, int kind, double* dm_inv, double* area) {
    auto* h = static_cast<Handle*>(hv);
    return guarded([&] { DISPATCH(h, {
        const auto& s = (kind == 1) ? S.stretch : S.shear;
        for (size_t t = 0; t < s.idx.size(); ++t) { for (int q = 0; q < 4; ++q) dm_inv[4 * t + q] = double(s.Dm_inv[t].m[q]); area[t] = double(s.area[t]); } }); });
}
// Last solved system: diag (9N), values (9 nnz), rhs (3N), dx (3N).
int orc_sim_get_system(void* hv, double* diag, double* values, double* rhs, double* dx) {
    auto* h = static_cast<Handle*>(hv);
    return guarded([&] { DISPATCH(h, {
        for (size_t i = 0; i < S.diag.size(); ++i) for (int q = 0; q < 9; ++q) diag[9 * i + q] = double(S.diag[i].m[q]);
        for (size_t i = 0; i < S.values.size(); ++i) for (int q = 0; q < 9; ++q) values[9 * i + q] = double(S.values[i].m[q]);
        put3(S.rhs, rhs);
        put3(S.dx, dx); }); });
}
int orc_sim_get_dx(void* hv, double* dx) {
    auto* h = static_cast<Handle*>(hv);
    return guarded([&] { DISPATCH(h, {
        using T = decltype(S.damping);
        for (int i = 0; i < S.n; ++i) {
            const V3<T> d = (static_cast<int>(S.dx.size()) == S.n) ? S.dx[i] : V3<T>();
            dx[3 * i] = double(d.x); dx[3 * i + 1] = double(d.y); dx[3 * i + 2] = double(d.z);
        } }); });
}
int orc_sim_set_dx(void* hv, const double* dx) {
    auto* h = static_cast<Handle*>(hv);
    return guarded([&] { DISPATCH(h, {
        using T = decltype(S.damping);
        S.dx.resize(S.n);
        for (int i = 0; i < S.n; ++i) S.dx[i] = V3<T>(T(dx[3 * i]), T(dx[3 * i + 1]), T(dx[3 * i + 2])); }); });
}
// Energy and gradient of E at the current positions (FD checks).
int orc_sim_energy(void* hv, double* out) {
    auto* h = static_cast<Handle*>(hv);
    return guarded([&] { DISPATCH(h, { *out = S.energy(S.pos); }); });
}
int orc_sim_gradient(void* hv, double* out) {
    auto* h = static_cast<Handle*>(hv);
    return guarded([&] { DISPATCH(h, {
        using T = decltype(S.damping);
        std::vector<V3<T>> g; S.gradient(S.pos, g); put3(g, out); }); });
}
// Hessian of E at current positions (no inertia) into diag/values.
int orc_sim_hessian(void* hv, double* diag, double* values) {
    auto* h = static_cast<Handle*>(hv);
    return guarded([&] { DISPATCH(h, {
        if (S.topology_dirty || S.topo_n != S.n) S.ensure_topology();
        S.hessian(S.pos);
        for (size_t i = 0; i < S.diag.size(); ++i) for (int q = 0; q < 9; ++q) diag[9 * i + q] = double(S.diag[i].m[q]);
        for (size_t i = 0; i < S.values.size(); ++i) for (int q = 0; q < 9; ++q) values[9 * i + q] = double(S.values[i].m[q]); }); });
}

// Standalone block-Jacobi PCG on a caller-provided BSR system (warm start in x).
// The standalone PCG on (A + C_off) when `contact` is given (the caller
// baked the same-slot terms into diag).
static int pcg_solve(int precision, int n, const int* row_offsets, const int* cols, const double* diag,
                     const double* values, const double* b, double* x, int max_iter, double tol, int* iters,
                     double* rr, double* bb, const ContactOp* contact, double* rho = nullptr) {
    auto run = [&](auto tag) {
        using T = decltype(tag);
        Sim<T> S;
        S.sidecar = contact;
        S.n = n;
        S.pat.n = n;
        S.pat.row_offsets.assign(row_offsets, row_offsets + n + 1);
        const int nnz = row_offsets[n];
        S.pat.col_indices.assign(cols, cols + nnz);
        S.diag.resize(n);
        S.values.resize(nnz);
        for (int i = 0; i < n; ++i) for (int q = 0; q < 9; ++q) S.diag[i].m[q] = T(diag[9 * i + q]);
        for (int i = 0; i < nnz; ++i) for (int q = 0; q < 9; ++q) S.values[i].m[q] = T(values[9 * i + q]);
        std::vector<V3<T>> bv(n), xv(n);
        for (int i = 0; i < n; ++i) {
            bv[i] = V3<T>(T(b[3 * i]), T(b[3 * i + 1]), T(b[3 * i + 2]));
            xv[i] = V3<T>(T(x[3 * i]), T(x[3 * i + 1]), T(x[3 * i + 2]));
        }
        S.max_iter = max_iter;
        S.tol = tol;
        *iters = S.pcg(bv, xv);
        *rr = S.last_rr;
        *bb = S.last_bb;
        if (rho) *rho = S.last_rho;
        put3(xv, x);
    };
    return guarded([&] { if (precision == 64) run(double{}); else run(float{}); });
}

int orc_pcg_solve(int precision, int n, const int* row_offsets, const int* cols, const double* diag,
                  const double* values, const double* b, double* x, int max_iter, double tol, int* iters,
                  double* rr, double* bb) {
    return pcg_solve(precision, n, row_offsets, cols, diag, values, b, x, max_iter, tol, iters, rr, bb, nullptr);
}

// orc_pcg_solve plus the final <r, z>.
int orc_pcg_solve_rho(int precision, int n, const int* row_offsets, const int* cols, const double* diag,
                      const double* values, const double* b, double* x, int max_iter, double tol, int* iters,
                      double* rr, double* bb, double* rho) {
    return pcg_solve(precision, n, row_offsets, cols, diag, values, b, x, max_iter, tol, iters, rr, bb, nullptr, rho);
}

int orc_pcg_solve_contact(int precision, int n, const int* row_offsets, const int* cols, const double* diag,
                          const double* values, const double* b, double* x, int max_iter, double tol, int count,
                          const int* pairs, const double* weights, const double* slips, double k, double mu,
                          double eps, int* iters, double* rr, double* bb) {
    const ContactOp op{pairs, weights, slips, count, k, mu, eps};
    return pcg_solve(precision, n, row_offsets, cols, diag, values, b, x, max_iter, tol, iters, rr, bb, &op);
}

// y = A x with the oracle's SpMV.
int orc_spmv(int precision, int n, const int* row_offsets, const int* cols, const double* diag,
             const double* values, const double* x, double* y) {
    auto run = [&](auto tag) {
        using T = decltype(tag);
        Sim<T> S;
        S.n = n;
        S.pat.n = n;
        S.pat.row_offsets.assign(row_offsets, row_offsets + n + 1);
        const int nnz = row_offsets[n];
        S.pat.col_indices.assign(cols, cols + nnz);
        S.diag.resize(n);
        S.values.resize(nnz);
        for (int i = 0; i < n; ++i) for (int q = 0; q < 9; ++q) S.diag[i].m[q] = T(diag[9 * i + q]);
        for (int i = 0; i < nnz; ++i) for (int q = 0; q < 9; ++q) S.values[i].m[q] = T(values[9 * i + q]);
        std::vector<V3<T>> xv(n), yv;
        for (int i = 0; i < n; ++i) xv[i] = V3<T>(T(x[3 * i]), T(x[3 * i + 1]), T(x[3 * i + 2]));
        S.spmv(xv, yv);
        put3(yv, y);
    };
    return guarded([&] { if (precision == 64) run(double{}); else run(float{}); });
}

}  // extern "C"
