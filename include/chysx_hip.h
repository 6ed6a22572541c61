/*
 * chysx_hip.h — C ABI of the MI355X (gfx950) implementation of ChysX's
 * implicit cloth time-step.
 *
 * The C++ host layer (ClothSimulator, BlockCSR3, PCGSolver, ClothBatch) and
 * any foreign-language binding call the device only through these entry
 * points.  Signatures use plain pointers, sizes and integer stream handles
 * (a hipStream_t cast to uintptr_t, 0 = the default stream); no C++ or
 * framework types cross this boundary.
 *
 * Every entry point returns CHX_OK (0) or an error code; the message of the
 * most recent error on the calling thread is chx_last_error_message().  The
 * code maps onto the reference's exception types:
 *   CHX_ERR_INVALID_ARGUMENT -> std::invalid_argument (Python ValueError)
 *   CHX_ERR_OUT_OF_RANGE     -> std::out_of_range     (Python IndexError)
 *   CHX_ERR_RUNTIME / DEVICE -> std::runtime_error    (Python RuntimeError)
 * (ChysX/src/cloth/cloth_simulator.cu:46-51, ChysX/src/sparse/block_csr.cu:197-201).
 *
 * Memory layout shared with the kernels:
 *   positions / velocities / vectors : AoS float3, 12 bytes per vertex
 *   3x3 blocks passed by the caller   : 9 floats, row-major, 36 bytes
 *     (chx_spmv_bsr3, chx_pcg_solve_bsr3)
 * which is the reference's Vec3f / Mat3f layout (ChysX/src/math/vec.cuh,
 * ChysX/src/math/matrix.cuh:130-133), so caller buffers are used in place.
 * The kernels index arrays with 32-bit byte offsets: every block array (and
 * every per-instance array of a cloth) must stay below 4 GiB, i.e. at most
 * 119,304,647 block rows or off-diagonal blocks per matrix.
 * The matrices a cloth instance assembles for itself (chx_instance diag,
 * values, minv, diag_c, values16) and the preconditioner of
 * chx_pcg_solve_bsr3's workspace are stored in grouped paired planes
 * instead: blocks in groups of CHX_BLOCK_GROUP (64), each group one
 * contiguous run of components (row-major r*3+col) in planes of pairs --
 * (c0,c1), (c2,c3), (c4,c5), (c6,c7) as 64 pairs each, then c8 as 64 words:
 * component c < 8 of block i at [(i / 64) * 576 + (c / 2) * 128 +
 * 2 * (i % 64) + c % 2], c8 at [(i / 64) * 576 + 512 + i % 64]; arrays
 * rounded up to whole groups.  A wavefront reads one pair of 64 consecutive
 * blocks as 512 contiguous bytes (five loads per block).  A cloth instance's
 * diagonal blocks and preconditioner are symmetric and keep three planes of
 * pairs per group, (xx,xy), (xz,yy), (yz,zz): component q at
 * [(i / 64) * 384 + (q / 2) * 128 + 2 * (i % 64) + q % 2].
 */
#ifndef CHYSX_HIP_H_
#define CHYSX_HIP_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    CHX_OK = 0,
    CHX_ERR_INVALID_ARGUMENT = 1,
    CHX_ERR_OUT_OF_RANGE = 2,
    CHX_ERR_RUNTIME = 3,
    CHX_ERR_DEVICE = 4
};

/* Contribution-list entry encoding (gather map of the deterministic
 * assembly), 16 bits: bits 15..14 element kind, 13..12 local row a, 11..10
 * local column b, 9..0 the element's position in its assembly unit's list
 * of elements of that kind (at most CHX_CONTRIB_MAX_LOCAL + 1 elements of a
 * kind per unit). */
#define CHX_CONTRIB_PIN 0u
#define CHX_CONTRIB_TRI 1u
#define CHX_CONTRIB_DIH 2u
#define CHX_CONTRIB_STVK 3u
#define CHX_CONTRIB_MAX_LOCAL 1023u

/* Packed element records (device arrays). Vertex indices are local to the
 * element's cloth instance. */
/* Pins: vertex and stiffness; the targets are a separate AoS float3 array
 * (chx_instance::pin_targets) so that update_pin_targets uploads only them. */
typedef struct chx_pin { int v; float k; } chx_pin;
typedef struct chx_tri { int v0, v1, v2, pad; float dm00, dm01, dm10, dm11; float kA; } chx_tri;
typedef struct chx_dih { int v0, v1, v2, v3; float rest; float k; int pad0, pad1; } chx_dih;
/* StVK membrane triangle (no reference counterpart; ClothMaterial's
 * lame_mu / lame_lambda hooks): rest shape as chx_tri with theta = 0,
 * muA = mu * area * thickness, lamA = lambda * area * thickness. */
typedef struct chx_stvk {
    int v0, v1, v2, pad;
    float dm00, dm01, dm10, dm11;
    float muA, lamA;
    int pad0, pad1;
} chx_stvk;

/* Static contact primitives (ChysX/src/collision/static_contact.h:82-110):
 * a plane dot(n, x) + d = 0 with unit n pointing to the free side, and an
 * oriented box (centre, half extents, unit frame columns ex, ey, ez). */
typedef struct chx_plane { float nx, ny, nz, d; } chx_plane;
typedef struct chx_box {
    float cx, cy, cz;
    float hx, hy, hz;
    float ex[3], ey[3], ez[3];
    float pad;
} chx_box;

typedef struct chx_contact_op {
    const int* pairs;                /* [max_contacts][4] particle indices               */
    const float* weights;            /* [max_contacts][8]: w0..w3, nx, ny, nz, depth     */
    const int* count;                /* device int: active contacts (clamped to          */
                                     /* max_contacts), or NULL for none                  */
    const float* slips;              /* [max_contacts][4] (ux, uy, uz, |u_t|), or NULL   */
    int max_contacts;
    float stiffness;
    float friction_mu;               /* 0: normal block only                             */
    float friction_eps;              /* slip regularisation eps_u [m]                    */
} chx_contact_op;

/* Assembly units.  The assembly kernel processes one unit per workgroup:
 * a row range of one instance -- whole slices of CHX_SLICE_ROWS rows, up to
 * four, or part of one slice -- whose elements' force / Hessian records
 * (every element with a vertex in the range) it computes into LDS, then
 * every block the unit owns: the diagonal blocks of its rows, then the upper
 * value slots of its rows (slice by slice; block index jj, then row), each
 * summing its contribution list, which the kernel stages in LDS with the
 * slots' offsets.  Record floats per element in LDS, and the LDS bytes a
 * unit's records (4 B a float), contribution entries (2 B each) and slot
 * offsets (2 B each) may use together: */
#define CHX_ASM_PIN_REC 4
#define CHX_ASM_TRI_REC 24
#define CHX_ASM_DIH_REC 16
#define CHX_ASM_STVK_REC 32
#define CHX_ASM_LDS_BYTES 53248
typedef struct chx_unit {
    int row_begin, row_end;          /* rows [row_begin, row_end) of the instance             */
    int elem_begin;                  /* first entry in unit_elems: pins, triangles, StVK,     */
    int n_pins, n_tris, n_stvks, n_dihs; /* dihedrals, ascending element index per kind      */
    int n_slots;                     /* blocks the unit owns (diagonal, then upper slots)     */
    int offs_begin;                  /* first of its n_slots + 1 entries in contrib_offsets   */
                                     /* (a multiple of 8)                                     */
    int word_begin;                  /* first of its entries in contribs (a multiple of 8)    */
    int n_words;                     /* its entries, padded with zeros to a multiple of 8     */
    int pad;
} chx_unit;

/* Number of float64 partial sums kept per row tile by the PCG kernel. */
#define CHX_TILE_PARTIALS 16

/* 32-bit words of the grid-barrier / control area `sync` of a batch. */
#define CHX_SYNC_WORDS 3008

/* Default rows per PCG row tile (the unit of workgroup assignment and of the
 * FP64 partial sums; the batch may use any multiple of CHX_STAGE_ROWS),
 * rows per staging step of the PCG's SpMV (the rows whose pattern and source
 * window the kernel loads into LDS at once), and elements / BSR slots per
 * element-pass and gather-pass work item. */
#define CHX_TILE_ROWS 256
#define CHX_STAGE_ROWS 256
#define CHX_ITEM_SIZE 256

/* Rows per slice of the sliced layout of the off-diagonal blocks. */
#define CHX_SLICE_ROWS 64
/* Blocks per group of the grouped block layout. */
#define CHX_BLOCK_GROUP 64

/* One cloth instance: every array is the instance's own device memory, with
 * indices local to the instance, so the off-diagonal pattern is exactly the
 * reference's (ChysX/src/sparse/block_csr.h:59-135).  Slots of the gather
 * map: [0, n_rows) diagonal blocks, [n_rows, n_rows + nnz) off-diagonal.
 *
 * The off-diagonal blocks are stored once per symmetric pair, as the upper
 * block A_ij (j > i) of row i, A_ji being its transpose.  Both the upper
 * blocks and each row's entries (all its off-diagonal columns) are sliced:
 * rows 64s..64s+63 form slice s.  Slice s owns 64 * width_s entry slots
 * [slice_offsets[s], slice_offsets[s+1]) (width_s = its longest row); the
 * j-th off-diagonal column of row 64s+l (CSR order) sits in entry slot
 * slice_offsets[s] + 64 j + l as (col << 6) | jj, where jj is the index of
 * the block among the upper blocks of its owner row (i for col > i, col for
 * col < i).  The upper blocks of slice s own 64 * max(1, upper width_s)
 * value slots from upper_offsets[s]; upper block jj of row 64s+l is value
 * slot upper_offsets[s] + 64 jj + l.  Entry slots past a row's length are
 * padding ((row << 6) | 0, unused).  block_slots maps the reference's CSR
 * index k (row_offsets / col_indices order) to the value slot of an upper
 * block, or to -1 - (the partner's value slot) for a lower one.  Limits:
 * fewer than 2^26 rows and at most 64 upper blocks per row. */
typedef struct chx_instance {
    float* pos;                      /* external AoS float3 positions (x_n in, x_{n+1} out) */
    float* vel;                      /* external AoS float3 velocities                      */
    const float* inv_mass;           /* external inverse masses, or NULL (mass 1 kg)         */
    const int* row_offsets;          /* [n_rows+1]                                           */
    const int* col_indices;          /* [nnz]                                                */
    float* diag;                     /* symmetric: upper triangles in groups of 64 blocks,  */
                                     /* three planes of pairs each; room for 9 * n_rows     */
                                     /* rounded up to 64 blocks                              */
    float* values;                   /* grouped paired planes over the n_upper upper value   */
                                     /* slots                                                */
    const chx_pin* pins;
    const float* pin_targets;        /* [n_pins*3] AoS float3 pin targets                    */
    const chx_tri* tris;             /* stretch elements, then shear elements                */
    const chx_dih* dihs;
    const chx_unit* units;           /* [n_units] assembly units, ascending rows             */
    const int* unit_elems;           /* element indices of the units' lists                  */
    const uint16_t* contrib_offsets; /* per unit: n_slots + 1 offsets of its slots' lists   */
                                     /* relative to its word_begin, padded to 8 entries      */
    const uint16_t* contribs;        /* per unit: its slots' contribution entries            */
    float* rhs;                      /* b [n_rows*3]                                         */
    float* x;                        /* dx: warm start in, solution out [n_rows*3]           */
    float* r;
    float* z;
    float* p0;
    float* p1;
    float* q;
    float* q1;                       /* second q buffer of the single-barrier solve          */
    float* minv;                     /* block-Jacobi preconditioner, symmetric like diag     */
    float* diag_c;                   /* BF16 precision mode, else NULL: the inner SpMV's     */
    uint16_t* values16;              /* matrix, grouped paired-plane FP32 compensated        */
                                     /* diagonal D + sum_j (V_ij - bf16(V_ij)) and BF16      */
                                     /* upper-block copies in the same layout               */
    const chx_stvk* stvks;           /* StVK membrane elements, or NULL                      */
    const int* windows;              /* [ceil(n_rows/CHX_STAGE_ROWS)][2]: smallest and       */
                                     /* largest column (own rows included) of each stage of  */
                                     /* CHX_STAGE_ROWS rows, or NULL (no LDS staging)        */
    const int* slice_offsets;        /* [ceil(n_rows/CHX_SLICE_ROWS)+1] first entry slot     */
    const int* slice_entries;        /* [n_padded] (col << 6) | jj of every entry slot       */
    const uint16_t* slice_entries16; /* the same entries in 16 bits when every one fits,     */
                                     /* else NULL: ((col - row + 2048) << 4) | jj, |col -    */
                                     /* row| < 2048, jj < 16 (row = the slot's row)          */
    const int* upper_offsets;        /* [ceil(n_rows/CHX_SLICE_ROWS)+1] first value slot     */
    const int* block_slots;          /* [nnz] value slot of CSR block k (-1 - slot: lower)   */
    const chx_plane* planes;         /* static contact primitives, or NULL                   */
    const chx_box* boxes;
    int n_rows, nnz, n_pins, n_tris, n_dihs;
    int tile_begin;                  /* this instance's row tiles in the batch tile list     */
    int tile_end;
    int n_stvks;
    int n_padded;                    /* entry slots (>= nnz)                                 */
    int n_upper;                     /* upper value slots (>= nnz / 2)                       */
    int n_units;                     /* assembly units                                       */
    float dt;                        /* substep                                              */
    float damping;                   /* v *= exp(-damping*dt) when > 0                       */
    float gx, gy, gz;                /* gravity                                              */
    int max_iter;                    /* PCG iteration cap                                    */
    int n_planes, n_boxes;           /* static contact: active when contact_stiffness > 0    */
    float contact_thickness;         /* h: penalty on depth = h - signed distance > 0        */
    float contact_stiffness;         /* k [N/m]                                              */
    float contact_friction;          /* Coulomb mu (0: off)                                  */
    float contact_friction_eps;      /* slip regularisation eps_u [m]                        */
    double tol;                      /* stop at ||r|| <= tol*||b||; <= 0: run max_iter       */
} chx_instance;

/* A batch of instances advanced together.  Work lists are device int4
 * arrays: row tiles {instance, row_begin, row_end, 0} (local rows, row_begin
 * a multiple of CHX_STAGE_ROWS, grouped by instance), element items {instance, kind
 * (0 pin, 1 triangle, 2 dihedral, 3 StVK triangle), begin, end} (at most
 * CHX_ITEM_SIZE each; energy only) and unit items {instance, unit, 0, 0}.  A
 * batch with inst == NULL holds one instance in `single` (passed by value in
 * the kernel arguments); its element and unit lists are implied by the
 * counts, and so is its tile list when `tiles` is NULL (tile_rows rows per
 * tile, CHX_TILE_ROWS when tile_rows is 0; n_tiles = ceil(n_rows / that)),
 * which saves the PCG a load of the tile record per tile and phase. */
typedef struct chx_cloth_batch {
    int n_instances;
    int n_tiles;
    int n_elem_items;
    int n_unit_items;
    int lanes_per_row;               /* SpMV lanes per block row: 4, 8 or 16          */
    int grid_blocks;                 /* persistent PCG workgroups (0 = auto)          */
    int bf16_blocks;                 /* 1: inner CG on the BF16 matrix (FP32          */
                                     /* accumulate), outer refinement on the FP32     */
                                     /* true residual when the inner residual meets   */
                                     /* the tolerance                                 */
    int entries16;                   /* 1: every instance has slice_entries16 and the PCG   */
                                     /* reads entries in 16 bits                         */
    int tile_rows;                   /* rows per tile: informational with a tile list, */
                                     /* the implied tiles' size without one (0:       */
                                     /* CHX_TILE_ROWS; a multiple of CHX_STAGE_ROWS)  */
    int no_lds_staging;              /* 1: the PCG SpMV gathers from global memory;   */
                                     /* 0: stages of CHX_STAGE_ROWS rows go through   */
                                     /* LDS (instances with windows)                  */
    chx_instance single;
    const chx_instance* inst;        /* device instance table, or NULL                */
    const int* tiles;
    const int* elem_items;
    const int* unit_items;
    double* partials;                /* [n_tiles*CHX_TILE_PARTIALS]                   */
    int* inst_iters;                 /* [n_instances] PCG iterations of the last solve */
    double* inst_rr;                 /* [n_instances] final ||r||^2                   */
    double* inst_bb;                 /* [n_instances] ||b||^2                         */
    int* inst_done;                  /* [n_instances] scratch                         */
    double* inst_rho;                /* [n_instances] final <r, z>, or NULL           */
    unsigned* sync;                  /* [CHX_SYNC_WORDS] grid barrier + control words */
    unsigned* status_host;           /* host-mapped word set to 1 when a solve aborts */
                                     /* (grid-barrier timeout), or NULL; cleared by   */
                                     /* chx_batch_check                               */
    chx_contact_op contact;          /* contact sidecar of a row-major solve (pairs     */
                                     /* NULL: none); its per-row index below           */
    const int* contact_rows;         /* [n_rows+1] first index entry of each row        */
    const int* contact_entries;      /* 4 c + slot, ascending per row                   */
    unsigned long long* trace;       /* optional solver trace, or NULL: entry e =      */
                                     /* {100 MHz real-time stamp, active instances} at */
                                     /* barrier e (0: after set-up; two-barrier form:  */
                                     /* 2k+1 / 2k+2 after phase A / B of iteration k;  */
                                     /* single-barrier form: k+1 after phase k),       */
                                     /* written by workgroup 0                         */
    int trace_entries;               /* capacity of `trace` in entries (2 words each) */
    unsigned wait_ticks;             /* bound of every grid-barrier wait in 100 MHz   */
                                     /* ticks (0: 1e8, one second); past it the solve */
                                     /* aborts                                        */
    int pcg_form;                    /* FP32 cloth instances: 0 one grid barrier per   */
                                     /* iteration (the two reductions merged), 1 two  */
                                     /* barriers per iteration                        */
    int pad2;
} chx_cloth_batch;

/* ---- errors and runtime ------------------------------------------------- */
const char* chx_last_error_message(void);
/* Number of visible devices; 0 when the HIP runtime finds none. */
int chx_device_count(int* count);
/* The calling thread's current device. */
int chx_get_device(int* device);
/* Compute units, the persistent PCG grid cap, and the runtime's occupancy
 * figure (workgroups per CU) for the PCG kernel on the current device. */
int chx_device_info(int* cu_count, int* max_pcg_blocks, int* occupancy_blocks_per_cu);
/* Co-resident workgroups of one form of the persistent PCG kernel on the
 * current device (the grid cap of a launch with those settings). */
int chx_pcg_grid_limit(int lanes_per_row, int bf16_blocks, int no_lds_staging, int* max_blocks);

/* Device memory: replaces CudaArray<T> (ChysX/src/memory/cuda_array.h:64-284,
 * cuda_array.cu:41-90). */
int chx_malloc(void** ptr, size_t bytes);
int chx_free(void* ptr);
/* Page-locked host memory, also mapped for the device (the pinned host
 * mirror of CudaArray<T>, cuda_array.cu:41-62). */
int chx_host_malloc(void** ptr, size_t bytes);
int chx_host_free(void* ptr);
/* Host-to-device copy queued on `stream` without waiting (`src` must be
 * page-locked and stay untouched until the copy ran); on the null stream it
 * waits, like the reference's cuda_copy_h2d (cuda_array.cu:69-79). */
int chx_memcpy_h2d_async(void* dst, const void* src, size_t bytes, uintptr_t stream);
/* Stream events (ordering host reuse of page-locked staging buffers). */
int chx_event_create(void** event);
int chx_event_destroy(void* event);
int chx_event_record(void* event, uintptr_t stream);
int chx_event_synchronize(void* event);
/* Host-to-device copy ordered on `stream`; returns once `src` may be reused. */
int chx_memcpy_h2d(void* dst, const void* src, size_t bytes, uintptr_t stream);
/* Device-to-host copy; returns after the data is on the host. */
int chx_memcpy_d2h(void* dst, const void* src, size_t bytes, uintptr_t stream);
int chx_memcpy_d2d(void* dst, const void* src, size_t bytes, uintptr_t stream);
/* Zero-fill with a kernel (safe inside stream capture). */
int chx_zero(void* dst, size_t bytes, uintptr_t stream);
int chx_stream_synchronize(uintptr_t stream);
/* Host-side profiler ranges (roctx; shown by rocprofv3 --marker-trace),
 * the role of the reference's NVTX ranges (ChysX/src/profile/nvtx_range.h:
 * 58-100): a nested range per calling thread.  chx_cloth_step itself marks
 * "step::hessian" (the assembly launch) and "step::pcg" (the PCG launch). */
int chx_profile_range_push(const char* name);
int chx_profile_range_pop(void);
int chx_device_synchronize(void);

/* ---- setup kernels -------------------------------------------------------- */
/* Rest shape (Dm^-1 rotated by theta, area) of n triangles from current
 * positions: replaces triangle_rest_shape_kernel
 * (ChysX/src/constraint/cloth/triangle_stretch_constraint.cu:50-110,319-356).
 * tris: device int[3n]; pos: device float3; dm_inv: device float[4n]; area: float[n]. */
int chx_triangle_rest_shape(const int* tris, const float* pos, int n, float cos_theta, float sin_theta,
                            float* dm_inv, float* area, uintptr_t stream);

/* Rest dihedral angles of n dihedrals (v0,v1,v2,v3): replaces
 * bending_rest_angle_kernel (ChysX/src/constraint/cloth/bending_constraint.cu:132-148,253-282). */
int chx_dihedral_rest_angle(const int* dihedrals, const float* pos, int n, float* rest, uintptr_t stream);

/* Area-weighted lumped inverse mass without atomics: each vertex sums
 * density/3 * area over its incident triangles (vertex_tri_offsets[n+1],
 * vertex_tris[] in ascending triangle order).  Replaces
 * scatter_triangle_mass_kernel + mass_to_inv_mass_kernel
 * (ChysX/src/cloth/cloth_simulator.cu:151-190,516-575). */
int chx_mass_area_weighted(const int* tris, const float* pos, const int* vertex_tri_offsets,
                           const int* vertex_tris, int n_vertices, float density_over_3,
                           float* inv_mass, uintptr_t stream);

/* Elastic energy of the batch's elements at the instances' current positions
 * (FP64 sum, synchronous).  Replaces Constraint::compute_energy
 * (pin_constraint.cu:120-147, triangle_stretch_constraint.cu:358-384,
 * bending_constraint.cu:284-308). kind_mask: 1 pins, 2 triangles, 4 dihedrals, 8 StVK triangles. */
int chx_batch_energy(const chx_cloth_batch* batch, int kind_mask, double* energy, uintptr_t stream);

/* ---- sparse operator and solver -------------------------------------------- */
/* y = alpha*A*x + beta*y on a block-CSR matrix with separate diagonal
 * (beta == 0: y is not read).  Replaces sparse::spmv
 * (ChysX/src/sparse/block_csr.h:143-148, block_csr.cu:54-83,210-243). */
int chx_spmv_bsr3(int n_rows, const float* diag, const int* row_offsets, const int* col_indices,
                  const float* values, const float* x, float* y, float alpha, float beta,
                  int lanes_per_row, uintptr_t stream);

/* Bytes of device workspace chx_pcg_solve_bsr3 needs for n_rows block rows. */
int chx_pcg_workspace_size(int n_rows, size_t* bytes);

/* Block-Jacobi PCG on a caller-owned block-CSR system, warm-started from x:
 * replaces PCGSolver::solve (ChysX/src/solver/pcg_solver.h:126-131,
 * pcg_solver.cu:282-318).  tol <= 0 runs exactly max_iterations (reference
 * semantics); otherwise stops when ||r||_2 <= tol*||b||_2 (FP64).  The
 * caller passes `workspace` (device memory of at least
 * chx_pcg_workspace_size bytes, no initialisation needed).  Iteration count
 * and FP64 ||r||^2, ||b||^2 are written to the host outputs (any may be
 * NULL); when all are NULL the call is asynchronous. */
int chx_pcg_solve_bsr3(int n_rows, const float* diag, const int* row_offsets, const int* col_indices,
                       const float* values, const float* b, float* x, int max_iterations, double tol,
                       int lanes_per_row, void* workspace, size_t workspace_bytes, int* iterations,
                       double* rr, double* bb, uintptr_t stream);

/* Largest system chx_pcg_solve_bsr3_scratch accepts (block rows). */
#define CHX_PCG_SCRATCH_MAX_ROWS 65536

/* chx_pcg_solve_bsr3 without a workspace argument, for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  No counterpart in the
 * reference.  The workspace is a device array of the library sized for
 * CHX_PCG_SCRATCH_MAX_ROWS rows (6.0 MiB, held by each device the library
 * runs on until the process ends); the call allocates and frees nothing and
 * rejects n_rows > CHX_PCG_SCRATCH_MAX_ROWS with CHX_ERR_INVALID_ARGUMENT.
 * All calls on one device share that array, which keeps the last call's
 * contents (the solve clears what must start at zero): calls that overlap
 * on one device race (serialise them, e.g. on one stream).  Same arguments
 * and results as chx_pcg_solve_bsr3 otherwise. */
int chx_pcg_solve_bsr3_scratch(int n_rows, const float* diag, const int* row_offsets, const int* col_indices,
                               const float* values, const float* b, float* x, int max_iterations, double tol,
                               int lanes_per_row, int* iterations, double* rr, double* bb, uintptr_t stream);

/* ---- contact SpMV sidecar ----------------------------------------------------
 * A COO operator added to the block-CSR system the PCG iterates against,
 * replacing collision::ContactSpMVOp / bake_contact_diag / apply_contact_spmv
 * (ChysX/src/collision/contact_spmv.h:59-135, contact_spmv.cu:46-268) and the
 * contact argument of PCGSolver::solve (pcg_solver.h:126-131).  Contact c
 * couples particles pairs[c][0..3] with weights w0..w3, normal n and depth:
 * H_c = k n n^T + a_c (I - n n^T), a_c = mu k depth f1(|u_t|) (friction, when
 * mu > 0 and slips are given), C[i][j] = sum_c w_i w_j H_c.  The bake adds the
 * same-slot terms to the diagonal blocks, the SpMV the cross-slot terms.  Each
 * call gathers per row from an inverted contact index built on the device in
 * fixed order: no float atomics, results reproducible run to run. */

/* Bytes of device workspace the contact calls need (inverted index). */
int chx_contact_workspace_size(int n_rows, int max_contacts, size_t* bytes);
/* diag (row-major 3x3 blocks, [n_rows][9]) += alpha * same-slot contact terms. */
int chx_contact_bake_diag(int n_rows, float* diag, const chx_contact_op* op, float alpha, void* workspace,
                          size_t workspace_bytes, uintptr_t stream);
/* y += alpha * cross-slot contact terms * x (AoS float3 [n_rows*3]). */
int chx_contact_spmv(int n_rows, const chx_contact_op* op, const float* x, float* y, float alpha, void* workspace,
                     size_t workspace_bytes, uintptr_t stream);
/* chx_pcg_solve_bsr3 on (A + C_off): the contact cross-slot terms join the
 * residual and every SpMV (the caller bakes the same-slot terms into diag). */
int chx_pcg_solve_bsr3_contact(int n_rows, const float* diag, const int* row_offsets, const int* col_indices,
                               const float* values, const float* b, float* x, int max_iterations, double tol,
                               int lanes_per_row, void* workspace, size_t workspace_bytes,
                               const chx_contact_op* op, void* contact_workspace, size_t contact_workspace_bytes,
                               int* iterations, double* rr, double* bb, uintptr_t stream);

/* ---- the implicit cloth substep -------------------------------------------- */
/* One implicit-Euler substep of every instance of a packed batch: one
 * assembly kernel (per unit: element forces and Hessian records into LDS,
 * then the deterministic gather of every block and right-hand side the unit
 * owns, into the fixed BSR pattern), the persistent block-Jacobi PCG (warm
 * start in batch->x) and finalize into the instances' external pos/vel.
 * Replaces the device work of ClothSimulator::step
 * (ChysX/src/cloth/cloth_simulator.cu:814-1261).  `batch` is a host pointer
 * to a descriptor of device pointers.  Asynchronous on `stream`. */
int chx_cloth_step(const chx_cloth_batch* batch, uintptr_t stream);

/* The two halves of chx_cloth_step, for callers that inspect the assembled
 * system before solving (debug_dump_last_solve) or time the parts. */
int chx_cloth_assemble(const chx_cloth_batch* batch, uintptr_t stream);
int chx_cloth_solve(const chx_cloth_batch* batch, int finalize, uintptr_t stream);

/* Reads the device status word of the batch (non-zero after a grid-barrier
 * timeout inside the persistent solver; it stays set across substeps until
 * read here) and clears it and the host-mapped status word; synchronous. */
int chx_batch_check(const chx_cloth_batch* batch, uintptr_t stream);

/* Statistics of the last chx_pcg_solve_bsr3 call that used `workspace`:
 * iterations, FP64 ||r||^2 and ||b||^2 of the final residual and the
 * preconditioner-weighted residual <r, z> of the final iteration (the value
 * PCGSolver::last_residual returns, pcg_solver.h:133-136,
 * pcg_solver.cu:319-323).  Any output may be NULL; synchronous on `stream`. */
int chx_pcg_last_stats(int n_rows, const void* workspace, size_t workspace_bytes, int* iterations, double* rr,
                       double* bb, double* rho, uintptr_t stream);

/* ---- hipGraph-captured substep ---------------------------------------------
 * No counterpart in the reference's cloth step (its PCGSolver captures only
 * its own iteration loop, pcg_solver.h:146-166).  Captures the device work of
 * chx_cloth_step for `batch` (element pass, gather pass, persistent PCG with
 * finalize: kernels only, no memset or copy nodes) on `stream` (not the null
 * stream) and instantiates it.  The graph keeps the descriptor's values of
 * capture time: per-step instance scalars and pointers read from the device
 * instance table (batch->inst) may change between replays; a change of the
 * descriptor itself (a single inline instance, pointers, sizes, launch shape)
 * needs a new capture. */
int chx_cloth_step_graph(const chx_cloth_batch* batch, uintptr_t stream, void** graph_exec);
/* Replays a captured substep on `stream`. */
int chx_graph_launch(void* graph_exec, uintptr_t stream);
int chx_graph_destroy(void* graph_exec);

#ifdef __cplusplus
}
#endif

#endif /* CHYSX_HIP_H_ */
