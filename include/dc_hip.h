/* dc_hip.h — C ABI of the MI355X dual-contouring backend (libdc_hip.so).
 *
 * This is the drop-in boundary for the reference's dense GPU backend. Each entry point names the
 * reference interface it replaces (file:line in the Rizzlord/Phanty-SDF-Cuda tree). Plain
 * pointers and sizes only; every call returns a dc_hip_status and never exits the process.
 *
 * Semantics follow the reference CPU backend (CpuDualContouringBackend, src/cpp/dc_backend.cpp:
 * 51-125): active cells are the cells with a strictly negative and a strictly positive corner;
 * vertex ids are their rank in (k, j, i) order; quads are emitted per axis in (i, j, k) order
 * with k fastest, oriented by normalized central differences; vertices are the truncated-SVD
 * QEF solution about the mass point with 4-cell averaged trilinear hermite normals.
 */
#ifndef DC_HIP_H
#define DC_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum dc_hip_status {
  DC_HIP_OK = 0,
  DC_HIP_ERR_NO_DEVICE = 1,      /* no HIP device, or the device index is out of range */
  DC_HIP_ERR_INVALID_ARG = 2,    /* NULL pointer, unknown option or precision */
  DC_HIP_ERR_INVALID_GRID = 3,   /* nx, ny or nz < 2, or sizes beyond the 32-bit index limits */
  DC_HIP_ERR_NONFINITE = 4,      /* a sample is NaN or +-Inf */
  DC_HIP_ERR_OUT_OF_MEMORY = 5,  /* device allocation failed */
  DC_HIP_ERR_HIP = 6,            /* a HIP runtime call failed (see dc_hip_last_error) */
  DC_HIP_ERR_NO_FORWARD = 7,     /* backward or result query without a successful forward */
  DC_HIP_ERR_INTERNAL = 8        /* a device-side consistency check failed */
} dc_hip_status;

typedef enum dc_hip_precision {
  /* FP64 hermite data and solve, the CPU backend's arithmetic; FP64 vertices available. */
  DC_HIP_PLAIN_FP64 = 0,
  /* FP32 assembly in mass-point-centred coordinates with compensated sums; cells whose condition
   * estimate or truncation margin is poor are solved again in FP64. */
  DC_HIP_FP32_FALLBACK = 1
} dc_hip_precision;

/* Replaces DenseSdfGrid (include/dc_backend.h:13-21) and DenseSdfGridDevice
 * (include/dc_cuda.h:24-29): values[i + nx*(j + ny*k)], sample (i,j,k) at
 * (ox + i*vx, oy + j*vy, oz + k*vz) evaluated in FP32. */
typedef struct dc_hip_grid {
  int32_t nx, ny, nz;
  float vx, vy, vz;
  float ox, oy, oz;
  const float* values;
} dc_hip_grid;

/* Mirrors the DualContouringStats fields the dense backends fill (include/dc_backend.h:44-78,
 * src/cuda/dc_cuda.cu:828-844), plus the per-pass timings of this backend. Times are device
 * (hipEvent) milliseconds except total_ms, which is host wall time of the call. */
typedef struct dc_hip_stats {
  int32_t nx, ny, nz;
  int64_t total_cells;
  int64_t active_cells;
  float upload_ms;
  float marking_ms;        /* marking pass: active plane + face planes (mark + transposes) */
  float compaction_ms;     /* active-cell scan */
  float face_prefix_sum_ms;/* pencil scan */
  float qef_ms;            /* vertex solve */
  float face_fill_ms;      /* quad emission */
  float face_emission_ms;  /* face_prefix_sum_ms + face_fill_ms */
  float download_ms;
  float total_ms;
  float backward_ms;
  int64_t vertex_count;
  int64_t face_count;
  int64_t qef_fallback_count; /* FP32 mode: cells solved again in FP64 */
  int64_t invalid_count;
  float normals_ms;        /* averaged hermite normals, once per face */
  float mark_kernel_ms;    /* the sample-reading kernel of the marking pass alone */
  int32_t graph_mode;      /* forward launch sequence: 0 eager, 1 captured now, 2 replayed */
} dc_hip_stats;

typedef struct dc_hip_context dc_hip_context;

/* Creates a context on HIP device `device` with its own stream. */
int dc_hip_create(int device, dc_hip_context** out);
void dc_hip_destroy(dc_hip_context* ctx);
/* Runs all later work on `stream` (a hipStream_t). NULL, the default, is the null stream: each
 * call runs there and returns when its work is complete, except dc_hip_backward with the
 * "timing" option off, which returns once its work is enqueued (see there). */
int dc_hip_set_stream(dc_hip_context* ctx, void* stream);
/* Options: "mark_rows" (the marking sign pass's wave tile: 4, 8, 16 or 32 sample rows of 64
 * samples; 1: whole rows, 8 samples per lane, where every grid's nx is a multiple of 8 and its
 * values are 16-byte aligned, else 16; 0, the default: whole rows where 16-row tiles would give
 * fewer than 8 waves per CU, else 16), "mark_waves" (1, 2, 4, default 4: waves per marking
 * block, for tiles of 4-32 rows), "gather_rows" (1, 2, 4, 8:
 * sample rows per wave of the adjoint's gradient gather; applies from the next forward),
 * "gather_blocks" (0, the default: one wave per row group of the gather; N > 0: at most N blocks
 * of 4 waves, each looping over row groups; the same gradient bit for bit), "timing" (0/1, default 1: per-pass
 * event timings in the stats), "graph" (0/1, default 1: with timing off, a forward plan that runs
 * twice in a row is captured as a hipGraph and replayed while its grids, sizes, pointers,
 * precision, options and stream stay the same), "qef_blocks" (persistent blocks of the solve),
 * "adjoint_lanes", "adjoint_face_lanes" (0, 1, 4, 16: lanes per active cell in the adjoint's
 * first and second per-cell kernels; 0, the default, picks by mesh size: several lanes where one
 * thread per cell would leave the device idle), "adjoint_fallback_lanes" (0, 1, 16: the same for
 * FP32 mode's FP64 cells; 0, the default: 16 while they are few),
 * "fallback_state" (FP32 mode: up to how many FP64-solved cells keep their solve states for the
 * adjoint; with more, the adjoint solves them all again; -1, the default: a quarter of the
 * active-cell capacity). */
int dc_hip_set_option(dc_hip_context* ctx, const char* key, int64_t value);

/* Replaces CudaDualContouringBackend::extract (src/cuda/dc_cuda.cu:614-649): host values are
 * uploaded, extracted, and the mesh stays in the context (copy it with dc_hip_copy_result). */
int dc_hip_extract(dc_hip_context* ctx, const dc_hip_grid* grid, int precision,
                   dc_hip_stats* stats);
/* Replaces CudaDualContouringBackend::extract_device (src/cuda/dc_cuda.cu:651-847): values is a
 * device pointer, read on the context's stream; the call returns when the mesh is complete. */
int dc_hip_extract_device(dc_hip_context* ctx, const dc_hip_grid* grid, int precision,
                          dc_hip_stats* stats);
/* No counterpart in the reference: an extract_device that allocates nothing, for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc). values is a device pointer; the mesh goes straight into the caller's device
 * buffers: vertices (3 * vertex_capacity floats), vertices_f64 (optional, plain FP64 mode only,
 * 3 * vertex_capacity doubles) and faces (4 * face_capacity int32). Every intermediate buffer
 * lives in the library's static device scratch (64 MiB per device, held until the process
 * ends): a grid whose marking planes plus capacity-sized lists exceed it (about 256^3 with
 * meshes of a few hundred thousand quads fit) is rejected with DC_HIP_ERR_INVALID_GRID before
 * anything runs. *n_vertices and *n_faces receive the mesh's counts; when they exceed the
 * capacities the call returns DC_HIP_ERR_INVALID_ARG with the counts set (nothing past the
 * capacities is written). Runs on the context's stream without per-pass timing and returns when
 * the mesh is complete. Calls on one device share the scratch: calls that overlap on one device
 * race. The context's last result (dc_hip_result_*, dc_hip_backward) is cleared. */
int dc_hip_extract_device_noalloc(dc_hip_context* ctx, const dc_hip_grid* grid, int precision,
                                  float* vertices, double* vertices_f64, int64_t vertex_capacity,
                                  int32_t* faces, int64_t face_capacity, int64_t* n_vertices,
                                  int64_t* n_faces);
/* Counts of the last successful extraction (single grid or whole batch). */
int dc_hip_result_counts(dc_hip_context* ctx, int64_t* n_vertices, int64_t* n_faces);
/* Copies the last result (any pointer may be NULL; host or device memory). vertices:
 * 3*n_vertices floats; vertices_f64: 3*n_vertices doubles (plain FP64 mode only); faces:
 * 4*n_faces int32. Returns when the copies are complete. */
int dc_hip_copy_result(dc_hip_context* ctx, float* vertices, double* vertices_f64,
                       int32_t* faces);
/* The last result's quads split into triangles by the reference's rule, triangulate_v1
 * (src/cpp/contouring.cpp:88-113, orientation :40-86): quad (a,b,c,d) -> triangles 2f = (a,b,c)
 * and 2f+1 = (a,c,d), each with its last two vertices swapped when its normal points against the
 * quad normal (b-a)x(d-a) (or (c-a)x(d-a) when that is degenerate), using the FP64 vertices in
 * plain mode and the FP32 ones otherwise. triangles: 6*n_faces int32, host or device memory;
 * vertex ids are grid-local as in the quads. Returns when the copy is complete. */
int dc_hip_triangulate(dc_hip_context* ctx, int32_t* triangles);
/* Device pointers of the last result, owned by the context and valid until its next call. */
int dc_hip_result_device(dc_hip_context* ctx, const float** vertices,
                         const double** vertices_f64, const int32_t** faces);

/* Batched extraction of n_grids device-resident grids of any sizes in one pipeline (no
 * counterpart in the reference, which extracts one grid per call). Results are concatenated
 * grid-major; per-grid vertex and face offsets (n_grids + 1 entries each, may be NULL) give the
 * ranges; face vertex ids are local to their grid. */
int dc_hip_extract_batch(dc_hip_context* ctx, const dc_hip_grid* grids, int32_t n_grids,
                         int precision, int64_t* vertex_offsets, int64_t* face_offsets,
                         dc_hip_stats* stats);

/* Adjoint of the last extraction at fixed topology (no counterpart in the reference):
 * grad_values[N_s] (device, overwritten) = dL/dvalues for grad_vertices[3*n_vertices]
 * (device) = dL/dvertices. The adjoint reads the last forward's samples again: they must be alive
 * and unchanged, and grad_values must not overlap them (DC_HIP_ERR_INVALID_ARG otherwise). Deterministic: per-cell stencil contributions are gathered in a
 * fixed order, no floating-point atomics. With the "timing" option on (the default) it returns
 * when the gradient is complete; with it off it returns once the work is enqueued on the
 * context's stream, the null stream included: later work on that stream sees the gradient, and
 * a host read must synchronise the stream first. */
int dc_hip_backward(dc_hip_context* ctx, const float* grad_vertices, float* grad_values,
                    dc_hip_stats* stats);

/* Tile-shape sweep of the marking pass (no counterpart in the reference, whose marking kernel has
 * one fixed launch shape, src/cuda/dc_cuda.cu:241-281 launched at :687-692). Every
 * (mark_rows, mark_waves) in {4, 8, 16, 32} x {1, 2, 4} (the whole-row form is not a candidate) is
 * timed on these device-resident grids
 * with the per-pass events (best of `repeats` runs after one warm-up run); the fastest is kept on
 * the context and returned. table_ms (12 floats, rows-major; may be NULL) receives every
 * candidate's marking time (sign and cell passes). The context's "timing" option is restored afterwards; the last
 * extraction's result stays in the context like any other extraction's. */
int dc_hip_autotune(dc_hip_context* ctx, const dc_hip_grid* grids, int32_t n_grids, int precision,
                    int32_t repeats, int32_t* best_rows, int32_t* best_waves, float* table_ms);

/* Counts and times of one mesh post-processing call (dc_hip_postprocess). */
typedef struct dc_hip_post_stats {
  int64_t welded_vertices;  /* after the first weld */
  int64_t welded_faces;     /* quads left by the first weld (fewer than 3 distinct vertices: dropped) */
  int64_t components;       /* edge-connected face components (remove_floaters; else 0) */
  int64_t kept_faces;       /* faces of the largest component (all welded faces without it) */
  int64_t boundary_edges;   /* directed boundary edges given to the hole walk (close_holes) */
  int64_t loops_filled;     /* boundary loops closed */
  int64_t added_vertices;   /* vertices made by the hole fill */
  int64_t added_faces;      /* faces made by the hole fill */
  int64_t vertex_count;     /* result */
  int64_t face_count;
  float walk_ms;            /* host time of the boundary-loop walk and fill */
  float total_ms;           /* host wall time of the call */
} dc_hip_post_stats;

/* Replaces postprocess_mesh (src/cpp/dc_backend.cpp:217-558, declared include/dc_backend.h:28,
 * bound at src/cpp/bindings.cpp:145-146) for a quad mesh: vertices welded on a quantised grid of
 * 1e-5 x the bounding-box diagonal (quads left with fewer than 3 distinct vertices dropped), with
 * remove_floaters only the largest edge-connected face component kept (ties: the one holding the
 * lowest face), with close_holes every boundary loop of 3..4096 edges closed by concentric rings
 * and a centre fan (triangles as quads with a repeated last vertex), vertices renumbered by first
 * use and, with close_holes, welded again. Output equals the reference's for the same input.
 * vertices (3 * n_vertices floats) and faces (4 * n_faces int32, indices < n_vertices) may be host
 * or device memory; both NULL with zero counts: the context's last extraction result (single
 * grid). With neither flag set, or an empty mesh, the result is the input. The result stays in the
 * context beside the last extraction (which stays valid for dc_hip_backward); *out_n_vertices and
 * *out_n_faces (may be NULL) receive its counts; stats may be NULL. Returns when complete. */
int dc_hip_postprocess(dc_hip_context* ctx, const float* vertices, int64_t n_vertices,
                       const int32_t* faces, int64_t n_faces, int close_holes,
                       int remove_floaters, int64_t* out_n_vertices, int64_t* out_n_faces,
                       dc_hip_post_stats* stats);
/* Copies the last post-processing result (host or device memory; either pointer may be NULL):
 * 3 * vertex_count floats, 4 * face_count int32. */
int dc_hip_postprocess_copy(dc_hip_context* ctx, float* vertices, int32_t* faces);
/* Device pointers of the last post-processing result, valid until the next dc_hip_postprocess. */
int dc_hip_postprocess_device(dc_hip_context* ctx, const float** vertices, const int32_t** faces);

/* Counts and times of one mesh-to-SDF call (dc_hip_mesh_to_sdf). */
typedef struct dc_hip_m2s_stats {
  int64_t triangles;    /* non-degenerate triangles (the others have no normal and are skipped) */
  int64_t bin_entries;  /* (bin, triangle) pairs: each triangle in the bins its grown box covers */
  int64_t bins;         /* max(1, nx/8) * max(1, ny/8) * max(1, nz/8) */
  int64_t active_bins;  /* bins with a triangle in their 3x3x3 neighbourhood (evaluated) */
  int32_t flood_passes; /* launches of the exterior flood (the last one changed nothing) */
  int32_t voxel_res;    /* the voxelize option's clamped resolution, else 0 */
  float eval_ms;        /* device time of the distance evaluation (far fill included) */
  float total_ms;       /* host wall time of the call */
} dc_hip_m2s_stats;

/* Replaces compute_mesh_sdf_device_cuda / compute_mesh_sdf_cuda (include/mesh_to_sdf.h:6-26,
 * src/cuda/mesh_to_sdf_cuda.cu:740-1002; bound at src/cpp/bindings.cpp:270-285 and called by
 * dc_cli's mesh input, src/cpp/dc_cli.cpp:120-203): the signed distance of the triangle mesh
 * (vertices: 3 * num_vertices floats; faces: 3 * num_faces int32 vertex indices; host or device
 * memory) sampled at (ox + i*vx, oy + j*vy, oz + k*vz) into values[i + nx*(j + ny*k)] (host or
 * device memory, nx*ny*nz floats). Samples within 2.5 bins (bin = (n*v)/max(1, n/8) per axis) of
 * a triangle of their 3x3x3 bin neighbourhood get the signed distance, its sign from the
 * angle-weighted normals of the nearest triangles; the others get +10 (exterior) or -10. Without
 * voxelize_first, far samples joined to the grid's border or to a positive sample through far
 * samples are exterior, and four relaxation passes follow (|f| > 1.5 min(v) takes
 * min(|f|, |f_neighbour| + v)). With voxelize_first, far samples are -10 inside the solid of a
 * voxelization at max(32, min(2048, voxel_res))^3 (shell voxels within 0.866 voxel diagonals of
 * a triangle, grown by one voxel, plus every voxel the border cannot reach) and +10 elsewhere.
 * Equal to the reference's result for the same input, with two choices where the reference's
 * result varies from run to run: each bin's triangles are visited in ascending order (the
 * reference fills bins with atomics), and floods run to their fixpoint (the reference stops its
 * sample flood after 40 unordered passes). A triangle index outside [0, num_vertices) returns
 * DC_HIP_ERR_INVALID_ARG, a non-finite vertex DC_HIP_ERR_NONFINITE, nx/ny/nz outside
 * [1, 4096] or a non-positive spacing DC_HIP_ERR_INVALID_GRID. stats may be NULL. Runs on the
 * context's stream and returns when values is complete. */
int dc_hip_mesh_to_sdf(dc_hip_context* ctx, const float* vertices, int64_t num_vertices,
                       const int32_t* faces, int64_t num_faces, int32_t nx, int32_t ny,
                       int32_t nz, float ox, float oy, float oz, float vx, float vy, float vz,
                       int voxelize_first, int32_t voxel_res, float* values,
                       dc_hip_m2s_stats* stats);

/* Device memory for results that leave the library, such as a mesh_to_sdf grid kept on the
 * device (the reference's compute_mesh_sdf_device_cuda returns cudaMalloc'd values,
 * src/cuda/mesh_to_sdf_cuda.cu:759-762, freed by free_mesh_sdf_device_cuda, :1004-1009):
 * allocated on the context's device, released with dc_hip_free. */
int dc_hip_malloc(dc_hip_context* ctx, int64_t bytes, void** ptr);
int dc_hip_free(void* ptr);

/* Normal of a crossing edge's hermite entry in the sparse extraction (the reference's
 * NormalComputationMode, include/dc_cuda.h:5-9). */
typedef enum dc_hip_normal_mode {
  DC_HIP_NORMAL_FINITE_DIFFERENCE = 0,    /* gradient of the cell's trilinear interpolant */
  DC_HIP_NORMAL_EDGE_GRADIENT = 1,        /* the ends' central differences, blended along the edge */
  DC_HIP_NORMAL_PRECOMPUTED_GRADIENT = 2  /* the reference passes no gradient grid to its kernels
                                             (dc_cuda_sparse.cu:1310, :1336): as mode 0 */
} dc_hip_normal_mode;

/* Counts and times of one sparse extraction, with the reference's DualContouringStats names
 * (include/dc_backend.h:44-78) for the fields its sparse backends fill (dc_cuda_sparse.cu:
 * 1436-1482). */
typedef struct dc_hip_sparse_stats {
  int32_t nx, ny, nz, brick_size;
  int64_t total_cells;    /* (nx-1)(ny-1)(nz-1) */
  int64_t active_cells;
  int64_t total_bricks;   /* ceil((nx-1)/B) ceil((ny-1)/B) ceil((nz-1)/B) */
  int64_t active_bricks;
  int64_t vertex_count, face_count;
  int64_t qef_fallback_count, clamp_count, invalid_count;
  int64_t ambiguous_cells, multi_vertex_cells, one_vertex_cells, two_vertex_cells;
  int64_t split_rejection_count, bad_qef_count, faces_skipped_due_to_missing_cluster;
  float upload_ms;
  float active_brick_marking_ms;     /* sample bit planes and brick flags (reads the grid) */
  float active_brick_compaction_ms;  /* 0: the bricks stay a flag table, no list is compacted */
  float active_cell_marking_ms;
  float active_cell_compaction_ms;   /* active-plane scan (ranks and the cell list) */
  float marking_ms, compaction_ms;
  float qef_ms;                      /* vertices (multi-vertex form: with their scan) */
  float face_count_ms;               /* the quad pass: each wave's quads into its staging run */
  float face_prefix_sum_ms;          /* scan of the per-wave quad counts */
  float face_fill_ms;                /* the staging runs copied into face order */
  float face_emission_ms;            /* the three above */
  float download_ms;                 /* 0 here: the mesh stays on the device */
  float total_ms;                    /* host wall time of the call */
  float sample_pass_ms;              /* the grid-reading kernel alone */
} dc_hip_sparse_stats;

/* Replaces CudaSparseDualContouringBackend::extract / ::extract_device and, with multi_vertex set,
 * CudaSparseMvdcDualContouringBackend (include/dc_cuda.h:37-62, src/cuda/dc_cuda_sparse.cu:
 * 1134-1485; bound at src/cpp/bindings.cpp:224-262, selected by dc_cli -b cuda-sparse /
 * cuda-sparse-mvdc, src/cpp/dc_cli.cpp:317-321). Bricks of brick_size cells per axis are active
 * when their samples hold a negative and a positive value; their cells are active when the 8
 * corners hold both signs or one lies within 0.45 vx of zero; each active cell solves one FP32
 * QEF (multi_vertex: two when its hermite normals split into two opposed clusters); a strict
 * sign change or a thin edge (same sign, both ends within 0.45 vx, opposed gradients) on a
 * cell's first-corner x, y or z edge emits a quad (a thin edge also its reverse). Vertices
 * follow the active cells in linear cell order, faces the active cells' x, y, z edges: the
 * result equals the reference's, bit for bit (the reference's nvcc contracts a*b+c into FMAs by
 * default; its FP32 values match this result's where it is built with contraction off).
 * grid->values: host memory (dc_hip_extract_sparse) or device memory (_device). brick_size >= 1,
 * normal_mode a dc_hip_normal_mode, else DC_HIP_ERR_INVALID_ARG; dimensions under 2 or above
 * 65536 DC_HIP_ERR_INVALID_GRID; a NaN or Inf sample DC_HIP_ERR_NONFINITE. The mesh stays on
 * the device until the next sparse call; the reference's chunk_size has no effect there either.
 * stats may be NULL. Runs on the context's stream and returns when the mesh is complete. */
int dc_hip_extract_sparse(dc_hip_context* ctx, const dc_hip_grid* grid, int32_t brick_size,
                          int normal_mode, int multi_vertex, dc_hip_sparse_stats* stats);
int dc_hip_extract_sparse_device(dc_hip_context* ctx, const dc_hip_grid* grid, int32_t brick_size,
                                 int normal_mode, int multi_vertex, dc_hip_sparse_stats* stats);
/* The last sparse mesh: counts and device pointers (valid until the next sparse call) ... */
int dc_hip_sparse_result_device(dc_hip_context* ctx, int64_t* n_vertices, int64_t* n_faces,
                                const float** vertices, const int32_t** faces);
/* ... or a copy into vertices (3 floats per vertex) and faces (4 int32 per quad), host or device
 * memory (the reference's download, dc_cuda_sparse.cu:1395-1405). Either pointer may be NULL. */
int dc_hip_sparse_copy_result(dc_hip_context* ctx, float* vertices, int32_t* faces);

const char* dc_hip_last_error(dc_hip_context* ctx);
const char* dc_hip_status_string(int status);

#ifdef __cplusplus
}
#endif

#endif /* DC_HIP_H */
