/*
 * natinterp3d_hip.h -- C ABI of the MI355X (gfx950) Sibson query path.
 *
 * The reference binds four C functions from Cython
 * (natinterp3d_cython.pyx:17-35):
 *
 *     buildNewMeshAndVertices   natural.h:16            -> nat_mesh_build
 *     freeMeshAndVertices       natural.h:18            -> nat_mesh_free
 *     getInsertionFreeWeights   natural_insertionfree.h:76-79
 *     getInsertionFreeWeightsParallel                    -> nat_get_weights
 *                               natural_insertionfree.h:81-84
 *
 * The entry points below keep their argument meaning, ownership and error
 * convention: queries are float64 [M,3] already centred by the caller, the
 * caller owns rowptr[M+1], the callee malloc()s values[nnz] and cols[nnz]
 * (release with free() or nat_free_host), and -1 means "nnz exceeds
 * INT32_MAX" (nothing allocated).  Other negative codes are HIP or argument
 * errors; nat_last_error() describes the last one on the calling thread.
 *
 * The remaining entry points have no reference counterpart; they serve
 * callers whose data already lives in device memory (PyTorch tensors).
 * All pointers named d_* are device pointers; `stream` is a hipStream_t
 * (NULL = the legacy default stream).  Calls on one nat_mesh are serialised
 * internally; different meshes may be used concurrently.
 */
#ifndef NATINTERP3D_HIP_H
#define NATINTERP3D_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct nat_mesh nat_mesh;

typedef struct {
    int32_t n_data;          /* N data points */
    int32_t n_tets;          /* T tets, including those touching the super simplex */
    int32_t kd_depth;        /* depth of the k-d tree over tet centroids */
    int32_t coplanar;        /* degeneracies the builder perturbed away */
    int32_t cospherical;
    int32_t edge_anomalies;  /* non-manifold edge groups met while linking tets */
    int32_t device;          /* HIP device the mesh lives on */
    int32_t num_cus;         /* compute units of that device */
    int64_t device_bytes;    /* device memory held by the mesh (not workspace) */
    int64_t workspace_bytes; /* device workspace currently held */
    int32_t n_problems;      /* 1, or the number of problems of a batch mesh */
    int32_t builder;         /* 0: host build (the reference's mesh), 1: GPU build */
    int32_t build_rounds;    /* GPU build: insertion rounds */
    double build_seconds;    /* wall time of the tetrahedralisation */
} nat_mesh_info;

/* Replaces buildNewMeshAndVertices (natural.c:114-118) plus the one-time
 * packMeshSimplices / precomputeOrientSigns (natural_insertionfree.c:397-476):
 * builds the Delaunay mesh of data_points [num_data_points,3] (float64,
 * centred) on the host, reproducing the reference mesh exactly, then uploads
 * it to `device` as immutable packed arrays.  device = -1 builds a host-only
 * mesh (no GPU needed) that supports nat_mesh_get_info and nat_mesh_export
 * only.  Returns 0, or a negative code (-2: a tet of non-positive
 * orientation, where the reference exits). */
int nat_mesh_build(const double *data_points, int num_data_points, int device,
                   nat_mesh **out);

/* Opt-in GPU Delaunay build (no reference counterpart; it replaces the
 * host insertion of buildMesh, delaunay.c:108-161, 561-750).  flags 0 is
 * nat_mesh_build.  flags NAT_BUILD_GPU: the tetrahedralisation of the same
 * points and the same super simplex is made on `device` (>= 0) in rounds of
 * parallel Bowyer-Watson insertions with exact predicates, then packed with a
 * balanced k-d tree over the tet centroids and uploaded.  For points in
 * general position the Delaunay tetrahedralisation is unique, so the tets
 * are the reference's; their order, the vertex order inside each tet and the
 * k-d tree are not, so rows agree with the reference's as per-row column
 * sets, with weights equal up to rounding (FP64 operation order).  The GPU
 * build does not perturb: repeated points or exactly degenerate cavities
 * return -3 (build with flags 0 instead); -4 when a cavity exceeds 512 tets
 * or 1024 boundary faces. */
#define NAT_BUILD_GPU 1
int nat_mesh_build_ex(const double *data_points, int num_data_points, int device, int flags,
                      nat_mesh **out);

/* Exact audit of a single-problem mesh's tetrahedralisation on its device
 * (no reference counterpart): counts[0] = faces with a neighbour tested,
 * [1] = tets of non-positive orientation, [2] = neighbour links that are not
 * mutual, [3] = faces whose neighbour's far vertex lies strictly inside the
 * tet's circumsphere (0 for a Delaunay mesh: local Delaunay on every face). */
int nat_mesh_delaunay_audit(const nat_mesh *m, int64_t *counts);

/* Replaces freeMeshAndVertices (natural.c:120-123). */
void nat_mesh_free(nat_mesh *m);

int nat_mesh_get_info(const nat_mesh *m, nat_mesh_info *info);

/* Copies the packed host mesh out (test and tooling access): verts[(N+4)*3]
 * (data slots 0..N-1, super vertices N..N+3), tet_v[T*4], tet_n[T*4],
 * kd_left[T], kd_right[T], kd_pos[T*3], kd_rect[6].  Any pointer may be NULL. */
int nat_mesh_export(const nat_mesh *m, double *verts, int32_t *tet_v, int32_t *tet_n,
                    int32_t *kd_left, int32_t *kd_right, double *kd_pos, double *kd_rect);

/* A mesh from packed arrays of the layout nat_mesh_export writes (no
 * reference counterpart: the reference has no mesh serialisation).  The
 * builder does not run: a mesh saved to disk and loaded back gives the same
 * rows as the one built from the keys.  degeneracies = {coplanar, cospherical,
 * edge anomalies} for nat_mesh_get_info, or NULL.  The arrays are checked:
 * indices in range and a k-d tree reaching every node from node 0 exactly
 * once (depth <= 4096); -1 otherwise.  device = -1: host-only mesh. */
int nat_mesh_from_packed(const double *verts, int num_data_points, const int32_t *tet_v,
                         const int32_t *tet_n, int num_tets, const int32_t *kd_left,
                         const int32_t *kd_right, const double *kd_pos, const double *kd_rect,
                         const int32_t *degeneracies, int device, nat_mesh **out);

/* Replaces getInsertionFreeWeights / getInsertionFreeWeightsParallel
 * (natural_insertionfree.c:761-940): host queries in, host CSR out, same
 * ownership and return codes.  The rows are computed on the GPU. */
int nat_get_weights(const double *query_points, int num_query_points, nat_mesh *m,
                    int num_data_points, double **weight_values, int **weight_col_inds,
                    int *weight_row_ptrs);

/* Device-resident form of nat_get_weights, in two steps.
 * nat_weights_compute runs every row of d_queries [M,3] into the mesh's
 * workspace and returns the exact nnz (int64) in *nnz_out (no INT32 limit
 * here).  nat_weights_emit then writes the CSR into caller-owned device
 * buffers d_vals[nnz] (float64), d_cols[nnz] (int32) and d_rowptr[M+1]
 * (int32; returns -1 when nnz > INT32_MAX, like the reference).  The emit
 * must follow the compute on the same mesh with no other call in between. */
int nat_weights_compute(nat_mesh *m, const double *d_queries, int num_query_points,
                        void *stream, int64_t *nnz_out);
int nat_weights_emit(nat_mesh *m, double *d_vals, int32_t *d_cols, int32_t *d_rowptr,
                     void *stream);

/* nat_weights_compute + nat_weights_emit in one call that allocates no
 * memory.  No reference counterpart: it is for callers that cannot allocate
 * -- code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc.  Its workspace is a fixed 64 MiB device array that each GPU the
 * library runs on holds until the process ends; every call on a GPU shares
 * it (its contents are those of the last call), so calls that overlap on one
 * GPU, from several threads or streams and on any mesh, race.  It takes at
 * most 32768 queries per call (-1 beyond) and writes the CSR of d_queries
 * [M,3] into caller-owned device buffers d_vals[capacity] (float64),
 * d_cols[capacity] (int32) and d_rowptr[M+1] (int32); *nnz_out = the
 * entries.  Returns 0; -7 when nnz > capacity (nothing written, *nnz_out
 * says how many are needed); -8 when the rows need more topology records or
 * pool entries than the array holds (fewer queries per call); -4 when a row
 * needs a cavity beyond 4096 tets (use nat_weights_compute).  The rows are
 * those of nat_weights_compute + nat_weights_emit, bit for bit, under the
 * mesh's precision and policy; the mesh's own workspace is not touched (a
 * later nat_weights_emit still refers to the last nat_weights_compute), and
 * the step-3 table is used only if an earlier compute built it. */
int nat_weights_noalloc(nat_mesh *m, const double *d_queries, int num_query_points,
                        double *d_vals, int32_t *d_cols, int32_t *d_rowptr, int64_t capacity,
                        void *stream, int64_t *nnz_out);

/* W @ values on the device, accumulating each row in stored order from zero
 * like scipy's CSR x dense product (natinterp3d.py:131-137):
 * d_out[M,D] = W[M,N] (CSR, d_rowptr int32[M+1], d_cols int32, d_vals)
 * x d_values[N,D].  dtype codes: 0 = float32, 1 = float64; vals_dtype is the
 * weight type, values_dtype the value type, out_dtype the accumulation and
 * output type (the NumPy result type of the two).  Rows are visited in
 * d_row_order when it is given (a permutation of 0..M-1; NULL: 0..M-1); the
 * result does not depend on it, the speed does: a spatial order such as
 * nat_weights_row_order's lets neighbouring rows share the rows of d_values
 * they read in cache. */
int nat_spmm_csr(int num_rows, int num_cols, int dim, const int32_t *d_rowptr,
                 const int32_t *d_cols, const void *d_vals, int vals_dtype,
                 const void *d_values, int values_dtype, const int32_t *d_row_order,
                 void *d_out, int out_dtype, void *stream);

/* Device pointer to the query order of the last nat_weights_compute on this
 * mesh (int32[M], a permutation of 0..M-1 in which neighbouring queries are
 * close in space), for nat_spmm_csr.  Valid until the next compute on the
 * mesh; returns -1 when num_query_points is not that compute's M.  (No
 * reference counterpart: scipy visits rows in index order.) */
int nat_weights_row_order(const nat_mesh *m, int num_query_points, const int32_t **out);

/* W @ values for the rows of the last nat_weights_compute on this mesh, read
 * from the mesh's row pool rather than an emitted CSR: d_out[M, dim] float64
 * = W (float64) x d_values[N, dim] (N = num_data_points, all problems of a
 * batch, checked against the mesh; values_dtype 0 = float32, 1 = float64;
 * dim 1..4), bit for bit what nat_spmm_csr gives on nat_weights_emit's CSR
 * (scipy's stored-order sums, natinterp3d.py:131-137).  The pool rows are the
 * CSR rows in compute order with the mesh's own vertex numbering, so the
 * product reads them nearly in sequence.  Valid until the next compute.
 * (No reference counterpart.) */
int nat_weights_spmm(nat_mesh *m, int num_query_points, const void *d_values,
                     int num_data_points, int values_dtype, int dim, double *d_out,
                     void *stream);

/* Sort the entries of every CSR row by column, in place (d_vals dtype codes
 * as above).  The reference's Interpolator.get_weights casts W to the query
 * dtype (natinterp3d.py:117); for any dtype other than float64 scipy's
 * astype canonicalises the matrix, which sorts each row's columns, and
 * W @ values then accumulates in that order.  Rows hold no duplicate
 * columns, so sorting is the whole canonicalisation. */
int nat_csr_sort_rows(int num_rows, const int32_t *d_rowptr, int32_t *d_cols, void *d_vals,
                      int vals_dtype, void *stream);

/* ---- batches (no reference counterpart: the reference builds and queries
 * one key set at a time).  Many independent problems share one mesh object
 * and every call runs all of their queries in the same kernel launches.
 *
 * nat_batch_build builds problem p's mesh from data_points[p]
 * [num_data_points[p], 3] (centred float64, as nat_mesh_build) on
 * num_threads host threads (<= 0: all cores) and uploads them together.
 * Columns are global: problem p's data point i is column
 * vertex_offsets[p] + i (nat_batch_offsets), so a batch's W is block
 * diagonal and W @ values works on the concatenated values. */
int nat_batch_build(const double *const *data_points, const int32_t *num_data_points,
                    int num_problems, int device, int num_threads, nat_mesh **out);
int nat_batch_offsets(const nat_mesh *m, int64_t *vertex_offsets, int64_t *tet_offsets);

/* nat_weights_compute for a batch: d_queries holds every problem's queries
 * back to back, problem p's at rows [query_offsets[p], query_offsets[p+1])
 * (host array of num_problems + 1).  Follow with nat_weights_emit. */
int nat_batch_weights_compute(nat_mesh *m, const double *d_queries,
                              const int64_t *query_offsets, void *stream, int64_t *nnz_out);

/* ---- graphs (no reference counterpart).  nat_graph_create runs one
 * weights computation of d_queries [M,3] (with query_offsets for a batch
 * mesh, else NULL) on `stream` to size every capacity (+25%), then captures
 * the whole launch sequence -- bucketing, start tets, topology, geometry, the
 * larger levels and the CSR assembly -- in a hipGraph on a private stream.
 * nat_graph_launch replays it on `stream` for whatever d_queries holds now
 * (same M and offsets) and returns nnz; it returns -8 when a capacity was
 * exceeded (compute eagerly instead, or make a new graph), -9 when an
 * eager call on the mesh has reallocated its workspace since, and -11 when
 * nat_mesh_set_precision or nat_mesh_set_predicates has changed the row
 * arithmetic (mode, thresholds or predicates)
 * since the capture: the captured kernels keep the capture's arithmetic, so
 * such a replay would not give the rows the mesh now promises.
 * nat_graph_copy_out copies the last replay's CSR to caller buffers.  The
 * graph uses the mesh's workspace: calls on the mesh are serialised with it. */
typedef struct nat_graph nat_graph;
int nat_graph_create(nat_mesh *m, const double *d_queries, int num_query_points,
                     const int64_t *query_offsets, void *stream, nat_graph **out);
int nat_graph_launch(nat_graph *g, void *stream, int64_t *nnz_out);
int nat_graph_copy_out(nat_graph *g, double *d_vals, int32_t *d_cols, int32_t *d_rowptr,
                       void *stream);
void nat_graph_free(nat_graph *g);

/* ---- backward (no reference counterpart: the reference has no autograd;
 * these serve the torch.autograd.Function of natinterp3d_amd).
 *
 * nat_topology_save copies what the last nat_weights_compute on this mesh
 * learned about its queries -- the cavity, boundary faces and neighbour
 * order of every row the level-0 kernels made, the start tet of every
 * query, and the query order -- into
 * caller-owned device memory, so that the backward pass recomputes the
 * geometry from it instead of searching again.  With d_dst NULL it only
 * reports the size in *bytes.  Returns -7 when capacity is too small. */
int nat_topology_save(nat_mesh *m, void *d_dst, int64_t capacity, int64_t *bytes,
                      void *stream);

/* Lets the next nat_weights_compute of num_query_points queries on this mesh
 * write its topology records straight into d_dst (16-byte aligned, capacity
 * bytes), the buffer a nat_topology_save(m, d_dst, ...) after that compute
 * fills: the save then copies only the per-query arrays, not the records
 * (≈5 GB at 10 M queries).  One-shot: any next compute consumes it.  When the
 * records do not fit, or M differs, the compute keeps them in its workspace
 * and the save copies them as without a target.  If the save goes to another
 * buffer, it copies the records from d_dst, which must stay valid until then.
 * d_dst NULL clears a pending target.  (No reference counterpart.) */
int nat_topology_target(nat_mesh *m, void *d_dst, int64_t capacity, int num_query_points);

/* dL/dq [M,3] float64 of y = W(q) @ values, from dL/dy [M,D] float64:
 * dL/dq_i = sum_j <dL/dy_i, values_j> d W_ij / d q_i, with the Sibson
 * coordinate derivative from the Voronoi facets between q_i and its
 * neighbours (d sigma_j/dq = facet moment about q / |p_j - q|).  d_topo is a
 * nat_topology_save of the forward call on the same centred queries, and
 * d_rowptr / d_vals that call's CSR (nat_weights_emit); rows without a saved
 * record are recomputed.  Empty and coincident rows get 0. */
int nat_grad_queries(nat_mesh *m, const void *d_topo, const double *d_queries,
                     int num_query_points, const int32_t *d_rowptr, const double *d_vals,
                     const double *d_gy, const double *d_values, int dim, double *d_dq,
                     void *stream);

/* Device pointer to the query order (int32[M], a permutation of 0..M-1 in
 * which neighbouring queries are close in space) held inside a saved
 * topology; reads and checks the topology's header. */
int nat_topology_row_order(const void *d_topo, int num_query_points, const int32_t **out);

/* dL/dvalues = W^T @ dL/dy: d_out [N,D] float64 (overwritten) from the
 * device CSR of nat_weights_emit and d_gy [M,D] float64.  Rows are visited
 * in d_row_order when it is given (NULL: 0..M-1); a spatial order such as
 * nat_topology_row_order's keeps the atomic additions of neighbouring rows
 * on the same cache lines.  Float64 atomic accumulation: the order of the
 * additions into one output is not fixed.  When the output is small (at most
 * 512 KB: few columns under many entries), 8 to 64 private copies of it (at
 * most 32 MB) are allocated and freed in stream order (hipMallocAsync) so
 * that the atomics spread over more addresses; they are summed into d_out. */
int nat_spmm_t_csr(int num_rows, int num_cols, int dim, const int32_t *d_rowptr,
                   const int32_t *d_cols, const double *d_vals, const double *d_gy,
                   const int32_t *d_row_order, double *d_out, void *stream);

/* Kernel timings of the last nat_weights_compute / nat_weights_emit on a
 * mesh, from HIP events recorded on the stream the kernels ran on (no
 * reference counterpart; for measurement).  Off by default. */
typedef struct {
    double order_ms;          /* Morton bucketing of the queries */
    double locate_ms;         /* level 0: start tets */
    double topo_ms;           /* level 0: cavities, boundaries, neighbour order */
    double geom_ms;           /* level 0: circumcentres, volumes, rows */
    double rows_ms[4];        /* per capacity level ([0] = topo + geom) */
    int32_t rows_items[4];    /* queries processed per level */
    int32_t levels;           /* levels launched */
    double emit_ms;           /* CSR scan + gather */
    double grad_ms[4];        /* nat_grad_queries: [0] from the saved records,
                                 [1..3] recomputed rows per capacity level */
    int32_t grad_items[4];
    int64_t mixed_fp32;       /* mixed precision: tets and boundary faces done in
                                 FP32 by the last compute (always counted) */
    int64_t mixed_fp64;       /* ... and recomputed in FP64 (ill-conditioned) */
    int64_t mixed_rows_fp64;  /* rows redone in FP64 (cancelling volumes) */
    double wave_ms;           /* wave-per-query level: forward */
    double grad_wave_ms;      /* ... and in nat_grad_queries (with its start tets) */
    int32_t wave_items;
    int32_t grad_wave_items;
    double topo_search_ms;    /* level 0 topology: the cavity search (in topo_ms) */
    double topo_order_ms;     /* ... and the order rules (in topo_ms) */
    int64_t robust_exact;     /* robust predicates (nat_mesh_set_predicates): tests the
                                 static error bound left open, decided exactly */
    int64_t robust_changed;   /* ... of which the exact sign differs from the FP64
                                 sign the reference would have used */
    int64_t robust_unresolved; /* ... too wide for the exact evaluator (FP64 sign kept) */
    double mixed_err_max;     /* mixed precision: max |w_mixed - w_fp64| over the
                                 sampled rows (nat_mesh_set_mixed_check) */
    int64_t mixed_err_rows;   /* ... number of rows sampled (0: no statement) */
    int64_t mixed_err_entries; /* ... and their entries */
} nat_timings;

int nat_mesh_set_timing(nat_mesh *m, int enable);

/* Robustness audit (no reference counterpart): re-tests the insphere
 * decisions behind the rows of a saved topology (nat_topology_save) -- the
 * cavity tets and the tets across its boundary faces -- under Shewchuk's
 * static error bound for the FP64 determinant.  *decisions counts them,
 * *uncertain those whose sign FP64 evaluation does not certify (where an
 * exact predicate could decide differently from the reference). */
int nat_predicate_audit(nat_mesh *m, const void *d_topo, const double *d_queries,
                        int num_query_points, int64_t *decisions, int64_t *uncertain,
                        void *stream);

/* Host evaluation of the same test for n point tuples (a, b, c, d, e), 15
 * doubles each: det[i] = the FP64 insphere determinant (the reference's
 * inspherefast expression), uncertain[i] = 1 when its sign is not certified
 * by the static error bound.  No GPU needed. */
int nat_insphere_classify(const double *points, int n, double *det, int32_t *uncertain);

/* Predicates of the query path (no reference counterpart; the reference's
 * are the plain FP64 determinants, predicates.c:1687-1710 and 3301-3347,
 * and its adaptive ones, predicates.c:3940-4265, are never called).  mode 0
 * (default): the plain FP64 signs, rows bit-identical to the reference.
 * mode 1 (robust): every insphere test of the cavity search and orient3d
 * test of the walk whose FP64 sign the static error bound (Shewchuk's
 * o3derrboundA / isperrboundA) does not certify is decided by the sign of
 * the exact determinant instead; nat_timings counts those tests, the
 * decisions that changed, and the tests whose exponent range is too wide
 * for the exact evaluator (they keep the FP64 sign).  The weights stay FP64
 * (or mixed) geometry over the robust cavity. */
int nat_mesh_set_predicates(nat_mesh *m, int mode);

/* Host evaluation of the robust decision for n tuples (no GPU needed):
 * kind 0 = orient3d(a, b, c, d), 12 doubles per tuple; kind 1 =
 * insphere(a, b, c, d, e), 15 doubles.  sign[i] = the sign of the exact
 * determinant (-1, 0, 1), or 2 when its exponent range is too wide;
 * uncertain[i] = 1 when the static error bound does not certify the FP64
 * sign (the tests the device evaluates exactly). */
int nat_predicate_exact(const double *points, int n, int kind, int32_t *sign,
                        int32_t *uncertain);

/* Row arithmetic of a mesh (no reference counterpart).  mode 0 (default):
 * FP64 with the reference's operation order, weights bit-identical.  mode 1:
 * mixed -- the cavity, boundary and neighbour order still come from the FP64
 * predicates (neighbour sets and CSR columns unchanged), the circumcentres
 * and subcell volumes of the rows made by the first capacity level run in
 * FP32 relative to the query; every tet or boundary face whose condition
 * estimate |det| / (product of edge lengths) is below cond_threshold, every
 * row whose subcell volumes cancel (sum |vol| > cancel_threshold * |sum vol|,
 * the factor by which FP32 rounding of the terms reaches the weights) and
 * every row of the larger levels is computed in FP64.  nat_timings reports
 * how many tets, faces and rows took each path.  The backward pass is FP64
 * in both modes (from the forward's weights). */
int nat_mesh_set_precision(nat_mesh *m, int mode, double cond_threshold,
                           double cancel_threshold);
int nat_mesh_get_timings(const nat_mesh *m, nat_timings *out);

/* Error statement of the mixed rows (no reference counterpart): every
 * stride-th query, in the spatial order the queries are processed in, is
 * computed again in FP64 from the same topology after each mixed
 * nat_weights_compute / nat_get_weights, and nat_timings reports the largest
 * |w_mixed - w_fp64| over those rows (mixed_err_max) with the number of rows
 * and entries it covers.  stride 0 turns it off; default 1024.  Graph
 * replays and nat_weights_noalloc make no statement. */
int nat_mesh_set_mixed_check(nat_mesh *m, int stride);

/* ---- row-path policy and its autotuner (no reference counterpart).
 * Which kernel makes a row changes speed, never the row: every policy gives
 * the same CSR, bit for bit.
 *
 * A query first goes to the group kernel -- the topology search, 16 lanes of
 * a wave for a group of four consecutive queries of the Morton order, its
 * tables in LDS, then the order kernels -- unless its problem's budget is 0;
 * the group gives up on a query once the group has visited more tets than
 * its queries' budgets (at most 96 each, 192 in all, the LDS tables' size),
 * or when the row has more than 64 cavity tets, 136 boundary faces or 32
 * neighbours.
 * Those queries go next, when wave_rows is set, to the wave-per-query level
 * (a wave of 64 lanes per query, tables in global memory, up to 2048 cavity
 * tets, 8192 boundary faces and 2048 neighbours), and then to the
 * thread-per-query levels (one thread per query, tables in global memory:
 * level 1 up to 4096 cavity tets and 2048 neighbours on up to 131072
 * threads, level 2 up to 32768 on 2048, level 3 up to 262144 on 64), starting
 * at first_level. */
typedef struct {
    int32_t first_level;  /* 1 or 2 */
    int32_t work_sort;    /* 1: the geometry kernel takes rows regrouped by
                             cavity size within windows of 4096 queries */
    int32_t topo_grid;    /* group-kernel grid = resident blocks x this (1..4) */
    int32_t geom_grid;    /* geometry-kernel grid = resident blocks x this (1..4) */
    int32_t wave_rows;    /* 1: the wave-per-query level runs */
    int32_t geom_lanes;   /* geometry kernel lanes per row: 2 (0 also
                             means 2) or 1 (one thread per row); a mesh
                             without the step-3 table always runs one
                             thread per row */
    int32_t reserved[2];
} nat_policy;

/* Sets the policy and, when budgets is given, the group kernel's visited-tet
 * budget per problem (budgets[P], 0..96; 0 = thread-per-query rows only).
 * A graph keeps the grid sizes, work sort and geometry lanes of its capture and reads the
 * budgets and first level at every replay. */
int nat_mesh_set_policy(nat_mesh *m, const nat_policy *policy, const uint8_t *budgets);
int nat_mesh_get_policy(const nat_mesh *m, nat_policy *policy, uint8_t *budgets);

typedef struct {
    double default_ms;       /* compute with the default policy (best of repeats) */
    double tuned_ms;         /* compute with the chosen policy */
    int32_t candidates;      /* policies timed */
    int32_t routed_problems; /* problems given budget 0 */
    int64_t routed_queries;  /* their queries */
    double left_group;       /* fraction of queries that left the group kernel
                                under the default policy */
} nat_tune_report;

/* Times row-path policies on the caller's queries (centred float64 [M,3] in
 * device memory; query_offsets[P+1] on the host for a batch mesh, NULL for a
 * single problem) and keeps the fastest on the mesh.  Candidates, tried one
 * dimension at a time from the default: problems whose queries mostly leave
 * the group kernel routed to the larger levels (several thresholds), the
 * wave-per-query level off, first_level 2, smaller budgets for the other
 * problems, the work sort off, larger grids.  Each candidate is timed as the best of `repeats` computes
 * (host clock, synchronised).  The last compute uses the chosen policy, so
 * nat_weights_emit may follow.  Earlier topologies and graphs stay valid. */
int nat_mesh_autotune(nat_mesh *m, const double *d_queries, int64_t num_query_points,
                      const int64_t *query_offsets, int repeats, void *stream,
                      nat_tune_report *report);

/* Test hook of the overflow convention: when the environment variable
 * NATINTERP3D_TEST_FORGE_NNZ holds a positive integer at a compute, the
 * nnz > INT32_MAX checks of nat_get_weights, nat_weights_compute's *nnz_out
 * and nat_weights_emit see the true total plus that many entries (so a test
 * reaches the -1 branches without 2^31 entries); the rows are unchanged. */

/* Release memory returned by nat_get_weights. */
void nat_free_host(void *p);

/* Description of the last error on the calling thread ("" if none). */
const char *nat_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
