// nat_internal.h -- device data layout and kernel launchers shared by the
// C ABI (nat_capi.cpp) and the kernels (sibson_kernels.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace nat {

// Per-problem parameters of the start-tet search (and of the Morton
// bucketing of its queries).  A mesh holds one problem, a batch many.
struct LocParams {
    double g0[3];        // grid origin
    double gh, ginv;     // cell edge and its inverse
    int gn[3];           // cells per axis
    int cell_base;       // first entry of this problem's cells in cell_start
    int kd_root;         // tet index of the k-d root (the builder's first tet)
    int pad_;
    double rmin[3], rmax[3];  // k-d root hyper-rectangle
    double bbox[6];           // data bounding box min[3], max[3]
};

// Immutable device mesh.  All arrays are packed SoA, 16- or 32-byte records
// so that one tet or one vertex is one aligned vector load.  Within each
// problem the data vertices are numbered by the Morton code of their
// position and the tets (k-d nodes) by that of their centroid, so that the
// records neighbouring queries read share cache lines; the host mesh keeps
// the builder's numbering and CSR columns are mapped back to key indices
// when rows are gathered.  Indices are labels to every kernel: orders come
// from vertex slots, adjacency slots and the k-d tree's shape, never from
// index comparisons.  A batch of independent problems is one mesh with
// global numbering: the data vertices of every problem first (problem p's at
// its vertex offset), then the four super vertices of each problem; tets and
// k-d nodes at the problem's tet offset.  The super-vertex test stays
// "slot >= n_data".
// One tet in one 128-byte line: its vertices' coordinates (p0..p3, x y z
// each, packed into three double4), its vertex slots (tet_v, orientation bit
// included) and its neighbours (tet_n).  The cavity search reads a tested
// tet with one request instead of a tet load followed by four vertex loads.
struct alignas(128) FatTet {
    double4 c0, c1, c2;  // x0 y0 z0 x1 | y1 z1 x2 y2 | z2 x3 y3 z3
    int4 v;
    int4 n;
};
static_assert(sizeof(FatTet) == 128, "one line per tet");

struct SphereTet;  // sphere.h: circumsphere and certificate constants of a tet

struct DevMesh {
    const double4 *verts;  // [N+4P] x,y,z,0; slots N.. are the super vertices
    const int4 *tet_v;     // [T] vertex slots p0..p3; bit 31 of .w set when
                           //     orient3d(p0,p1,p2,p3) <= 0
    const int4 *tet_n;     // [T] neighbour across face 0..3, -1 = none
    const FatTet *fat;     // [T] the same tets with their vertex coordinates inline
    const SphereTet *sph;  // [T] circumsphere, certificate constants and neighbours per
                           //     tet: the cavity search's in-sphere shortcut (sphere.h)
    const double4 *kd;     // [T] k-d node i = tet i: centroid x,y,z and
                           //     .w = bit pattern of int2(left, right)
    int n_data;            // N (all problems)
    int n_tets;            // T (all problems)
    // Uniform grid over the same tet centroids, for the exact nearest-centroid
    // search that replaces the k-d descent whenever its answer is certified.
    const int *cell_start;    // per problem [ncell + 1] from LocParams::cell_base:
                              // records of cell c at [cell_start[c], cell_start[c+1])
    const double4 *cell_pts;  // [T] centroid x,y,z; .w = bit pattern of the tet index (int64)
    const LocParams *lp;      // [P]
    const int *prob;          // per query: its problem (set per call; NULL = problem 0)
    // Per problem, the row-path policy: low byte = budget of tets the 16-lane
    // group kernel may visit (0 sends the problem's queries straight to the
    // thread-per-query levels), high byte = first thread-per-query level for
    // the queries that leave the group kernel.  Read at launch time.
    const unsigned short *route;
    // Step-3 subcell volumes, which depend on the mesh alone: per tet, per
    // vertex VI, per choice of vertex order of the three face circumcentres
    // it uses (the face mask's bits of the other three vertices, as the
    // record's fcmask), at [(t * 16 + table_pos(fcmask, dm)) * 4 + VI] with dm
    // = sphere_dm(sph[t]): bit K set when the circumcentre of face K in the
    // neighbour's vertex order differs from the one in the tet's own order
    // (only those bits of a mask change a value); NULL until built (computed
    // inline then).
    const double *old_vol;
    // Predicates of the query path: 0 = the plain FP64 signs, 1 = FP64 where
    // the static error bound certifies the sign, else the exact sign.  With
    // 1, robust_ctr[0..2] count exact evaluations, decisions that differ
    // from the FP64 sign, and tests too wide for the exact evaluator.
    int robust;
    unsigned long long *robust_ctr;
};

__device__ __forceinline__ unsigned route_of(const DevMesh &m, int qi) {
    return m.route[m.prob ? m.prob[qi] : 0];
}

__device__ __forceinline__ const LocParams &loc_of(const DevMesh &m, int qi) {
    return m.lp[m.prob ? m.prob[qi] : 0];
}

// Longest row the level-0 kernels handle (topology LDS neighbour list and
// the geometry kernels' per-thread LDS accumulators); longer rows go to
// level 1.  Threads per block of the geometry kernels.
constexpr int kLevel0Row = 32;
constexpr int kGeomThreads = 128;

// Per-thread capacities.  A query that needs more marks itself for the next,
// larger level.
struct Caps {
    int kd;          // k-d descent stack frames (>= tree depth)
    int stack;       // cavity-search stack
    int vis_log2;    // log2 of the visited-set hash size
    int cav;         // cavity tets
    int bnd;         // boundary faces
    int nb;          // natural neighbours (row length)
    int nbmap_log2;  // log2 of the vertex -> slot hash size
};

// Per-thread scratch.  Stacks and lists are interleaved (element k of thread
// t at [k * nthr + t]); hash tables are contiguous per thread (slot h of
// thread t at [t << log2 | h]) so that one query's probes stay on a few
// cache lines.
struct Scratch {
    int *kd_node;     // [kd]   node | phase << 30
    double *kd_saved; // [kd]   saved rectangle bound
    int *stack;       // [stack]
    int2 *vis;        // [1 << vis_log2]  (tet, stamp)
    int *vis_pos;     // [1 << vis_log2]  cavity position, -1 = tested, outside
    int *cav;         // [cav]
    int *bnd;         // [bnd]  cavity position << 2 | face
    int *nb_idx;      // [nb]
    double *nb_vol;   // [nb]
    int2 *nbmap;      // [1 << nbmap_log2]  (vertex, stamp)
    int *nbslot;      // [1 << nbmap_log2]  slot of the vertex
    int *epoch;       // [1]    per-thread stamp counter, persists across launches
    int nthr;
};

// Per-query topology records written by the topology kernel and consumed by
// the geometry kernel (16-byte units, one contiguous record per query).
struct Topo {
    int4 *rec;                   // record pool
    long long rec_cap;           // in int4 units
    unsigned long long *rec_used;
    long long *rec_off;          // [M] record offset, -1 = row made by another level
    // [items] per item of the search: (query, record offset low, high,
    // C | B << 8 | room << 16), .w = -1 when the order kernels have nothing
    // to do for it; NULL: they read the offset and the header instead
    int4 *itab = nullptr;
};

struct RowOut {
    double *pool_vals;           // row entries in completion order
    int *pool_cols;
    long long pool_cap;
    unsigned long long *pool_used;
    long long *row_off;          // [M] offset of each row in the pool
    int *row_cnt;                // [M] row length
    double *row_total;           // [M] total stolen volume of each row (level 0 rows)
    int *ovf_list;               // queries that exceeded the level's capacities
    unsigned int *ovf_count;
    // when a topology will be saved (save.rec set): where the wave-per-query
    // rows put their big records (header .w = 2) for the backward
    Topo save{};
};

// Backward inputs: dL/dy rows and the interpolated values, both float64;
// the forward's CSR (rows in neighbour-slot order) and row totals.
struct GradIn {
    const double *gy;         // [M, D]
    const double *values;     // [N, D]
    int D;
    double *dq;               // [M, 3] out: dL/dq
    const int *rowptr;        // [M + 1]
    const double *w;          // [nnz] weights
    const double *row_total;  // [M] total stolen volume (rows with a record)
    const int4 *rec;          // the saved records (big records of the wave-per-query rows)
    const long long *rec_off; // [M]
};

size_t scratch_bytes_per_thread(const Caps &c);
void scratch_carve(void *base, const Caps &c, int nthr, Scratch *s);

// the query order: a stable sort of the keys problem-major, then 3*bits-bit
// Morton cells of the problem's box (equal keys in index order), in
// d_work of morton_sort_bytes(M) bytes
size_t morton_sort_bytes(int M);
hipError_t launch_morton_sort(const double *d_q, int M, const LocParams *d_lp,
                              const int *d_prob, int P, int bits, int *d_order, void *d_work,
                              hipStream_t st);
// prob[i] = p with offsets[p] <= i < offsets[p+1]
hipError_t launch_problem_ids(const long long *d_offsets, int P, int M, int *d_prob,
                              hipStream_t st);
int rows_blocks_per_cu();  // resident 256-thread blocks of the fused row kernel per CU
int topo_blocks_per_cu();   // 64-thread blocks of the topology search kernel
int geom_blocks_per_cu();       // kGeomThreads-thread blocks
int geom_rows_blocks_per_cu(bool mixed);  // blocks of the two-lanes-per-row geometry kernel
int grad_geom_blocks_per_cu();  // resident blocks of the gradient geometry kernel
// fused row kernel: used for the larger capacity levels
// d_n_items != NULL: the item count is read on the device (graph form)
hipError_t launch_rows(const DevMesh &m, const Caps &caps, const Scratch &sc, int level,
                       int nblocks, const double *d_q, const int *d_items, int n_items,
                       const unsigned *d_n_items, unsigned int *d_work, const RowOut &out,
                       hipStream_t st);
// zero-fill of n 32-bit words (a kernel, so that captured graphs hold no memset)
hipError_t launch_zero_u32(void *d_p, size_t n, hipStream_t st);
// level 0: start tets, topology records, then rows from them
int locate_blocks_per_cu();  // 256-thread blocks
hipError_t launch_locate(const DevMesh &m, const Caps &caps, const Scratch &sc, int nblocks,
                         const double *d_q, const int *d_items, int n_items, int *d_start,
                         hipStream_t st);
// topology: the quad search (64-thread blocks of 4 quads of 4 queries, 16 lanes
// per quad), then one thread per query for the order rules
hipError_t launch_topo_search(const DevMesh &m, int nblocks, const double *d_q,
                              const int *d_items, int n_items, const int *d_start,
                              unsigned int *d_work, const Topo &topo, unsigned int *ovf_count,
                              int *ovf_list, hipStream_t st);
hipError_t launch_topo_order(const DevMesh &m, const int *d_items, int n_items, const Topo &topo,
                             unsigned int *ovf_count, int *ovf_list, hipStream_t st);
hipError_t launch_topo(const DevMesh &m, int nblocks, const double *d_q, const int *d_items,
                       int n_items, const int *d_start, unsigned int *d_work, const Topo &topo,
                       unsigned int *ovf_count, int *ovf_list, hipStream_t st);
// the items regrouped by record size within spatial windows (geometry order)
hipError_t launch_work_sort(const int *d_items, int n, const Topo &topo, int *d_out,
                            hipStream_t st);
// mix != NULL: mixed precision (FP32 where the condition estimate is at least
// mixed_tau; rows with sum |vol| > mixed_kappa |sum vol| redone in FP64),
// counts of FP32 / FP64 tets and faces and of redone rows added to mix[0..2]
// nblocks: grid of the thread-per-row kernel (no step-3 table, or
// nblocks_rows = 0); nblocks_rows: grid of the two-lanes-per-row kernel
// (FP64, or mixed precision when mix is given).
hipError_t launch_geom(const DevMesh &m, int nblocks, int nblocks_rows, const double *d_q,
                       const int *d_items, int n_items, unsigned int *d_work, const Topo &topo,
                       const RowOut &out, float mixed_tau, float mixed_kappa,
                       unsigned long long *mix, hipStream_t st);
// backward: dL/dq from the topology records, then the fused kernel for the
// queries without one (and those that overflow it); W^T g
hipError_t launch_grad_geom(const DevMesh &m, int nblocks, const double *d_q, const int *d_items,
                            int n_items, unsigned int *d_work, const Topo &topo,
                            const GradIn &g, int *ovf_list, unsigned int *ovf_count,
                            hipStream_t st);
// the cavity search's per-tet circumspheres (out: n_tets records), from fat
hipError_t launch_sphere(const DevMesh &m, SphereTet *out, hipStream_t st);
// the step-3 table of DevMesh::old_vol (out: n_tets * 64 doubles)
hipError_t launch_old_vol(const DevMesh &m, double *out, SphereTet *sph, hipStream_t st);
// wave-per-query rows (cavities beyond the group kernel): per-wave scratch
// of wave_scratch_size() bytes for each of nblocks one-wave blocks
size_t wave_scratch_size();
int wave_blocks_per_cu();
hipError_t launch_wave_rows(const DevMesh &m, char *wscr, int nblocks, const double *d_q,
                            const int *d_items, int n_items, const unsigned *d_n_items,
                            const int *d_start, unsigned int *d_work, const RowOut &out,
                            int *ovf_list, unsigned int *ovf_count, hipStream_t st);
hipError_t launch_wave_grad(const DevMesh &m, char *wscr, int nblocks, const double *d_q,
                            const int *d_items, int n_items, const int *d_start,
                            unsigned int *d_work, const GradIn &g, int *ovf_list,
                            unsigned int *ovf_count, hipStream_t st);
hipError_t launch_grad_rows(const DevMesh &m, const Caps &caps, const Scratch &sc, int level,
                            int nblocks, const double *d_q, const int *d_items, int n_items,
                            unsigned int *d_work, const GradIn &g, int *ovf_list,
                            unsigned int *ovf_count, hipStream_t st);
// insphere decisions behind saved rows re-tested under the static error bound
hipError_t launch_predicate_audit(const DevMesh &m, const double *d_q, int n, const Topo &topo,
                                  unsigned long long *d_counts, hipStream_t st);
// W^T g into out [N, D]; with P > 1 and priv (P x N x D doubles) the blocks
// accumulate into P private copies, summed into out at the end
hipError_t launch_spmm_t(int M, int N, int D, const int *rowptr, const int *cols,
                         const double *vals, const double *gy, const int *order, double *out,
                         double *priv, int P, hipStream_t st);
// error statement of mixed rows: save the sampled rows (order[i * stride],
// i < n) aside, then, after an FP64 geometry pass over them, compare and put
// the mixed rows back; out[0..2] = max |dw| bits, rows, entries
hipError_t launch_mixed_check_save(const int *d_order, int n, int stride,
                                   const long long *row_off, const int *row_cnt,
                                   const double *row_total, int *items, long long *s_off,
                                   int *s_cnt, double *s_total, hipStream_t st);
hipError_t launch_mixed_check_compare(int n, const int *items, const long long *rec_off,
                                      long long *row_off, int *row_cnt, double *row_total,
                                      const long long *s_off, const int *s_cnt,
                                      const double *s_total, const double *pool_vals,
                                      const int *pool_cols, long long pool_cap,
                                      unsigned long long *out, hipStream_t st);
// exclusive scan of n int32 counts into int64 offsets; total in out[n]
hipError_t launch_scan_counts(const int *d_in, long long *d_out, int n, long long *d_tmp,
                              hipStream_t st);
size_t scan_tmp_elems(int n);
// rows reaching past cap entries (pool or output) are skipped; columns pass
// through vmap (device vertex index -> key index)
hipError_t launch_gather_csr(const long long *d_rowptr64, const long long *d_row_off,
                             const int *d_row_cnt, const double *pool_vals,
                             const int *pool_cols, const int *vmap, int M, long long cap,
                             double *d_vals, int *d_cols, int *d_rowptr32, hipStream_t st);
// W @ x from the row pool of a compute (rows in order, device columns, x in
// device vertex order; xdt 0 = float32, 1 = float64), D = 1..4
hipError_t launch_pool_spmm(int M, int D, const long long *row_off, const int *row_cnt,
                            const int *pool_cols, const double *pool_vals, const void *x_dev,
                            int xdt, const int *order, double *y, hipStream_t st);
// out[i, :] = in[idx[i], :], rows of D elements of elt (4 or 8) bytes, i < n
hipError_t launch_gather_rows_bytes(const void *in, const int *idx, int n, int D, int elt,
                                    void *out, hipStream_t st);
// out[i, :] = in[idx[i], :], rows of D doubles, i < n
hipError_t launch_gather_rows(const double *in, const int *idx, int n, int D, double *out,
                              hipStream_t st);
hipError_t launch_sort_rows(int M, const int *rowptr, int *cols, void *vals, int vdt,
                            hipStream_t st);
// rows in order (a permutation of 0..M-1) when it is given
hipError_t launch_spmm(int M, int D, const int *rowptr, const int *cols, const void *vals,
                       int vdt, const void *x, int xdt, const int *order, void *y, int ydt,
                       hipStream_t st);

}  // namespace nat
