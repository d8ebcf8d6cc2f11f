/*
 * sibson_oracle.h -- CPU restatement of natinterp3d's insertion-free Sibson
 * query path, used ONLY as the checker for the HIP implementation (tests/,
 * __graft_entry__.smoke, bench.py's CPU-baseline leg).  Never linked into
 * the product.
 *
 * The mesh is passed in the packed layout the reference's packMeshSimplices
 * produces (natural_insertionfree.c:397-450): tets in linked-list order,
 * vertex slots 0..N-1 for data points and N..N+3 for the super-simplex
 * vertices (reference indices -1..-4, delaunay.c:858-862), k-d node i = tet i.
 */
#ifndef SIBSON_ORACLE_H
#define SIBSON_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    int n_data;             /* N: data points (super vertices are slots N..N+3) */
    int n_tets;             /* T */
    const double *verts;    /* [(N+4)*3] final (possibly perturbed) coordinates */
    const int *tet_v;       /* [T*4] vertex slots, p[0..3] order */
    const int *tet_n;       /* [T*4] neighbour tets, s[0..3] order, -1 = none */
    const int *kd_left;     /* [T] */
    const int *kd_right;    /* [T] */
    const double *kd_pos;   /* [T*3] tet centroids */
    const double *kd_rect;  /* [6] min[3], max[3] of all centroids */
} orc_mesh;

/* Per-query work counters, summed over a call (the inputs of the algorithmic
 * byte and flop counts). */
typedef struct {
    long long queries;
    long long kd_nodes;      /* k-d nodes entered */
    long long walk_tests;    /* orient3d tests in the walk */
    long long insphere;      /* insphere tests in the cavity search */
    long long cavity;        /* cavity tets */
    long long boundary;      /* boundary faces */
    long long face_cc;       /* distinct face circumcentres */
    long long sub3;          /* old (cavity) subcell volumes */
    long long sub4;          /* new (virtual tet) subcell volumes */
    long long nnz;           /* CSR entries */
    long long dist_verts;    /* distinct vertices read by cavity tets */
    long long max_cavity;    /* largest cavity seen */
    long long max_visited;   /* largest visited set seen */
    long long max_stack;     /* deepest cavity-search stack */
    long long max_boundary;  /* most boundary faces */
    long long max_nnz;       /* longest row */
    long long coincident;    /* rows produced by the coincident-vertex path */
    long long empty_rows;    /* rows with no entries */
    long long tets_read;     /* distinct tets read (walk path and cavity search) */
    long long verts_read;    /* distinct vertices of those tets */
} orc_counters;

/* Same contract as getInsertionFreeWeights[Parallel]
 * (natural_insertionfree.h:76-84): queries already centred, *vals and *cols
 * malloc'd here (release with orc_free), rowptr[M+1] owned by the caller.
 * Returns 0, or -1 when nnz exceeds INT32_MAX.  num_threads <= 0 runs serially.
 * counters may be NULL. */
int orc_get_weights(const orc_mesh *m, const double *q, int M, int num_threads,
                    double **vals, int **cols, int *rowptr, orc_counters *counters);

/* Start tet of the cavity search for each query (walk result, -1 = outside). */
void orc_locate(const orc_mesh *m, const double *q, int M, int *start_tet);

/* Per-row cavity tets in search order (for debugging the device kernels):
 * writes up to cap tets for query q, returns the cavity size. */
int orc_cavity(const orc_mesh *m, const double *q, int *out, int cap);

/* The two fast predicates and the circumcentre helpers, exposed for tests. */
double orc_orient3d(const double *a, const double *b, const double *c, const double *d);
double orc_insphere(const double *a, const double *b, const double *c, const double *d,
                    const double *e);
void orc_cc_tri(const double *a, const double *b, const double *c, double *out);
void orc_cc_tet(const double *a, const double *b, const double *c, const double *d, double *out);
double orc_subcell(const double *pk, const double *pa, const double *pb, const double *pc,
                   const double *ct, const double *fab, const double *fac, const double *fbc);

void orc_free(void *p);

#ifdef __cplusplus
}
#endif
#endif
