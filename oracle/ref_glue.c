/*
 * ref_glue.c -- accessor glue for the reference build in oracle/_ref/.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is compiled together with the
 * reference's own single-translation-unit build (src/natinterp3d/unity.c,
 * included below from where it lies) into oracle/_ref/libnatref.so.  It adds
 * nothing to the reference algorithm: it exposes the reference mesh, k-d tree
 * and CSR output through plain C entry points so that tests can compare the
 * product's mesh builder and the oracle restatement against the real code.
 *
 * Nothing in the product (natinterp3d_amd/) may load this library.
 */
#include "unity.c"

#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef struct {
    mesh *m;
    vertex *ps;
    int n;
} ref_handle;

/* Build the reference mesh (natural.c:114-118).  The reference perturbs
 * degenerate points with the process-global libc rand() (delaunay.c:513-522);
 * reset_rng=1 puts that generator into the state of a fresh process
 * (srand(1)), which is the state the product's builder restates. */
void *ref_build(const double *pts, int n, int reset_rng) {
    if (reset_rng) srand(1);
    ref_handle *h = (ref_handle *)malloc(sizeof(ref_handle));
    h->n = n;
    buildNewMeshAndVertices((double *)pts, n, &h->m, &h->ps);
    return h;
}

void ref_free(void *hp) {
    ref_handle *h = (ref_handle *)hp;
    freeMeshAndVertices(h->m, h->ps);
    free(h);
}

int ref_num_tets(void *hp) {
    ref_handle *h = (ref_handle *)hp;
    return getNumSimplicies(h->m);
}

void ref_degeneracies(void *hp, int *coplanar, int *cospherical) {
    ref_handle *h = (ref_handle *)hp;
    *coplanar = numPlanarDegenerecies(h->m);
    *cospherical = numSphericalDegenerecies(h->m);
}

static int vslot(ref_handle *h, vertex *v) {
    if (v >= h->ps && v < h->ps + h->n) return (int)(v - h->ps);
    for (int i = 0; i < 4; i++)
        if (v == &h->m->superVerticies[i]) return h->n + i;
    return -1000000;
}

typedef struct { uintptr_t p; int idx; } ptr_idx;

static int cmp_ptr_idx(const void *a, const void *b) {
    uintptr_t x = ((const ptr_idx *)a)->p, y = ((const ptr_idx *)b)->p;
    return x < y ? -1 : (x > y ? 1 : 0);
}

static int lookup(ptr_idx *tab, int n, const void *p) {
    int lo = 0, hi = n - 1;
    uintptr_t key = (uintptr_t)p;
    while (lo <= hi) {
        int mid = (lo + hi) / 2;
        if (tab[mid].p == key) return tab[mid].idx;
        if (tab[mid].p < key) lo = mid + 1; else hi = mid - 1;
    }
    return -1000000;
}

/* Export the mesh in linked-list order (the order packMeshSimplices uses,
 * natural_insertionfree.c:413-425) as packed arrays:
 *   verts [(n+4)*3]   final (possibly perturbed) coordinates; the 4 super
 *                     vertices (reference index -1..-4) occupy slots n..n+3
 *   tet_v [T*4]       vertex slots in p[0..3] order
 *   tet_n [T*4]       neighbour tet index in s[0..3] order, -1 for NULL
 *   kd_left/right [T] child node index (-1 for none); node i is tet i
 *   kd_pos [T*3]      node position (tet centroid)
 *   kd_data [T]       node payload (must equal the node index)
 *   kd_rect [6]       root hyper-rectangle min[3], max[3]
 * Must be called before the first ref_get_weights, which repacks the mesh. */
int ref_export(void *hp, double *verts, int *tet_v, int *tet_n,
               int *kd_left, int *kd_right, double *kd_pos, int *kd_data,
               double *kd_rect) {
    ref_handle *h = (ref_handle *)hp;
    mesh *m = h->m;
    if (m->packed_simplices) return -1;
    int T = getNumSimplicies(m);
    for (int i = 0; i < h->n; i++)
        for (int d = 0; d < 3; d++) verts[i * 3 + d] = h->ps[i].v[d];
    for (int i = 0; i < 4; i++)
        for (int d = 0; d < 3; d++) verts[(h->n + i) * 3 + d] = m->superVerticies[i].v[d];

    ptr_idx *tab = (ptr_idx *)malloc(sizeof(ptr_idx) * (T > 0 ? T : 1));
    listNode *iter = topOfLinkedList(m->tets);
    simplex *s;
    int idx = 0;
    while ((s = nextElement(m->tets, &iter))) {
        tab[idx].p = (uintptr_t)s;
        tab[idx].idx = idx;
        idx++;
    }
    qsort(tab, T, sizeof(ptr_idx), cmp_ptr_idx);
    iter = topOfLinkedList(m->tets);
    idx = 0;
    while ((s = nextElement(m->tets, &iter))) {
        for (int j = 0; j < 4; j++) {
            tet_v[idx * 4 + j] = vslot(h, s->p[j]);
            tet_n[idx * 4 + j] = s->s[j] ? lookup(tab, T, s->s[j]) : -1;
        }
        if (m->simplicies_kd[idx] != s) { free(tab); return -2; }
        idx++;
    }
    free(tab);

    struct kdtree *kd = m->kd;
    for (int i = 0; i < T; i++) {
        struct kdnode *nd = &kd->node_pool[i];
        kd_left[i] = nd->left ? (int)(nd->left - kd->node_pool) : -1;
        kd_right[i] = nd->right ? (int)(nd->right - kd->node_pool) : -1;
        for (int d = 0; d < 3; d++) kd_pos[i * 3 + d] = nd->pos[d];
        kd_data[i] = nd->data;
        if (nd->dir != 0 && nd->dir != 1 && nd->dir != 2) return -3;
    }
    if (T > 0 && kd->root != &kd->node_pool[0]) return -4;
    for (int d = 0; d < 3; d++) {
        kd_rect[d] = T > 0 ? kd->rect->min[d] : 0.0;
        kd_rect[3 + d] = T > 0 ? kd->rect->max[d] : 0.0;
    }
    return 0;
}

/* Reference query path (natural_insertionfree.c:761-940).  num_threads <= 0
 * selects the serial getInsertionFreeWeights, otherwise the OpenMP one.
 * *vals and *cols are malloc'd by the reference; release with ref_free_buf. */
int ref_get_weights(void *hp, const double *q, int M, int num_threads,
                    double **vals, int **cols, int *rowptr) {
    ref_handle *h = (ref_handle *)hp;
    if (num_threads <= 0)
        return getInsertionFreeWeights((double *)q, M, h->m, h->n, vals, cols, rowptr);
    return getInsertionFreeWeightsParallel((double *)q, M, h->m, num_threads, h->n,
                                           vals, cols, rowptr);
}

void ref_free_buf(void *p) { free(p); }

/* The libc generator the reference perturbs with, for pinning the builder's
 * private restatement of it. */
void ref_srand(unsigned s) { srand(s); }
int ref_rand(void) { return rand(); }
