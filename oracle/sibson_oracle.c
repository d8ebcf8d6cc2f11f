/*
 * sibson_oracle.c -- CPU restatement of natinterp3d's insertion-free Sibson
 * query path.
 *
 * TEST INFRASTRUCTURE ONLY: this is the checker the HIP implementation is
 * compared against (tests/, __graft_entry__.smoke, bench.py's CPU-baseline
 * leg).  It is never linked into natinterp3d_amd.
 *
 * Each function names the reference lines it restates (paths relative to
 * src/natinterp3d/ of isarandi/natinterp3d).  Arithmetic follows the
 * reference's operation order exactly and is compiled with
 * -ffp-contract=off, so results are bit-identical to the reference on the
 * same packed mesh.  Set semantics replace the reference's pointer hash
 * tables (their iteration order never reaches the output).
 */
#include "sibson_oracle.h"

#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ---------------------------------------------------------------- predicates */

/* orient3dfast, predicates.c:1687-1710 */
double orc_orient3d(const double *pa, const double *pb, const double *pc, const double *pd) {
    double adx = pa[0] - pd[0], bdx = pb[0] - pd[0], cdx = pc[0] - pd[0];
    double ady = pa[1] - pd[1], bdy = pb[1] - pd[1], cdy = pc[1] - pd[1];
    double adz = pa[2] - pd[2], bdz = pb[2] - pd[2], cdz = pc[2] - pd[2];
    return adx * (bdy * cdz - bdz * cdy) + bdx * (cdy * adz - cdz * ady) +
           cdx * (ady * bdz - adz * bdy);
}

/* inspherefast, predicates.c:3301-3347 */
double orc_insphere(const double *pa, const double *pb, const double *pc, const double *pd,
                    const double *pe) {
    double aex = pa[0] - pe[0], bex = pb[0] - pe[0], cex = pc[0] - pe[0], dex = pd[0] - pe[0];
    double aey = pa[1] - pe[1], bey = pb[1] - pe[1], cey = pc[1] - pe[1], dey = pd[1] - pe[1];
    double aez = pa[2] - pe[2], bez = pb[2] - pe[2], cez = pc[2] - pe[2], dez = pd[2] - pe[2];
    double ab = aex * bey - bex * aey;
    double bc = bex * cey - cex * bey;
    double cd = cex * dey - dex * cey;
    double da = dex * aey - aex * dey;
    double ac = aex * cey - cex * aey;
    double bd = bex * dey - dex * bey;
    double abc = aez * bc - bez * ac + cez * ab;
    double bcd = bez * cd - cez * bd + dez * bc;
    double cda = cez * da + dez * ac + aez * cd;
    double dab = dez * ab + aez * bd + bez * da;
    double alift = aex * aex + aey * aey + aez * aez;
    double blift = bex * bex + bey * bey + bez * bez;
    double clift = cex * cex + cey * cey + cez * cez;
    double dlift = dex * dex + dey * dey + dez * dez;
    return (dlift * abc - clift * dab) + (blift * cda - alift * bcd);
}

/* ------------------------------------------------ vector helpers, delaunay.c:980-1020 */

static inline void vsub(const double *a, const double *b, double *o) {
    o[0] = a[0] - b[0]; o[1] = a[1] - b[1]; o[2] = a[2] - b[2];
}
static inline void vcross(const double *b, const double *c, double *o) {
    o[0] = b[1] * c[2] - b[2] * c[1];
    o[1] = b[2] * c[0] - b[0] * c[2];
    o[2] = b[0] * c[1] - b[1] * c[0];
}
static inline double vdot(const double *a, const double *b) {
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}

/* circumCenterTri3D, natural_insertionfree.c:24-49 */
void orc_cc_tri(const double *a, const double *b, const double *c, double *out) {
    double ab[3], ac[3], n[3];
    vsub(b, a, ab);
    vsub(c, a, ac);
    vcross(ab, ac, n);
    double n2 = vdot(n, n);
    if (n2 < 1e-30) {
        out[0] = (a[0] + b[0] + c[0]) / 3.0;
        out[1] = (a[1] + b[1] + c[1]) / 3.0;
        out[2] = (a[2] + b[2] + c[2]) / 3.0;
        return;
    }
    double ab2 = vdot(ab, ab), ac2 = vdot(ac, ac);
    double t[3], cr[3];
    t[0] = ab2 * ac[0] - ac2 * ab[0];
    t[1] = ab2 * ac[1] - ac2 * ab[1];
    t[2] = ab2 * ac[2] - ac2 * ab[2];
    vcross(t, n, cr);
    double inv = 1.0 / (2.0 * n2);
    out[0] = a[0] + cr[0] * inv;
    out[1] = a[1] + cr[1] * inv;
    out[2] = a[2] + cr[2] * inv;
}

/* circumCenterTet, natural_insertionfree.c:54-84 */
void orc_cc_tet(const double *a, const double *b, const double *c, const double *d, double *out) {
    double ba[3], ca[3], da[3], x1[3], x2[3], x3[3], s[3];
    vsub(b, a, ba);
    vsub(c, a, ca);
    vsub(d, a, da);
    vcross(ba, ca, x1);
    vcross(da, ba, x2);
    vcross(ca, da, x3);
    double l1 = vdot(da, da), l2 = vdot(ca, ca), l3 = vdot(ba, ba);
    for (int i = 0; i < 3; i++) {
        double m1 = x1[i] * l1, m2 = x2[i] * l2, m3 = x3[i] * l3;
        double t = m1 + m2;
        s[i] = m3 + t;
    }
    double denom = 2.0 * vdot(ba, x3);
    if (fabs(denom) < 1e-30) {
        for (int i = 0; i < 3; i++) out[i] = (a[i] + b[i] + c[i] + d[i]) * 0.25;
        return;
    }
    double inv = 1.0 / denom;
    for (int i = 0; i < 3; i++) {
        double t = s[i] * inv;
        out[i] = t + a[i];
    }
}

/* voronoiSubcellVolume, natural_insertionfree.c:92-146 */
double orc_subcell(const double *pk, const double *pa, const double *pb, const double *pc,
                   const double *c_tet, const double *c_fab, const double *c_fac,
                   const double *c_fbc) {
    double ta[3], tb[3], tc[3];
    vsub(pa, pk, ta);
    vsub(pb, pk, tb);
    vsub(pc, pk, tc);
    double ox = tb[1] * tc[2] - tb[2] * tc[1];
    double oy = tb[2] * tc[0] - tb[0] * tc[2];
    double oz = tb[0] * tc[1] - tb[1] * tc[0];
    double sv = ta[0] * ox + ta[1] * oy + ta[2] * oz;
    const double *a = pa, *b = pb, *fab = c_fab, *fac = c_fac, *fbc = c_fbc;
    if (sv < 0) {
        a = pb; b = pa;
        fac = c_fbc; fbc = c_fac;
    }
    double df[3], dg[3], dh[3], dt[3];
    vsub(fab, pk, df);
    vsub(fac, pk, dg);
    vsub(fbc, pk, dh);
    vsub(c_tet, pk, dt);
    double C1[3], C2[3], C3[3];
    vcross(df, dt, C1);
    vcross(dt, dg, C2);
    vcross(dh, dt, C3);
    double dab[3], dac[3], dbc[3];
    vsub(a, b, dab);
    vsub(a, pc, dac);
    vsub(b, pc, dbc);
    double vol = dab[0] * C1[0] + dab[1] * C1[1] + dab[2] * C1[2]
               + dac[0] * C2[0] + dac[1] * C2[1] + dac[2] * C2[2]
               + dbc[0] * C3[0] + dbc[1] * C3[1] + dbc[2] * C3[2];
    return vol / 12.0;
}

/* ----------------------------------------------------------------- k-d tree */

/* hyperrect_dist_sq, kdtree.c:800-814 */
static double rect_dist(const double *rmin, const double *rmax, const double *pos) {
    double r = 0;
    for (int i = 0; i < 3; i++) {
        if (pos[i] < rmin[i]) {
            double t = rmin[i] - pos[i];
            r += t * t;
        } else if (pos[i] > rmax[i]) {
            double t = rmax[i] - pos[i];
            r += t * t;
        }
    }
    return r;
}

/* kd_nearest_i, kdtree.c:377-434; split dimension = depth mod 3 (insert_rec,
 * kdtree.c:213). */
static void kd_near(const orc_mesh *m, int node, int dir, const double *pos, int *best,
                    double *best_d, double *rmin, double *rmax, long long *count) {
    const double *np = &m->kd_pos[node * 3];
    (*count)++;
    double dummy = pos[dir] - np[dir];
    int nearer, farther;
    double *nc, *fc;
    if (dummy <= 0) {
        nearer = m->kd_left[node]; farther = m->kd_right[node];
        nc = &rmax[dir]; fc = &rmin[dir];
    } else {
        nearer = m->kd_right[node]; farther = m->kd_left[node];
        nc = &rmin[dir]; fc = &rmax[dir];
    }
    if (nearer >= 0) {
        double sv = *nc;
        *nc = np[dir];
        kd_near(m, nearer, (dir + 1) % 3, pos, best, best_d, rmin, rmax, count);
        *nc = sv;
    }
    double d = 0;
    for (int i = 0; i < 3; i++) {
        double t = np[i] - pos[i];
        d += t * t;
    }
    if (d < *best_d) {
        *best = node;
        *best_d = d;
    }
    if (farther >= 0) {
        double sv = *fc;
        *fc = np[dir];
        if (rect_dist(rmin, rmax, pos) < *best_d)
            kd_near(m, farther, (dir + 1) % 3, pos, best, best_d, rmin, rmax, count);
        *fc = sv;
    }
}

/* kd_nearest3_data, kdtree.c:543-564 */
static int kd_nearest(const orc_mesh *m, const double *pos, long long *count) {
    if (m->n_tets <= 0) return -1;
    double rmin[3] = {m->kd_rect[0], m->kd_rect[1], m->kd_rect[2]};
    double rmax[3] = {m->kd_rect[3], m->kd_rect[4], m->kd_rect[5]};
    const double *r = &m->kd_pos[0];
    double dx = r[0] - pos[0], dy = r[1] - pos[1], dz = r[2] - pos[2];
    double best_d = dx * dx + dy * dy + dz * dz;
    int best = 0;
    kd_near(m, 0, 0, pos, &best, &best_d, rmin, rmax, count);
    return best;
}

/* getFaceVerticies3, delaunay.c:444-467: face i of a tet, shared with s[i]. */
static const int FACE3[4][3] = {{0, 1, 2}, {3, 1, 0}, {0, 2, 3}, {3, 2, 1}};

static inline const double *V(const orc_mesh *m, int slot) { return &m->verts[slot * 3]; }

/* findContainingSimplex, delaunay.c:383-407.  The tets the walk reads are
 * recorded in path[0..cap) (statistics only). */
static int locate(const orc_mesh *m, const double *q, long long *kd_count, long long *tests,
                  int *path, int *npath, int cap) {
    int s = kd_nearest(m, q, kd_count);
    if (s < 0) return -1;
    if (path && *npath < cap) path[(*npath)++] = s;
    /* A Delaunay mesh never cycles; the bound only stops a corrupted mesh
     * (on which the reference would loop forever). */
    long long steps = 0, limit = 4LL * m->n_tets + 1024;
    for (int i = 0; i < 4; i++) {
        const int *tv = &m->tet_v[s * 4];
        (*tests)++;
        if (orc_orient3d(V(m, tv[FACE3[i][0]]), V(m, tv[FACE3[i][1]]), V(m, tv[FACE3[i][2]]), q) < 0) {
            int nb = m->tet_n[s * 4 + i];
            if (nb < 0) return -1;
            s = nb;
            i = -1;
            if (path && *npath < cap) path[(*npath)++] = s;
            if (++steps > limit) return -1;
        }
    }
    return s;
}

/* ------------------------------------------------------------ per-thread state */

typedef struct {
    int *vis;          /* [T] stamp: tet popped by the cavity search */
    int *cav;          /* [T] stamp: tet in the cavity */
    int *vstamp;       /* [N+4] stamp: vertex counted */
    int *rstamp;       /* [T] stamp: tet read (statistics) */
    int *rvstamp;      /* [N+4] stamp: vertex of a read tet (statistics) */
    int *nbr_map;      /* [N] vertex -> slot, -1 (natural_insertionfree.c:285-287) */
    int stamp;
    int *stack; int stack_cap;
    int *cavity; int cav_cap;
    int *bnd; int bnd_cap;           /* 3 vertex slots per boundary face */
    int *nbr; double *stolen; int nbr_cap;
    int *fkey; double *fcc; int *fgen; int fcap; int fgen_cur;
    /* result arena (natural_insertionfree.c:744-756) */
    double *av; int *ai; long long aused, acap;
    orc_counters c;
} qstate;

static void *grow(void *p, int *cap, int need, size_t elt) {
    if (need <= *cap) return p;
    int nc = *cap ? *cap : 64;
    while (nc < need) nc *= 2;
    *cap = nc;
    return realloc(p, (size_t)nc * elt);
}

static void qstate_init(qstate *s, const orc_mesh *m) {
    memset(s, 0, sizeof(*s));
    s->vis = (int *)calloc((size_t)m->n_tets + 1, sizeof(int));
    s->cav = (int *)calloc((size_t)m->n_tets + 1, sizeof(int));
    s->vstamp = (int *)calloc((size_t)m->n_data + 4, sizeof(int));
    s->rstamp = (int *)calloc((size_t)m->n_tets + 1, sizeof(int));
    s->rvstamp = (int *)calloc((size_t)m->n_data + 4, sizeof(int));
    s->nbr_map = (int *)malloc(sizeof(int) * ((size_t)m->n_data + 1));
    memset(s->nbr_map, 0xff, sizeof(int) * ((size_t)m->n_data + 1));
    s->fcap = 256;
    s->fkey = (int *)malloc(sizeof(int) * 3 * s->fcap);
    s->fcc = (double *)malloc(sizeof(double) * 3 * s->fcap);
    s->fgen = (int *)calloc(s->fcap, sizeof(int));
    s->acap = 4096;
    s->av = (double *)malloc(sizeof(double) * s->acap);
    s->ai = (int *)malloc(sizeof(int) * s->acap);
}

static void qstate_free(qstate *s) {
    free(s->vis); free(s->cav); free(s->vstamp); free(s->nbr_map);
    free(s->rstamp); free(s->rvstamp);
    free(s->stack); free(s->cavity); free(s->bnd); free(s->nbr); free(s->stolen);
    free(s->fkey); free(s->fcc); free(s->fgen);
    free(s->av); free(s->ai);
}

/* addOrGetNeighbor, natural_insertionfree.c:373-382 */
static int add_nbr(qstate *s, int vidx, int *count) {
    int slot = s->nbr_map[vidx];
    if (slot >= 0) return slot;
    slot = (*count)++;
    int cap = s->nbr_cap;
    s->nbr = (int *)grow(s->nbr, &cap, *count, sizeof(int));
    cap = s->nbr_cap;
    s->stolen = (double *)grow(s->stolen, &cap, *count, sizeof(double));
    s->nbr_cap = cap;
    s->nbr[slot] = vidx;
    s->stolen[slot] = 0.0;
    s->nbr_map[vidx] = slot;
    return slot;
}

/* faceCCLookup, natural_insertionfree.c:233-253: keyed by the vertex set, the
 * value computed with the FIRST caller's vertex order. */
static const double *face_cc(qstate *s, const orc_mesh *m, int va, int vb, int vc) {
    int k0 = va, k1 = vb, k2 = vc, t;
    if (k0 > k1) { t = k0; k0 = k1; k1 = t; }
    if (k1 > k2) { t = k1; k1 = k2; k2 = t; }
    if (k0 > k1) { t = k0; k0 = k1; k1 = t; }
    unsigned h = (unsigned)((uint64_t)k0 * 0x9e3779b97f4a7c15ULL ^ (uint64_t)k1 * 0x517cc1b727220a95ULL ^
                            (uint64_t)k2 * 0x6c62272e07bb0142ULL);
    int mask = s->fcap - 1;
    h &= mask;
    for (;;) {
        if (s->fgen[h] != s->fgen_cur) {
            orc_cc_tri(V(m, va), V(m, vb), V(m, vc), &s->fcc[h * 3]);
            s->fkey[h * 3] = k0; s->fkey[h * 3 + 1] = k1; s->fkey[h * 3 + 2] = k2;
            s->fgen[h] = s->fgen_cur;
            s->c.face_cc++;
            return &s->fcc[h * 3];
        }
        if (s->fkey[h * 3] == k0 && s->fkey[h * 3 + 1] == k1 && s->fkey[h * 3 + 2] == k2)
            return &s->fcc[h * 3];
        h = (h + 1) & mask;
    }
}

/* statistics: a tet record and its vertices read by this query */
static void note_read(qstate *s, const orc_mesh *m, int t, int st) {
    if (s->rstamp[t] == st) return;
    s->rstamp[t] = st;
    s->c.tets_read++;
    for (int j = 0; j < 4; j++) {
        int p = m->tet_v[t * 4 + j];
        if (s->rvstamp[p] != st) { s->rvstamp[p] = st; s->c.verts_read++; }
    }
}

/* getWeightsSingleQueryIF, natural_insertionfree.c:591-738.  Appends the row
 * to the thread's arena and returns its length. */
static int one_query(qstate *s, const orc_mesh *m, const double *q) {
    const int N = m->n_data;
    s->stamp++;
    const int st = s->stamp;
    int ncav = 0, top = 0, nb_count = 0, nbnd = 0;
    long long visited = 0;
    s->c.queries++;

    /* findCavity, natural_insertionfree.c:486-551 */
    int path[64], npath = 0;
    int start = locate(m, q, &s->c.kd_nodes, &s->c.walk_tests, path, &npath, 64);
    for (int i = 0; i < npath; i++) note_read(s, m, path[i], st);
    if (start >= 0) {
        s->stack = (int *)grow(s->stack, &s->stack_cap, 1, sizeof(int));
        s->stack[top++] = start;
        while (top > 0) {
            int cur = s->stack[--top];
            if (s->vis[cur] == st) continue;
            s->vis[cur] = st;
            visited++;
            note_read(s, m, cur, st);
            const int *tv = &m->tet_v[cur * 4];
            /* orientation bit, precomputeOrientSigns (natural_insertionfree.c:457-476) */
            int pos = orc_orient3d(V(m, tv[0]), V(m, tv[1]), V(m, tv[2]), V(m, tv[3])) > 0;
            double ins = pos ? orc_insphere(V(m, tv[0]), V(m, tv[1]), V(m, tv[2]), V(m, tv[3]), q)
                             : orc_insphere(V(m, tv[1]), V(m, tv[0]), V(m, tv[2]), V(m, tv[3]), q);
            s->c.insphere++;
            if (ins > 0) {
                s->cavity = (int *)grow(s->cavity, &s->cav_cap, ncav + 1, sizeof(int));
                s->cavity[ncav++] = cur;
                s->cav[cur] = st;
                for (int i = 0; i < 4; i++) {
                    int nb = m->tet_n[cur * 4 + i];
                    if (nb >= 0 && s->vis[nb] != st) {
                        s->stack = (int *)grow(s->stack, &s->stack_cap, top + 1, sizeof(int));
                        s->stack[top++] = nb;
                        if (top > s->c.max_stack) s->c.max_stack = top;
                    }
                }
            }
        }
    }
    if (visited > s->c.max_visited) s->c.max_visited = visited;
    if (ncav > s->c.max_cavity) s->c.max_cavity = ncav;
    s->c.cavity += ncav;

    int *row_idx = NULL;
    double *row_val = NULL;

    if (ncav == 0) {
        /* coincident-vertex path, natural_insertionfree.c:596-622 */
        long long dummy_k = 0, dummy_t = 0;
        int t0 = locate(m, q, &dummy_k, &dummy_t, NULL, NULL, 0);
        if (t0 >= 0) {
            int coin[4], nco = 0;
            for (int i = 0; i < 4; i++) {
                int p = m->tet_v[t0 * 4 + i];
                const double *pv = V(m, p);
                if (p < N && pv[0] == q[0] && pv[1] == q[1] && pv[2] == q[2]) coin[nco++] = p;
            }
            if (nco > 0) {
                double w = 1.0 / nco;
                for (int i = 0; i < nco; i++) {
                    int slot = add_nbr(s, coin[i], &nb_count);
                    s->stolen[slot] = w;
                }
                s->c.coincident++;
            }
        }
        goto emit;
    }

    /* extractBoundaryAndNeighbors, natural_insertionfree.c:559-584 */
    for (int ci = 0; ci < ncav; ci++) {
        int t = s->cavity[ci];
        for (int fi = 0; fi < 4; fi++) {
            int nb = m->tet_n[t * 4 + fi];
            if (nb < 0 || s->cav[nb] != st) {
                const int *tv = &m->tet_v[t * 4];
                int v1 = tv[FACE3[fi][0]], v2 = tv[FACE3[fi][1]], v3 = tv[FACE3[fi][2]];
                s->bnd = (int *)grow(s->bnd, &s->bnd_cap, 3 * (nbnd + 1), sizeof(int));
                s->bnd[3 * nbnd] = v1; s->bnd[3 * nbnd + 1] = v2; s->bnd[3 * nbnd + 2] = v3;
                nbnd++;
                if (v1 < N) add_nbr(s, v1, &nb_count);
                if (v2 < N) add_nbr(s, v2, &nb_count);
                if (v3 < N) add_nbr(s, v3, &nb_count);
            }
        }
    }
    s->c.boundary += nbnd;
    if (nbnd > s->c.max_boundary) s->c.max_boundary = nbnd;
    if (nb_count == 0) goto emit;

    /* distinct vertices read (for the byte count only) */
    for (int ci = 0; ci < ncav; ci++)
        for (int j = 0; j < 4; j++) {
            int p = m->tet_v[s->cavity[ci] * 4 + j];
            if (s->vstamp[p] != st) { s->vstamp[p] = st; s->c.dist_verts++; }
        }

    /* face-CC cache sized for 4 faces per cavity tet at <75% load (:628-636) */
    {
        int need = ncav * 4;
        if (need * 4 >= s->fcap * 3) {
            while (need * 4 >= s->fcap * 3) s->fcap *= 2;
            free(s->fkey); free(s->fcc); free(s->fgen);
            s->fkey = (int *)malloc(sizeof(int) * 3 * s->fcap);
            s->fcc = (double *)malloc(sizeof(double) * 3 * s->fcap);
            s->fgen = (int *)calloc(s->fcap, sizeof(int));
        }
        s->fgen_cur++;
        if (s->fgen_cur == 0) s->fgen_cur = 1;
    }

    /* Step 3: old contributions, natural_insertionfree.c:642-676 */
    for (int ci = 0; ci < ncav; ci++) {
        const int *tv = &m->tet_v[s->cavity[ci] * 4];
        double ct[3];
        orc_cc_tet(V(m, tv[0]), V(m, tv[1]), V(m, tv[2]), V(m, tv[3]), ct);
        const double *fcc[4];
        fcc[0] = face_cc(s, m, tv[1], tv[2], tv[3]);
        fcc[1] = face_cc(s, m, tv[0], tv[2], tv[3]);
        fcc[2] = face_cc(s, m, tv[0], tv[1], tv[3]);
        fcc[3] = face_cc(s, m, tv[0], tv[1], tv[2]);
        for (int vi = 0; vi < 4; vi++) {
            int pk = tv[vi];
            if (pk >= N) continue;
            int slot = s->nbr_map[pk];
            if (slot < 0) continue;
            int o0 = (vi + 1) & 3, o1 = (vi + 2) & 3, o2 = (vi + 3) & 3, t;
            if (o2 < o0) { t = o0; o0 = o2; o2 = t; }
            if (o1 > o2) { t = o1; o1 = o2; o2 = t; }
            if (o0 > o1) { t = o0; o0 = o1; o1 = t; }
            double vol = orc_subcell(V(m, pk), V(m, tv[o0]), V(m, tv[o1]), V(m, tv[o2]), ct,
                                     fcc[o2], fcc[o1], fcc[o0]);
            s->stolen[slot] += vol;
            s->c.sub3++;
        }
    }

    /* Step 4: new contributions, natural_insertionfree.c:681-719 */
    for (int bi = 0; bi < nbnd; bi++) {
        int fv[3] = {s->bnd[3 * bi], s->bnd[3 * bi + 1], s->bnd[3 * bi + 2]};
        double cv[3];
        orc_cc_tet(q, V(m, fv[0]), V(m, fv[1]), V(m, fv[2]), cv);
        const double *cf = face_cc(s, m, fv[0], fv[1], fv[2]);
        for (int k = 0; k < 3; k++) {
            int pk = fv[k];
            if (pk >= N) continue;
            int slot = s->nbr_map[pk];
            if (slot < 0) continue;
            int pa = fv[(k + 1) % 3], pb = fv[(k + 2) % 3];
            double ca[3], cb[3];
            orc_cc_tri(V(m, pk), q, V(m, pa), ca);
            orc_cc_tri(V(m, pk), q, V(m, pb), cb);
            double vol = orc_subcell(V(m, pk), q, V(m, pa), V(m, pb), cv, ca, cb, cf);
            s->stolen[slot] -= vol;
            s->c.sub4++;
        }
    }

    /* Step 5: normalise, natural_insertionfree.c:721-737 */
    {
        double total = 0.0;
        for (int i = 0; i < nb_count; i++) total += s->stolen[i];
        if (total > 1e-30) {
            double inv = 1.0 / total;
            for (int i = 0; i < nb_count; i++) s->stolen[i] *= inv;
        } else {
            for (int i = 0; i < nb_count; i++) s->nbr_map[s->nbr[i]] = -1;
            nb_count = 0;
        }
    }

emit:
    row_idx = s->nbr;
    row_val = s->stolen;
    if (s->aused + nb_count > s->acap) {
        while (s->aused + nb_count > s->acap) s->acap *= 2;
        s->av = (double *)realloc(s->av, sizeof(double) * s->acap);
        s->ai = (int *)realloc(s->ai, sizeof(int) * s->acap);
    }
    for (int i = 0; i < nb_count; i++) {
        s->av[s->aused + i] = row_val[i];
        s->ai[s->aused + i] = row_idx[i];
        s->nbr_map[row_idx[i]] = -1;
    }
    s->aused += nb_count;
    s->c.nnz += nb_count;
    if (nb_count > s->c.max_nnz) s->c.max_nnz = nb_count;
    if (nb_count == 0) s->c.empty_rows++;
    return nb_count;
}

static void add_counters(orc_counters *a, const orc_counters *b) {
    a->queries += b->queries; a->kd_nodes += b->kd_nodes; a->walk_tests += b->walk_tests;
    a->insphere += b->insphere; a->cavity += b->cavity; a->boundary += b->boundary;
    a->face_cc += b->face_cc; a->sub3 += b->sub3; a->sub4 += b->sub4; a->nnz += b->nnz;
    a->dist_verts += b->dist_verts;
    a->tets_read += b->tets_read;
    a->verts_read += b->verts_read;
    a->coincident += b->coincident; a->empty_rows += b->empty_rows;
    if (b->max_cavity > a->max_cavity) a->max_cavity = b->max_cavity;
    if (b->max_visited > a->max_visited) a->max_visited = b->max_visited;
    if (b->max_stack > a->max_stack) a->max_stack = b->max_stack;
    if (b->max_boundary > a->max_boundary) a->max_boundary = b->max_boundary;
    if (b->max_nnz > a->max_nnz) a->max_nnz = b->max_nnz;
}

/* getInsertionFreeWeights[Parallel], natural_insertionfree.c:761-940.  Query
 * order does not change any row (each row depends only on its query), so the
 * loop runs in input order. */
int orc_get_weights(const orc_mesh *m, const double *q, int M, int num_threads, double **vals,
                    int **cols, int *rowptr, orc_counters *counters) {
    int T = num_threads > 0 ? num_threads : 1;
    qstate *st = (qstate *)malloc(sizeof(qstate) * T);
    for (int t = 0; t < T; t++) qstate_init(&st[t], m);
    int *qt = (int *)malloc(sizeof(int) * (M > 0 ? M : 1));
    long long *qo = (long long *)malloc(sizeof(long long) * (M > 0 ? M : 1));
    int *qc = (int *)malloc(sizeof(int) * (M > 0 ? M : 1));

#pragma omp parallel for schedule(dynamic, 64) num_threads(T)
    for (int i = 0; i < M; i++) {
#ifdef _OPENMP
        int tid = omp_get_thread_num();
#else
        int tid = 0;
#endif
        qstate *s = &st[tid];
        long long off = s->aused;
        int c = one_query(s, m, &q[(size_t)i * 3]);
        qt[i] = tid; qo[i] = off; qc[i] = c;
    }

    int rc = 0;
    long long total = 0;
    for (int i = 0; i < M; i++) {
        rowptr[i] = (int)total;
        total += qc[i];
    }
    if (total > 2147483647LL) {
        rc = -1;
    } else {
        rowptr[M] = (int)total;
        *vals = (double *)malloc(sizeof(double) * (total > 0 ? total : 1));
        *cols = (int *)malloc(sizeof(int) * (total > 0 ? total : 1));
        for (int i = 0; i < M; i++) {
            memcpy(&(*vals)[rowptr[i]], &st[qt[i]].av[qo[i]], sizeof(double) * qc[i]);
            memcpy(&(*cols)[rowptr[i]], &st[qt[i]].ai[qo[i]], sizeof(int) * qc[i]);
        }
    }
    if (counters) {
        memset(counters, 0, sizeof(*counters));
        for (int t = 0; t < T; t++) add_counters(counters, &st[t].c);
    }
    for (int t = 0; t < T; t++) qstate_free(&st[t]);
    free(st); free(qt); free(qo); free(qc);
    return rc;
}

void orc_locate(const orc_mesh *m, const double *q, int M, int *start_tet) {
    for (int i = 0; i < M; i++) {
        long long a = 0, b = 0;
        start_tet[i] = locate(m, &q[(size_t)i * 3], &a, &b, NULL, NULL, 0);
    }
}

int orc_cavity(const orc_mesh *m, const double *q, int *out, int cap) {
    qstate s;
    qstate_init(&s, m);
    one_query(&s, m, q);
    /* one_query leaves the cavity list of the last query in s.cavity */
    int n = 0;
    for (int i = 0; i < s.cav_cap && i < (int)s.c.cavity; i++) {
        if (n < cap) out[n] = s.cavity[i];
        n++;
    }
    qstate_free(&s);
    return n;
}

void orc_free(void *p) { free(p); }
