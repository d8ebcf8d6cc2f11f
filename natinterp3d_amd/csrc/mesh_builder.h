// mesh_builder.h -- host Bowyer-Watson tetrahedralisation that reproduces the
// reference's mesh exactly (tet order, vertex order inside each tet,
// neighbour slots, k-d tree over tet centroids, degeneracy perturbations).
#pragma once
#include <cstdint>
#include <vector>

namespace nat {

struct PackedMesh {
    int n_data = 0;                 // N data points; super vertices are slots N..N+3
    int n_tets = 0;                 // T
    std::vector<double> verts;      // [(N+4)*3] final (possibly perturbed) coordinates
    std::vector<int32_t> tet_v;     // [T*4] vertex slots, p[0..3]
    std::vector<int32_t> tet_n;     // [T*4] neighbour across face i, -1 = none
    std::vector<int32_t> kd_left;   // [T] k-d children (node i = tet i), -1 = none
    std::vector<int32_t> kd_right;  // [T]
    std::vector<double> kd_pos;     // [T*3] tet centroids
    double kd_rect[6] = {0, 0, 0, 0, 0, 0};  // min[3], max[3] of the centroids
    int kd_depth = 0;               // deepest node (root = 1)
    int coplanar = 0;               // coplanar degeneracies perturbed away
    int cospherical = 0;            // cospherical degeneracies perturbed away
    int edge_anomalies = 0;         // non-manifold edge groups met while linking
};

// Returns 0 on success; -2 when a tet of non-positive orientation is met
// (the reference aborts the process there), -3 when a point lies outside the
// super simplex, -1 for invalid arguments.
int build_packed_mesh(const double *points, int n, PackedMesh *out);

// The reference's super simplex for points [n, 3] (delaunay.c:832-863 over
// the bounding box of getRange, delaunay.c:1086-1111): out[12] = its four
// vertices, which become slots n .. n + 3.
void super_vertices(const double *points, int n, double *out);

// A mesh from another builder (same vertex slots, positively oriented tets,
// neighbour across face i as the reference's face order): tets renumbered in
// the preorder of a balanced k-d tree over their centroids (median splits,
// split axis = depth mod 3, node i = tet i, root 0), k-d arrays filled.  The
// reference's tree is grown in tet-creation order; any tree over the same
// centroids gives the query path an exact nearest centroid.  Tets whose
// first vertex is -1 are freed slots and are dropped.
void pack_balanced(int n, std::vector<double> verts, const std::vector<int32_t> &tet_v,
                   const std::vector<int32_t> &tet_n, int threads, PackedMesh *out);

}  // namespace nat
