// delaunay_build.h -- host side of the opt-in GPU Delaunay build.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

namespace nat {

struct GpuBuild {
    int n = 0;                      // data points
    int n_tets = 0;                 // tets of the tetrahedralisation (super tet's vertices included)
    int rounds = 0;                 // insertion rounds
    double seconds = 0.0;           // wall time of the device build
    std::vector<double> verts;      // [(n + 4) * 3] points, then the super tet's vertices
    std::vector<int32_t> tet_v;     // [n_tets * 4] positively oriented
    std::vector<int32_t> tet_n;     // [n_tets * 4] neighbour across face i, -1 = none
};

// Delaunay tetrahedralisation of points [n, 3] and the reference's super tet
// on the current device.  0 on success; -3 for degenerate input (a repeated
// point, a point on a cavity face plane, a flat bounding box), -4 for a
// capacity or convergence failure, -10 for a HIP error; *msg says which.
int gpu_build_tets(const double *points, int n, GpuBuild *out, std::string *msg);

// Exact audit of a tetrahedralisation on the current device: counts[0] faces
// with a neighbour tested, [1] tets of non-positive orientation, [2] broken
// neighbour links, [3] faces whose neighbour's far vertex is strictly inside
// the tet's circumsphere.
int gpu_delaunay_audit(const double *verts, int nv, const int32_t *tet_v, const int32_t *tet_n,
                       int T, long long counts[4], std::string *msg);

}  // namespace nat
