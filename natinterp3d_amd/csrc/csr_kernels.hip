// csr_kernels.hip -- gfx950 kernels around the row computation:
//   morton_*        spatial bucketing of the queries (order only; no result
//                   depends on it)                            natural.c:18-96
//   scan/gather     CSR assembly (count, scan, fill)   natural_insertionfree.c:895-928
//   sort_rows       row canonicalisation done by scipy's astype   natinterp3d.py:117
//   spmm            W @ values in stored order                 natinterp3d.py:131-137
#include <hip/hip_runtime.h>

#include "nat_internal.h"

namespace nat {

// --------------------------------------------------------- Morton bucketing

__device__ __forceinline__ unsigned spread3(unsigned v) {
    v &= 0x3ff;
    v = (v | (v << 16)) & 0x030000ff;
    v = (v | (v << 8)) & 0x0300f00f;
    v = (v | (v << 4)) & 0x030c30c3;
    v = (v | (v << 2)) & 0x09249249;
    return v;
}

// Morton key of each query: problem-major, then the 3*bits-bit Morton code of
// its cell in the problem's data bounding box.
__global__ void morton_key_kernel(const double *__restrict__ q, int M,
                                  const LocParams *__restrict__ lp, const int *__restrict__ prob,
                                  int bits, int *__restrict__ keys) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= M) return;
    const int p = prob ? prob[i] : 0;
    const LocParams &L = lp[p];
    const int lim = (1 << bits) - 1;
    auto cell = [&](double v, int d) {
        const double r = L.bbox[3 + d] - L.bbox[d];
        const double s = r > 0 ? static_cast<double>(1 << bits) / r : 1.0;
        double c = (v - L.bbox[d]) * s;
        int k = (c >= 0.0 && c < 2147483647.0) ? static_cast<int>(c) : 0;  // NaN -> 0
        if (c >= 2147483647.0) k = lim;
        return k > lim ? lim : k;
    };
    const unsigned ix = cell(q[3 * static_cast<size_t>(i)], 0);
    const unsigned iy = cell(q[3 * static_cast<size_t>(i) + 1], 1);
    const unsigned iz = cell(q[3 * static_cast<size_t>(i) + 2], 2);
    keys[i] = (p << (3 * bits)) |
              static_cast<int>(spread3(ix) | (spread3(iy) << 1) | (spread3(iz) << 2));
}

// The query order is the stable sort of the keys (equal keys in index order):
// an LSD radix sort, 8 bits per pass.  Each pass counts the digits of every
// tile of kSortTile consecutive items, scans the counts digit-major (digit d
// of tile t lands after every smaller digit and after digit d of the earlier
// tiles), and scatters each tile in index order.  Every count is written, no
// buffer needs clearing, and the order is a function of the keys alone: the
// same queries always form the same runs of the order.
constexpr int kSortTile = 1024;  // one 64-lane wave, 16 rounds

__global__ void __launch_bounds__(64) radix_hist_kernel(const int *__restrict__ keys, int M,
                                                        int shift, int ntiles,
                                                        int *__restrict__ hist) {
    __shared__ int h[256];
    const int lane = threadIdx.x;
    const int tile = blockIdx.x;
    for (int d = lane; d < 256; d += 64) h[d] = 0;
    __syncthreads();
    const int base = tile * kSortTile;
    for (int r = 0; r < kSortTile / 64; r++) {
        const int i = base + r * 64 + lane;
        if (i < M) atomicAdd(&h[(keys[i] >> shift) & 255], 1);
    }
    __syncthreads();
    for (int d = lane; d < 256; d += 64) hist[d * ntiles + tile] = h[d];
}

// vals_in NULL: the values are the item indices
__global__ void __launch_bounds__(64) radix_scatter_kernel(
    const int *__restrict__ keys_in, const int *__restrict__ vals_in, int M, int shift, int ntiles,
    const long long *__restrict__ offs, int *__restrict__ keys_out, int *__restrict__ vals_out) {
    __shared__ int base[256];
    const int lane = threadIdx.x;
    const int tile = blockIdx.x;
    for (int d = lane; d < 256; d += 64) base[d] = static_cast<int>(offs[d * ntiles + tile]);
    __syncthreads();
    const unsigned long long lt = (1ull << lane) - 1ull;
    for (int r = 0; r < kSortTile / 64; r++) {
        const int i = tile * kSortTile + r * 64 + lane;
        const bool valid = i < M;
        const int k = valid ? keys_in[i] : 0;
        const int v = valid ? (vals_in ? vals_in[i] : i) : 0;
        const unsigned d = (static_cast<unsigned>(k) >> shift) & 255u;
        // the lanes holding the same digit, by its 8 bits
        unsigned long long same = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; b++) {
            const unsigned long long bal = __ballot((d >> b) & 1u);
            same &= ((d >> b) & 1u) ? bal : ~bal;
        }
        const int rank = __popcll(same & lt);
        const int pos = base[d] + rank;  // every lane reads before the leader writes
        if (valid && rank == 0) base[d] += __popcll(same);
        if (valid) {
            keys_out[pos] = k;
            vals_out[pos] = v;
        }
    }
}

__global__ void iota_kernel(int *__restrict__ out, int M) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < M) out[i] = i;
}

__global__ void zero_u32_kernel(unsigned *__restrict__ p, size_t n) {
    const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0u;
}

hipError_t launch_zero_u32(void *d_p, size_t n, hipStream_t st) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(zero_u32_kernel, dim3(static_cast<unsigned>((n + 255) / 256)), dim3(256), 0,
                       st, static_cast<unsigned *>(d_p), n);
    return hipGetLastError();
}

namespace {
size_t rnd256(size_t b) { return (b + 255) & ~static_cast<size_t>(255); }
}  // namespace

// Error statement of mixed rows, step 1: the sampled queries (every
// stride-th of the processing order) and their mixed rows' pool offset,
// length and total, kept aside while the FP64 pass rewrites them.
__global__ void mixed_check_save_kernel(const int *__restrict__ order, int n, int stride,
                                        const long long *__restrict__ row_off,
                                        const int *__restrict__ row_cnt,
                                        const double *__restrict__ row_total,
                                        int *__restrict__ items, long long *__restrict__ s_off,
                                        int *__restrict__ s_cnt, double *__restrict__ s_total) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int qi = order[static_cast<size_t>(i) * stride];
    items[i] = qi;
    s_off[i] = row_off[qi];
    s_cnt[i] = row_cnt[qi];
    s_total[i] = row_total[qi];
}

// Step 2, after the FP64 pass over items: per sampled row with a record, the
// largest |w_mixed - w_fp64| (columns must agree; a differing column or
// length counts as an infinite error), then the mixed row's pool offset,
// length and total are put back.  out[0] = max error (bits of a
// non-negative double, so unsigned order is numeric order), out[1] = rows,
// out[2] = entries.
__global__ void mixed_check_compare_kernel(int n, const int *__restrict__ items,
                                           const long long *__restrict__ rec_off,
                                           long long *__restrict__ row_off,
                                           int *__restrict__ row_cnt,
                                           double *__restrict__ row_total,
                                           const long long *__restrict__ s_off,
                                           const int *__restrict__ s_cnt,
                                           const double *__restrict__ s_total,
                                           const double *__restrict__ pool_vals,
                                           const int *__restrict__ pool_cols, long long pool_cap,
                                           unsigned long long *out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int qi = items[i];
    if (rec_off[qi] >= 0) {
        const long long o64 = row_off[qi], o32 = s_off[i];
        const int c64 = row_cnt[qi], c32 = s_cnt[i];
        if (o64 >= 0 && o32 >= 0 && o64 + c64 <= pool_cap && o32 + c32 <= pool_cap) {
            double err = 0.0;
            if (c64 != c32) {
                err = __longlong_as_double(0x7ff0000000000000ll);
            } else {
                for (int j = 0; j < c64; j++) {
                    if (pool_cols[o64 + j] != pool_cols[o32 + j])
                        err = __longlong_as_double(0x7ff0000000000000ll);
                    else
                        err = fmax(err, fabs(pool_vals[o64 + j] - pool_vals[o32 + j]));
                }
            }
            atomicMax(out, static_cast<unsigned long long>(__double_as_longlong(err)));
            atomicAdd(out + 1, 1ull);
            atomicAdd(out + 2, static_cast<unsigned long long>(c32));
        }
    }
    row_off[qi] = s_off[i];
    row_cnt[qi] = s_cnt[i];
    row_total[qi] = s_total[i];
}

hipError_t launch_mixed_check_save(const int *d_order, int n, int stride,
                                   const long long *row_off, const int *row_cnt,
                                   const double *row_total, int *items, long long *s_off,
                                   int *s_cnt, double *s_total, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(mixed_check_save_kernel, dim3((n + 255) / 256), dim3(256), 0, st, d_order,
                       n, stride, row_off, row_cnt, row_total, items, s_off, s_cnt, s_total);
    return hipGetLastError();
}

hipError_t launch_mixed_check_compare(int n, const int *items, const long long *rec_off,
                                      long long *row_off, int *row_cnt, double *row_total,
                                      const long long *s_off, const int *s_cnt,
                                      const double *s_total, const double *pool_vals,
                                      const int *pool_cols, long long pool_cap,
                                      unsigned long long *out, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(mixed_check_compare_kernel, dim3((n + 255) / 256), dim3(256), 0, st, n,
                       items, rec_off, row_off, row_cnt, row_total, s_off, s_cnt, s_total,
                       pool_vals, pool_cols, pool_cap, out);
    return hipGetLastError();
}

size_t morton_sort_bytes(int M) {
    const size_t Mz = static_cast<size_t>(M > 0 ? M : 1);
    const size_t nt = (Mz + kSortTile - 1) / kSortTile;
    const int nh = static_cast<int>(256 * nt);
    return 3 * rnd256(4 * Mz) + rnd256(4 * static_cast<size_t>(nh)) +
           rnd256(8 * (static_cast<size_t>(nh) + 1)) + rnd256(8 * scan_tmp_elems(nh));
}

hipError_t launch_morton_sort(const double *d_q, int M, const LocParams *d_lp,
                              const int *d_prob, int P, int bits, int *d_order, void *d_work,
                              hipStream_t st) {
    if (M <= 0) return hipSuccess;
    const size_t Mz = static_cast<size_t>(M);
    const int ntiles = static_cast<int>((Mz + kSortTile - 1) / kSortTile);
    const int nh = 256 * ntiles;
    char *w = static_cast<char *>(d_work);
    int *keys[2], *vals_tmp;
    keys[0] = reinterpret_cast<int *>(w);
    w += rnd256(4 * Mz);
    keys[1] = reinterpret_cast<int *>(w);
    w += rnd256(4 * Mz);
    vals_tmp = reinterpret_cast<int *>(w);
    w += rnd256(4 * Mz);
    int *hist = reinterpret_cast<int *>(w);
    w += rnd256(4 * static_cast<size_t>(nh));
    long long *offs = reinterpret_cast<long long *>(w);
    w += rnd256(8 * (static_cast<size_t>(nh) + 1));
    long long *tmp = reinterpret_cast<long long *>(w);
    hipLaunchKernelGGL(morton_key_kernel, dim3((M + 255) / 256), dim3(256), 0, st, d_q, M, d_lp,
                       d_prob, bits, keys[0]);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    // key bits: problem index above the 3*bits Morton bits
    const long long nkeys = static_cast<long long>(P) << (3 * bits);
    int keybits = 0;
    while ((1ll << keybits) < nkeys) keybits++;
    const int passes = (keybits + 7) / 8;
    if (passes == 0) {
        hipLaunchKernelGGL(iota_kernel, dim3((M + 255) / 256), dim3(256), 0, st, d_order, M);
        return hipGetLastError();
    }
    // values ping-pong so that the last pass writes d_order
    for (int k = 0; k < passes; k++) {
        const int shift = 8 * k;
        int *vin = k == 0 ? nullptr : ((passes - k) % 2 == 0 ? d_order : vals_tmp);
        int *vout = (passes - 1 - k) % 2 == 0 ? d_order : vals_tmp;
        hipLaunchKernelGGL(radix_hist_kernel, dim3(ntiles), dim3(64), 0, st, keys[k & 1], M,
                           shift, ntiles, hist);
        e = hipGetLastError();
        if (e != hipSuccess) return e;
        e = launch_scan_counts(hist, offs, nh, tmp, st);
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL(radix_scatter_kernel, dim3(ntiles), dim3(64), 0, st, keys[k & 1], vin,
                           M, shift, ntiles, offs, keys[(k + 1) & 1], vout);
        e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

__global__ void problem_ids_kernel(const long long *__restrict__ off, int P, int M,
                                   int *__restrict__ prob) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= M) return;
    int lo = 0, hi = P - 1;  // last p with off[p] <= i
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (off[mid] <= i) lo = mid; else hi = mid - 1;
    }
    prob[i] = lo;
}

hipError_t launch_problem_ids(const long long *d_offsets, int P, int M, int *d_prob,
                              hipStream_t st) {
    if (M <= 0) return hipSuccess;
    hipLaunchKernelGGL(problem_ids_kernel, dim3((M + 255) / 256), dim3(256), 0, st, d_offsets, P,
                       M, d_prob);
    return hipGetLastError();
}

// ---------------------------------------------------------------- scanning

constexpr int kScanBlock = 256;
constexpr int kScanItems = 8;
constexpr int kScanTile = kScanBlock * kScanItems;

__device__ __forceinline__ long long block_excl_scan(long long v, long long &total) {
    __shared__ long long wsum[kScanBlock / 64];
    const int lane = __lane_id();
    const int w = threadIdx.x >> 6;
    long long x = v;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const long long y = __shfl_up(x, d, 64);
        if (lane >= d) x += y;
    }
    if (lane == 63) wsum[w] = x;
    __syncthreads();
    long long off = 0, tot = 0;
#pragma unroll
    for (int j = 0; j < kScanBlock / 64; j++) {
        const long long s = wsum[j];
        if (j < w) off += s;
        tot += s;
    }
    __syncthreads();
    total = tot;
    return off + x - v;
}

__global__ void scan_tiles_kernel(const int *__restrict__ in, int n, long long *__restrict__ tsum) {
    const int base = blockIdx.x * kScanTile;
    long long s = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; k++) {
        const int i = base + k * kScanBlock + threadIdx.x;
        if (i < n) s += in[i];
    }
    long long total;
    block_excl_scan(s, total);
    if (threadIdx.x == 0) tsum[blockIdx.x] = total;
}

__global__ void scan_sums_kernel(long long *tsum, int ntiles, long long *grand) {
    long long carry = 0;
    for (int b = 0; b < ntiles; b += kScanBlock) {
        const int i = b + threadIdx.x;
        const long long v = i < ntiles ? tsum[i] : 0;
        long long total;
        const long long e = block_excl_scan(v, total);
        if (i < ntiles) tsum[i] = carry + e;
        carry += total;
    }
    if (threadIdx.x == 0) *grand = carry;
}

__global__ void scan_apply_kernel(const int *__restrict__ in, int n,
                                  const long long *__restrict__ tsum, long long *__restrict__ out) {
    const int base = blockIdx.x * kScanTile + threadIdx.x * kScanItems;
    long long v[kScanItems];
    long long s = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; k++) {
        const int i = base + k;
        v[k] = i < n ? in[i] : 0;
        s += v[k];
    }
    long long total;
    long long off = block_excl_scan(s, total) + tsum[blockIdx.x];
#pragma unroll
    for (int k = 0; k < kScanItems; k++) {
        const int i = base + k;
        if (i < n) out[i] = off;
        off += v[k];
    }
}

size_t scan_tmp_elems(int n) { return static_cast<size_t>((n + kScanTile - 1) / kScanTile) + 1; }

hipError_t launch_scan_counts(const int *d_in, long long *d_out, int n, long long *d_tmp,
                              hipStream_t st) {
    const int ntiles = (n + kScanTile - 1) / kScanTile;
    if (ntiles > 0) {
        hipLaunchKernelGGL(scan_tiles_kernel, dim3(ntiles), dim3(kScanBlock), 0, st, d_in, n,
                           d_tmp);
    }
    hipLaunchKernelGGL(scan_sums_kernel, dim3(1), dim3(kScanBlock), 0, st, d_tmp, ntiles,
                       d_out + n);
    if (ntiles > 0) {
        hipLaunchKernelGGL(scan_apply_kernel, dim3(ntiles), dim3(kScanBlock), 0, st, d_in, n,
                           d_tmp, d_out);
    }
    return hipGetLastError();
}

// ------------------------------------------------------------ CSR gather

// 16 lanes per row: each row's pool range and CSR range are contiguous, so
// the lanes of a row read and write consecutive entries.  Rows reaching past
// cap (a replay whose pool overflowed, reported separately) are skipped.
// Pool columns are device vertex indices; vmap turns them into the caller's
// (key) indices.
constexpr int kGatherLanes = 16;

__global__ void __launch_bounds__(256) gather_csr_kernel(
    const long long *__restrict__ rowptr64, const long long *__restrict__ row_off,
    const int *__restrict__ row_cnt, const double *__restrict__ pv, const int *__restrict__ pc,
    const int *__restrict__ vmap, int M, long long cap, double *__restrict__ vals,
    int *__restrict__ cols, int *__restrict__ rowptr32) {
    const long long t = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    const long long i = t / kGatherLanes;
    const int lane = static_cast<int>(t % kGatherLanes);
    if (i > M) return;
    if (lane == 0) rowptr32[i] = static_cast<int>(min(rowptr64[i], cap));
    if (i == M) return;
    const long long dst = rowptr64[i], src = row_off[i];
    const int c = row_cnt[i];
    if (src < 0 || src + c > cap || dst + c > cap) return;
    for (int j = lane; j < c; j += kGatherLanes) {
        vals[dst + j] = pv[src + j];
        cols[dst + j] = vmap[pc[src + j]];
    }
}

hipError_t launch_gather_csr(const long long *d_rowptr64, const long long *d_row_off,
                             const int *d_row_cnt, const double *pool_vals,
                             const int *pool_cols, const int *vmap, int M, long long cap,
                             double *d_vals, int *d_cols, int *d_rowptr32, hipStream_t st) {
    const long long threads = (static_cast<long long>(M) + 1) * kGatherLanes;
    hipLaunchKernelGGL(gather_csr_kernel, dim3(static_cast<unsigned>((threads + 255) / 256)),
                       dim3(256), 0, st, d_rowptr64, d_row_off, d_row_cnt, pool_vals, pool_cols,
                       vmap, M, cap, d_vals, d_cols, d_rowptr32);
    return hipGetLastError();
}

// out[i, :] = in[idx[i], :] for i < n (rows of D doubles): the caller's
// per-key values in device vertex order.
template <typename T>
__global__ void __launch_bounds__(256) gather_rows_kernel(const T *__restrict__ in,
                                                          const int *__restrict__ idx, int n,
                                                          int D, T *__restrict__ out) {
    const long long t = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= static_cast<long long>(n) * D) return;
    const long long i = t / D;
    const int d = static_cast<int>(t - i * D);
    out[t] = in[static_cast<long long>(idx[i]) * D + d];
}

hipError_t launch_gather_rows(const double *in, const int *idx, int n, int D, double *out,
                              hipStream_t st) {
    const long long total = static_cast<long long>(n) * D;
    if (total <= 0) return hipSuccess;
    hipLaunchKernelGGL(gather_rows_kernel<double>,
                       dim3(static_cast<unsigned>((total + 255) / 256)), dim3(256), 0, st, in,
                       idx, n, D, out);
    return hipGetLastError();
}

// ------------------------------------------------------ row canonicalisation

// Insertion sort of each row by column (rows are short; columns are unique).
template <typename TV>
__global__ void sort_rows_kernel(int M, const int *__restrict__ rowptr, int *__restrict__ cols,
                                 TV *__restrict__ vals) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= M) return;
    const int b = rowptr[i], e = rowptr[i + 1];
    for (int j = b + 1; j < e; j++) {
        const int c = cols[j];
        const TV v = vals[j];
        int k = j - 1;
        while (k >= b && cols[k] > c) {
            cols[k + 1] = cols[k];
            vals[k + 1] = vals[k];
            k--;
        }
        cols[k + 1] = c;
        vals[k + 1] = v;
    }
}

hipError_t launch_sort_rows(int M, const int *rowptr, int *cols, void *vals, int vdt,
                            hipStream_t st) {
    if (M <= 0) return hipSuccess;
    if (vdt == 0)
        hipLaunchKernelGGL(sort_rows_kernel<float>, dim3((M + 255) / 256), dim3(256), 0, st, M,
                           rowptr, cols, static_cast<float *>(vals));
    else
        hipLaunchKernelGGL(sort_rows_kernel<double>, dim3((M + 255) / 256), dim3(256), 0, st, M,
                           rowptr, cols, static_cast<double *>(vals));
    return hipGetLastError();
}

// ------------------------------------------------------------------- SpMM

// y[i,d] = sum over the row's entries in stored order of w * x[col,d],
// starting from zero, in the output type (scipy csr_matvec(s)).  One thread
// per (row, d); rows taken in `order` when given.
template <typename TW, typename TX, typename TY>
__global__ void spmm_kernel(int M, int D, const int *__restrict__ rowptr,
                            const int *__restrict__ cols, const TW *__restrict__ w,
                            const TX *__restrict__ x, const int *__restrict__ order,
                            TY *__restrict__ y) {
    const long long gid = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (gid >= static_cast<long long>(M) * D) return;
    const int i = order ? order[gid / D] : static_cast<int>(gid / D);
    const int d = static_cast<int>(gid % D);
    if (i < 0 || i >= M) return;
    TY acc = 0;
    const int b = rowptr[i], e = rowptr[i + 1];
    for (int j = b; j < e; j++) {
        const TY a = static_cast<TY>(w[j]);
        const TY v = static_cast<TY>(x[static_cast<size_t>(cols[j]) * D + d]);
        acc += a * v;
    }
    y[static_cast<size_t>(i) * D + d] = acc;
}

// The same sums for D <= 4 with 16 lanes per row: lane l forms the D
// products of entries b + l, b + 16 + l, ... (the row's weights and columns
// read by consecutive lanes), and lane d of the row adds column d's products
// in stored order from LDS -- the same roundings and the same sequence of
// additions as the loop above.  Rows in `order` (a spatial order keeps the
// rows of x that neighbouring rows share in L2).
constexpr int kSpmmLanes = 16;
constexpr int kSpmmRows = 256 / kSpmmLanes;

__device__ __forceinline__ void lanes_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}

// One row, entries [b, e), by its 16 lanes (prod: the row's LDS staging).
template <typename TW, typename TX, typename TY, int D>
__device__ __forceinline__ void spmm_row16(TY (*prod)[D], int lane, int b, int e,
                                           const int *__restrict__ cols,
                                           const TW *__restrict__ w, const TX *__restrict__ x,
                                           TY *__restrict__ yrow) {
    TY acc = 0;
    for (int c0 = b; c0 < e; c0 += kSpmmLanes) {
        const int j = c0 + lane;
        if (j < e) {
            const TY a = static_cast<TY>(w[j]);
            const size_t xb = static_cast<size_t>(cols[j]) * D;
#pragma unroll
            for (int d = 0; d < D; d++) prod[lane][d] = a * static_cast<TY>(x[xb + d]);
        }
        lanes_sync();
        if (lane < D) {
            const int n = min(kSpmmLanes, e - c0);
            for (int t = 0; t < n; t++) acc += prod[t][lane];
        }
        lanes_sync();
    }
    if (lane < D) yrow[lane] = acc;
}

template <typename TW, typename TX, typename TY, int D>
__global__ void __launch_bounds__(256) spmm_rows_kernel(int M, const int *__restrict__ rowptr,
                                                        const int *__restrict__ cols,
                                                        const TW *__restrict__ w,
                                                        const TX *__restrict__ x,
                                                        const int *__restrict__ order,
                                                        TY *__restrict__ y) {
    __shared__ TY prod[kSpmmRows][kSpmmLanes][D];
    const int lane = static_cast<int>(threadIdx.x) & (kSpmmLanes - 1);
    const int r = static_cast<int>(threadIdx.x) / kSpmmLanes;
    const long long k = static_cast<long long>(blockIdx.x) * kSpmmRows + r;
    int i = -1;
    if (k < M) i = order ? order[k] : static_cast<int>(k);
    if (i < 0 || i >= M) return;  // the row's 16 lanes leave together
    spmm_row16<TW, TX, TY, D>(prod[r], lane, rowptr[i], rowptr[i + 1], cols, w, x,
                              y + static_cast<size_t>(i) * D);
}

// The same product from the row pool of a compute: row i at
// [row_off[i], row_off[i] + row_cnt[i]) with device vertex columns, x in
// device vertex order, rows in the compute's order -- the pool rows are the
// emitted CSR rows (same entries, same order), so the sums are the same.
template <typename TX, int D>
__global__ void __launch_bounds__(256) pool_spmm_kernel(int M, const long long *__restrict__ row_off,
                                                        const int *__restrict__ row_cnt,
                                                        const int *__restrict__ cols,
                                                        const double *__restrict__ w,
                                                        const TX *__restrict__ x,
                                                        const int *__restrict__ order,
                                                        double *__restrict__ y) {
    __shared__ double prod[kSpmmRows][kSpmmLanes][D];
    const int lane = static_cast<int>(threadIdx.x) & (kSpmmLanes - 1);
    const int r = static_cast<int>(threadIdx.x) / kSpmmLanes;
    const long long k = static_cast<long long>(blockIdx.x) * kSpmmRows + r;
    if (k >= M) return;
    const int i = order[k];
    if (i < 0 || i >= M) return;
    const long long b = row_off[i];
    spmm_row16<double, TX, double, D>(prod[r], lane, static_cast<int>(0),
                                      row_cnt[i], cols + b, w + b, x,
                                      y + static_cast<size_t>(i) * D);
}

template <typename TX>
static hipError_t pool_spmm_t(int M, int D, const long long *off, const int *cnt, const int *c,
                              const double *w, const void *x, const int *order, double *y,
                              hipStream_t st) {
    const dim3 grid(static_cast<unsigned>((M + kSpmmRows - 1) / kSpmmRows));
    const TX *xt = static_cast<const TX *>(x);
    switch (D) {
        case 1:
            hipLaunchKernelGGL((pool_spmm_kernel<TX, 1>), grid, dim3(256), 0, st, M, off, cnt, c,
                               w, xt, order, y);
            break;
        case 2:
            hipLaunchKernelGGL((pool_spmm_kernel<TX, 2>), grid, dim3(256), 0, st, M, off, cnt, c,
                               w, xt, order, y);
            break;
        case 3:
            hipLaunchKernelGGL((pool_spmm_kernel<TX, 3>), grid, dim3(256), 0, st, M, off, cnt, c,
                               w, xt, order, y);
            break;
        case 4:
            hipLaunchKernelGGL((pool_spmm_kernel<TX, 4>), grid, dim3(256), 0, st, M, off, cnt, c,
                               w, xt, order, y);
            break;
        default:
            return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_pool_spmm(int M, int D, const long long *row_off, const int *row_cnt,
                            const int *pool_cols, const double *pool_vals, const void *x_dev,
                            int xdt, const int *order, double *y, hipStream_t st) {
    if (M <= 0) return hipSuccess;
    return xdt == 1 ? pool_spmm_t<double>(M, D, row_off, row_cnt, pool_cols, pool_vals, x_dev,
                                          order, y, st)
                    : pool_spmm_t<float>(M, D, row_off, row_cnt, pool_cols, pool_vals, x_dev,
                                         order, y, st);
}

hipError_t launch_gather_rows_bytes(const void *in, const int *idx, int n, int D, int elt,
                                    void *out, hipStream_t st) {
    const long long total = static_cast<long long>(n) * D;
    if (total <= 0) return hipSuccess;
    const dim3 grid(static_cast<unsigned>((total + 255) / 256));
    if (elt == 8)
        hipLaunchKernelGGL(gather_rows_kernel<double>, grid, dim3(256), 0, st,
                           static_cast<const double *>(in), idx, n, D, static_cast<double *>(out));
    else
        hipLaunchKernelGGL(gather_rows_kernel<float>, grid, dim3(256), 0, st,
                           static_cast<const float *>(in), idx, n, D, static_cast<float *>(out));
    return hipGetLastError();
}

template <typename TW, typename TX, typename TY>
static hipError_t spmm_t(int M, int D, const int *rp, const int *c, const void *w, const void *x,
                         const int *order, void *y, hipStream_t st) {
    const long long n = static_cast<long long>(M) * D;
    if (n == 0) return hipSuccess;
    const dim3 rows_grid(static_cast<unsigned>((M + kSpmmRows - 1) / kSpmmRows));
    const TW *wt = static_cast<const TW *>(w);
    const TX *xt = static_cast<const TX *>(x);
    TY *yt = static_cast<TY *>(y);
    switch (D) {
        case 1:
            hipLaunchKernelGGL((spmm_rows_kernel<TW, TX, TY, 1>), rows_grid, dim3(256), 0, st, M,
                               rp, c, wt, xt, order, yt);
            break;
        case 2:
            hipLaunchKernelGGL((spmm_rows_kernel<TW, TX, TY, 2>), rows_grid, dim3(256), 0, st, M,
                               rp, c, wt, xt, order, yt);
            break;
        case 3:
            hipLaunchKernelGGL((spmm_rows_kernel<TW, TX, TY, 3>), rows_grid, dim3(256), 0, st, M,
                               rp, c, wt, xt, order, yt);
            break;
        case 4:
            hipLaunchKernelGGL((spmm_rows_kernel<TW, TX, TY, 4>), rows_grid, dim3(256), 0, st, M,
                               rp, c, wt, xt, order, yt);
            break;
        default:
            hipLaunchKernelGGL((spmm_kernel<TW, TX, TY>),
                               dim3(static_cast<unsigned>((n + 255) / 256)), dim3(256), 0, st, M,
                               D, rp, c, wt, xt, order, yt);
    }
    return hipGetLastError();
}

hipError_t launch_spmm(int M, int D, const int *rp, const int *c, const void *w, int vdt,
                       const void *x, int xdt, const int *order, void *y, int ydt,
                       hipStream_t st) {
    const int code = vdt * 4 + xdt * 2 + ydt;
    switch (code) {
        case 0: return spmm_t<float, float, float>(M, D, rp, c, w, x, order, y, st);
        case 1: return spmm_t<float, float, double>(M, D, rp, c, w, x, order, y, st);
        case 3: return spmm_t<float, double, double>(M, D, rp, c, w, x, order, y, st);
        case 5: return spmm_t<double, float, double>(M, D, rp, c, w, x, order, y, st);
        case 7: return spmm_t<double, double, double>(M, D, rp, c, w, x, order, y, st);
        default: return hipErrorInvalidValue;
    }
}

}  // namespace nat
