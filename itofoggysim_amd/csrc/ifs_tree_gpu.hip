// Octree build on the GPU (gfx950), bit-exact with the host restatement of
// libtree (ifs_tree.cpp): scaling, Morton universal index at the threshold
// level, sort, l_max by the s-rule and the I4 interaction lists.
//
//   scaling   lib_tree.f90:320-374: x_min, D = max_d (max - min)(1 + 2^-52); p' = (p - x_min)/D
//   Morton    lib_tree_helper_functions.f90:202-364 (IEEE fraction bits, :532-614;
//             interleave with x the most significant bit of each triple, :1204-1240)
//   l_max     lib_tree.f90:1188-1242: smallest level with at most s elements per box
//   I4        lib_tree.f90:994-1143: children of the parent's I2 minus own I2
// All work is integer except the scaling, whose three IEEE operations (subtract,
// multiply, divide) round the same on the device as on the host.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <vector>

#include "../../include/ifs.h"
#include "ifs_engine.h"
#include "ifs_tree.h"

namespace ifs {
namespace {

constexpr int kLth = 21;

__device__ __forceinline__ uint64_t coord_bits(double f) {
    if (!(f > 0.0)) return 0;
    uint64_t bits = (uint64_t)__double_as_longlong(f);
    int64_t expo = (int64_t)(bits >> 52);
    uint64_t mant = (bits << 11) | (1ULL << 63);
    int64_t shift = 1023 - expo - 1;
    if (shift >= 64) return 0;
    if (shift < 0) return ~0ULL;
    return mant >> shift;
}

__global__ void k_minmax(const double *__restrict__ p, int n, double *__restrict__ out) {
    // one block: out[0..2] = min, out[3..5] = max (exact, order independent)
    __shared__ double lo[3][256], hi[3][256];
    double l[3] = {INFINITY, INFINITY, INFINITY}, h[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int i = threadIdx.x; i < n; i += blockDim.x)
        for (int d = 0; d < 3; ++d) {
            double v = p[3 * i + d];
            l[d] = fmin(l[d], v);
            h[d] = fmax(h[d], v);
        }
    for (int d = 0; d < 3; ++d) {
        lo[d][threadIdx.x] = l[d];
        hi[d][threadIdx.x] = h[d];
    }
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s)
            for (int d = 0; d < 3; ++d) {
                lo[d][threadIdx.x] = fmin(lo[d][threadIdx.x], lo[d][threadIdx.x + s]);
                hi[d][threadIdx.x] = fmax(hi[d][threadIdx.x], hi[d][threadIdx.x + s]);
            }
        __syncthreads();
    }
    if (threadIdx.x < 3) {
        out[threadIdx.x] = lo[threadIdx.x][0];
        out[3 + threadIdx.x] = hi[threadIdx.x][0];
    }
}

__global__ void k_morton(const double *__restrict__ p, int n, double x0, double y0, double z0, double D,
                         int64_t *__restrict__ key, int32_t *__restrict__ idx) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double mn[3] = {x0, y0, z0};
    uint64_t b[3];
    for (int d = 0; d < 3; ++d) {
        double f = __ddiv_rn(__dsub_rn(p[3 * i + d], mn[d]), D);
        b[d] = coord_bits(f) >> (64 - kLth);
    }
    int64_t u = 0;
    for (int s = kLth - 1; s >= 0; --s)
        u = (u << 3) | (int64_t)(((b[0] >> s) & 1) << 2) | (int64_t)(((b[1] >> s) & 1) << 1) | (int64_t)((b[2] >> s) & 1);
    key[i] = u;
    idx[i] = i;
}

// flag[l] = 1 if some box at level l holds more than s of the sorted keys:
// keys are sorted, so a box with s+1 elements shows as key_l[i] == key_l[i-s].
__global__ void k_overfull(const int64_t *__restrict__ key, int n, int s, int32_t *__restrict__ flag) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    int l = blockIdx.y + 1;
    if (i < s || i >= n) return;
    int sh = 3 * (kLth - l);
    if ((key[i] >> sh) == (key[i - s] >> sh)) flag[l] = 1;
}

__device__ __forceinline__ void demorton(int64_t n, int l, int64_t v[3]) {
    v[0] = v[1] = v[2] = 0;
    for (int i = l - 1; i >= 0; --i) {
        int64_t t = (n >> (3 * i)) & 7;
        v[0] = (v[0] << 1) | ((t >> 2) & 1);
        v[1] = (v[1] << 1) | ((t >> 1) & 1);
        v[2] = (v[2] << 1) | (t & 1);
    }
}
__device__ __forceinline__ int64_t morton3(const int64_t v[3], int l) {
    int64_t n = 0;
    for (int i = l - 1; i >= 0; --i)
        n = (n << 3) | (((v[0] >> i) & 1) << 2) | (((v[1] >> i) & 1) << 1) | ((v[2] >> i) & 1);
    return n;
}
__device__ __forceinline__ bool present(const int64_t *__restrict__ boxes, int64_t nb, int64_t b) {
    int64_t lo = 0, hi = nb;
    while (lo < hi) {
        int64_t mid = (lo + hi) >> 1;
        if (boxes[mid] < b) lo = mid + 1;
        else hi = mid;
    }
    return lo < nb && boxes[lo] == b;
}

// I4 of every box of a sorted box list at level l (k = 1), restricted to the
// boxes present, ascending: out[b * 189 + j], count[b].
__global__ void k_i4(const int64_t *__restrict__ boxes, int64_t nb, int l, int64_t *__restrict__ out,
                     int32_t *__restrict__ count) {
    int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= nb) return;
    const int64_t b = boxes[q];
    int64_t v[3], pv[3];
    demorton(b, l, v);
    demorton(b >> 3, l - 1, pv);
    const int64_t up = ((int64_t)1 << (l - 1)) - 1;
    // children of the parent's I2 (parent neighbourhood at level l-1) in ascending Morton
    // order: iterate the parent-level cube and the 8 children, then sort the short list
    int64_t cand[216];
    int nc = 0;
    for (int dx = -1; dx <= 1; ++dx)
        for (int dy = -1; dy <= 1; ++dy)
            for (int dz = -1; dz <= 1; ++dz) {
                int64_t w[3] = {pv[0] + dx, pv[1] + dy, pv[2] + dz};
                if (w[0] < 0 || w[1] < 0 || w[2] < 0 || w[0] > up || w[1] > up || w[2] > up) continue;
                int64_t pm = morton3(w, l - 1);
                for (int j = 0; j < 8; ++j) {
                    int64_t ch = 8 * pm + j;
                    int64_t cv[3];
                    demorton(ch, l, cv);
                    int64_t ax = cv[0] - v[0], ay = cv[1] - v[1], az = cv[2] - v[2];
                    bool own = ax >= -1 && ax <= 1 && ay >= -1 && ay <= 1 && az >= -1 && az <= 1;
                    if (!own && present(boxes, nb, ch)) cand[nc++] = ch;
                }
            }
    // insertion sort (<= 189 entries)
    for (int i = 1; i < nc; ++i) {
        int64_t x = cand[i];
        int j = i - 1;
        while (j >= 0 && cand[j] > x) {
            cand[j + 1] = cand[j];
            --j;
        }
        cand[j + 1] = x;
    }
    for (int i = 0; i < nc; ++i) out[q * 189 + i] = cand[i];
    for (int i = nc; i < 189; ++i) out[q * 189 + i] = -1;
    count[q] = nc;
}

}  // namespace
}  // namespace ifs

#define TRY(x)                                  \
    do {                                        \
        if ((x) != hipSuccess) return IFS_E_HIP; \
    } while (0)

extern "C" int ifs_tree_build_gpu(const double *pts, int32_t n, int32_t s, int64_t *uindex_sorted, int32_t *perm,
                                  int32_t levels[2]) {
    if (!pts || n <= 0 || s <= 0 || !uindex_sorted || !perm || !levels) return IFS_E_ARG;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return IFS_E_NODEVICE;
    ifs::DevBuf dp, dmm, dk, di, dk2, di2, dflag, tmp;
    int rc;
    if ((rc = dp.alloc(24ull * n)) || (rc = dmm.alloc(48)) || (rc = dk.alloc(8ull * n)) || (rc = di.alloc(4ull * n)) ||
        (rc = dk2.alloc(8ull * n)) || (rc = di2.alloc(4ull * n)) || (rc = dflag.alloc(4 * 32)))
        return rc;
    TRY(hipMemcpy(dp.p, pts, 24ull * n, hipMemcpyHostToDevice));
    hipLaunchKernelGGL(ifs::k_minmax, dim3(1), dim3(256), 0, 0, dp.as<double>(), n, dmm.as<double>());
    double mm[6];
    TRY(hipMemcpy(mm, dmm.p, 48, hipMemcpyDeviceToHost));
    double D = 0.0;
    for (int d = 0; d < 3; ++d) D = std::max(D, (mm[3 + d] - mm[d]) * (1.0 + 0x1p-52));
    if (!(D > 0.0)) D = 1.0;
    hipLaunchKernelGGL(ifs::k_morton, dim3((n + 255) / 256), dim3(256), 0, 0, dp.as<double>(), n, mm[0], mm[1],
                       mm[2], D, dk.as<int64_t>(), di.as<int32_t>());
    size_t tbytes = 0;
    TRY(hipcub::DeviceRadixSort::SortPairs(nullptr, tbytes, dk.as<int64_t>(), dk2.as<int64_t>(), di.as<int32_t>(),
                                           di2.as<int32_t>(), n));
    if ((rc = tmp.alloc(std::max<size_t>(tbytes, 16)))) return rc;
    TRY(hipcub::DeviceRadixSort::SortPairs(tmp.p, tbytes, dk.as<int64_t>(), dk2.as<int64_t>(), di.as<int32_t>(),
                                           di2.as<int32_t>(), n));
    TRY(hipMemset(dflag.p, 0, 4 * 32));
    if (n > s)
        hipLaunchKernelGGL(ifs::k_overfull, dim3((n + 255) / 256, ifs::kLth), dim3(256), 0, 0, dk2.as<int64_t>(), n,
                           s, dflag.as<int32_t>());
    int32_t flag[32];
    TRY(hipMemcpy(flag, dflag.p, 4 * 32, hipMemcpyDeviceToHost));
    int lmax = ifs::kLth;
    for (int l = 1; l <= ifs::kLth; ++l)
        if (!flag[l]) {
            lmax = l;
            break;
        }
    TRY(hipMemcpy(uindex_sorted, dk2.p, 8ull * n, hipMemcpyDeviceToHost));
    TRY(hipMemcpy(perm, di2.p, 4ull * n, hipMemcpyDeviceToHost));
    levels[0] = 2;  // l_min for the k = 1 neighbourhood (lib_tree.f90:1162-1186)
    levels[1] = std::max(lmax, 2);
    return IFS_OK;
}

extern "C" int ifs_tree_i4_gpu(const int64_t *boxes, int64_t nb, int32_t level, int64_t *out, int32_t *count) {
    if (!boxes || nb < 0 || level < 2 || level > 21 || !out || !count) return IFS_E_ARG;
    if (nb == 0) return IFS_OK;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return IFS_E_NODEVICE;
    for (int64_t i = 1; i < nb; ++i)
        if (boxes[i] <= boxes[i - 1]) return IFS_E_ARG;  // ascending, distinct
    ifs::DevBuf db, dout, dc;
    int rc;
    if ((rc = db.alloc(8ull * nb)) || (rc = dout.alloc(8ull * 189 * nb)) || (rc = dc.alloc(4ull * nb))) return rc;
    TRY(hipMemcpy(db.p, boxes, 8ull * nb, hipMemcpyHostToDevice));
    hipLaunchKernelGGL(ifs::k_i4, dim3((unsigned)((nb + 127) / 128)), dim3(128), 0, 0, db.as<int64_t>(), nb, level,
                       dout.as<int64_t>(), dc.as<int32_t>());
    TRY(hipGetLastError());
    TRY(hipMemcpy(out, dout.p, 8ull * 189 * nb, hipMemcpyDeviceToHost));
    TRY(hipMemcpy(count, dc.p, 4ull * nb, hipMemcpyDeviceToHost));
    return IFS_OK;
}
