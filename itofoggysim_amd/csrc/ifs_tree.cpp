// Octree (host, integer work): bit-exact restatement of libtree.
#include "ifs_tree.h"

#include <algorithm>
#include <cstring>

namespace ifs {
namespace tree {

Scaling scale_points(const double *pts, int n, std::vector<double> &scaled) {
    Scaling S{};
    double D[3];
    for (int d = 0; d < 3; ++d) {
        double lo = pts[d], hi = pts[d];
        for (int i = 1; i < n; ++i) {
            lo = std::min(lo, pts[3 * i + d]);
            hi = std::max(hi, pts[3 * i + d]);
        }
        S.x_min[d] = lo;
        D[d] = (hi - lo) * (1.0 + 0x1p-52);  // lib_tree.f90:351
    }
    S.D = std::max(D[0], std::max(D[1], D[2]));  // same scaling for all dimensions (:361-363)
    if (!(S.D > 0.0)) S.D = 1.0;                 // a single point: any positive edge
    scaled.resize((size_t)3 * n);
    for (int i = 0; i < n; ++i)
        for (int d = 0; d < 3; ++d) scaled[3 * i + d] = (pts[3 * i + d] - S.x_min[d]) / S.D;
    return S;
}

uint64_t coordinate_binary(double f) {
    if (!(f > 0.0)) return 0;
    uint64_t bits;
    std::memcpy(&bits, &f, 8);
    int64_t expo = (int64_t)(bits >> 52);
    uint64_t mant = (bits << 11) | (1ULL << 63);
    int64_t shift = 1023 - expo - 1;
    if (shift >= 64) return 0;
    if (shift < 0) return ~0ULL;  // f >= 1 does not occur for scaled points
    return mant >> shift;
}

int64_t universal_index(const double *p, int l) {
    uint64_t b[3];
    for (int d = 0; d < 3; ++d) b[d] = coordinate_binary(p[d]) >> (64 - l);
    int64_t n = 0;
    for (int i = l - 1; i >= 0; --i)
        n = (n << 3) | (int64_t)(((b[0] >> i) & 1) << 2) | (int64_t)(((b[1] >> i) & 1) << 1) | (int64_t)((b[2] >> i) & 1);
    return n;
}

static void deinterleave(int64_t n, int l, int64_t v[3]) {
    v[0] = v[1] = v[2] = 0;
    for (int i = l - 1; i >= 0; --i) {
        int64_t t = (n >> (3 * i)) & 7;
        v[0] = (v[0] << 1) | ((t >> 2) & 1);
        v[1] = (v[1] << 1) | ((t >> 1) & 1);
        v[2] = (v[2] << 1) | (t & 1);
    }
}

static int64_t interleave(const int64_t v[3], int l) {
    int64_t n = 0;
    for (int i = l - 1; i >= 0; --i)
        n = (n << 3) | (((v[0] >> i) & 1) << 2) | (((v[1] >> i) & 1) << 1) | ((v[2] >> i) & 1);
    return n;
}

void children(int64_t n, int64_t out[8]) {
    for (int j = 0; j < 8; ++j) out[j] = 8 * n + j;
}

void centre(int64_t n, int l, double out[3]) {
    int64_t v[3];
    deinterleave(n, l, v);
    double h = 1.0;
    for (int i = 0; i < l; ++i) h *= 0.5;
    for (int d = 0; d < 3; ++d) out[d] = h * ((double)v[d] + 0.5);
}

void neighbours(int64_t k, int64_t n, int l, int64_t out[26]) {
    int64_t v[3];
    deinterleave(n, l, v);
    const int64_t upper = (l >= 63) ? INT64_MAX : ((int64_t)1 << l) - 1;
    int64_t c[3][3];
    for (int d = 0; d < 3; ++d) {
        c[d][0] = v[d] - k < 0 ? -1 : v[d] - k;
        c[d][1] = v[d];
        c[d][2] = v[d] + k > upper ? -1 : v[d] + k;
    }
    for (int i = 0; i < 26; ++i) out[i] = -1;
    int cnt = 0;
    for (int a = 0; a < 3; ++a) {
        if (c[0][a] < 0) continue;
        for (int b = 0; b < 3; ++b) {
            if (c[1][b] < 0) continue;
            for (int e = 0; e < 3; ++e) {
                if (c[2][e] < 0) continue;
                int64_t w[3] = {c[0][a], c[1][b], c[2][e]};
                int64_t m = interleave(w, l);
                if (m != n) out[cnt++] = m;
            }
        }
    }
}

std::vector<int64_t> domain_i2(int64_t n, int l, int64_t k) {
    std::vector<int64_t> r{n};
    for (int64_t kk = 1; kk <= k; ++kk) {
        int64_t nb[26];
        neighbours(kk, n, l, nb);
        for (int i = 0; i < 26; ++i)
            if (nb[i] >= 0) r.push_back(nb[i]);
    }
    return r;
}

std::vector<int64_t> domain_i4(int64_t n, int l, int64_t k) {
    std::vector<int64_t> own = domain_i2(n, l, k);
    std::vector<int64_t> par = domain_i2(parent(n), l - 1, k);
    std::vector<int64_t> r;
    for (int64_t p : par) {
        int64_t ch[8];
        children(p, ch);
        for (int j = 0; j < 8; ++j) r.push_back(ch[j]);
    }
    std::sort(r.begin(), r.end());
    std::sort(own.begin(), own.end());
    std::vector<int64_t> out;
    std::set_difference(r.begin(), r.end(), own.begin(), own.end(), std::back_inserter(out));
    return out;
}

int level_min(int64_t k) {
    int lg = 0;
    while ((int64_t)1 << (lg + 1) <= k + 1) ++lg;
    return 1 + lg;
}

int level_max(const std::vector<int64_t> &u, int64_t s) {
    const int64_t N = (int64_t)u.size();
    int lmax = 1;
    for (int64_t i = 0; i + s < N; ++i) {
        int64_t a = u[i], b = u[i + s];
        for (int j = kThresholdLevel; j >= 1; --j) {
            a = parent(a);
            b = parent(b);
            if (a == b) {
                lmax = std::max(lmax, j);
                break;
            }
        }
    }
    return lmax;
}

}  // namespace tree
}  // namespace ifs

// ---------------------------------------------------------------------------
// C ABI of the octree (host only; no device work).  These replace libtree's
// lib_tree_hf_get_universal_index, lib_tree_get_neighbours, lib_tree_get_domain_i4
// and lib_tree_get_level_max for bit-exact checks.
// ---------------------------------------------------------------------------
#include "../../include/ifs.h"

extern "C" {

int ifs_tree_scale(const double *pts, int32_t n, double *scaled, double *x_min_D) {
    if (!pts || !scaled || n < 1) return IFS_E_ARG;
    std::vector<double> s;
    ifs::tree::Scaling S = ifs::tree::scale_points(pts, n, s);
    std::memcpy(scaled, s.data(), sizeof(double) * 3 * n);
    if (x_min_D) {
        x_min_D[0] = S.x_min[0];
        x_min_D[1] = S.x_min[1];
        x_min_D[2] = S.x_min[2];
        x_min_D[3] = S.D;
    }
    return IFS_OK;
}

int ifs_tree_universal_index(const double *scaled, int32_t n, int32_t level, int64_t *out) {
    if (!scaled || !out || n < 0 || level < 1 || level > ifs::tree::kThresholdLevel) return IFS_E_ARG;
    for (int i = 0; i < n; ++i) out[i] = ifs::tree::universal_index(scaled + 3 * i, level);
    return IFS_OK;
}

int ifs_tree_neighbours(int64_t k, int64_t box, int32_t level, int64_t out[26]) {
    if (!out || level < 1 || k < 1) return IFS_E_ARG;
    ifs::tree::neighbours(k, box, level, out);
    return IFS_OK;
}

int ifs_tree_centre(int64_t box, int32_t level, double out[3]) {
    if (!out || level < 0) return IFS_E_ARG;
    ifs::tree::centre(box, level, out);
    return IFS_OK;
}

int ifs_tree_domain_i4(int64_t box, int32_t level, int64_t *out, int32_t cap) {
    if (!out || level < 2) return IFS_E_ARG;
    std::vector<int64_t> v = ifs::tree::domain_i4(box, level, 1);
    int cnt = (int)v.size();
    for (int i = 0; i < cnt && i < cap; ++i) out[i] = v[i];
    return cnt;
}

int ifs_tree_level_max(const int64_t *sorted_uindex, int64_t n, int64_t s) {
    if ((!sorted_uindex && n > 0) || s < 1) return IFS_E_ARG;
    std::vector<int64_t> u(sorted_uindex, sorted_uindex + n);
    return ifs::tree::level_max(u, s);
}

}  // extern "C"
