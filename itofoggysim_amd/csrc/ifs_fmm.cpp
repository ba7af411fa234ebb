// MLFMM plan and passes.
//
// Restates libmlfmm's algorithm for the Mie plug-in
// (libmlfmm/src/lib/lib_ml_fmm.f90:231-262 and the callbacks of
// FoggySim/src/lib/mie_theory/lib_mie_ms_ml_fmm_interface.f90), with one
// translation convention throughout: every translation uses
// d = dsign * (origin(source) - origin(target)) in Xu's coefficients (dsign = -1
// by default, +1 for the reference's direct-path sign), regular waves (j_n)
// for outgoing->outgoing (P2M, M2M) and regular->regular (L2L, L2P), and the
// sphere-interaction kind (h1_n) for outgoing->regular (M2L, near field).
//
//   upward   : C(leaf) = sum P2M(sphere)              (lib_ml_fmm.f90:387-508)
//              C(parent) = sum M2M(C(child))          (:520-640)
//   downward : D(box) = sum_{I4(box)} M2L(C(src))     (:691-847)
//              D(box) += L2L(D(parent))               (:861-982)
//   final    : v_j = L2P(D(leaf)) + sum_{E2, i != j} P2P(x_i)  (:994-1204)
// Boxes carry a per-level truncation L = kD + 1.8 d0^(2/3) (kD)^(1/3), D the box edge.
#include "ifs_fmm.h"

#include <algorithm>
#include <array>
#include <cmath>
#include <map>
#include <unordered_map>
#include <vector>

#include "ifs_engine.h"
#include "ifs_tree.h"

namespace ifs {
namespace fmm {

struct Level {
    int l = 0;
    int L = 0;
    double edge = 0.0;
    std::vector<int64_t> boxes;                 // non-empty boxes, ascending Morton index
    std::unordered_map<int64_t, int> index;     // Morton -> position
    std::vector<double> centre;                 // 3 per box (unscaled)
    int64_t off = 0;                            // pool offset of the first box (complex)
    int64_t per_box() const { return 2LL * L * (L + 2); }
};

struct Fmm {
    int l_min = 2, l_max = 2;
    std::vector<Level> lv;                      // index l (0..l_max), only l_min..l_max used
    std::vector<int64_t> leaf_of;               // sphere -> leaf Morton index
    DevBuf C, D;                                // multipole / local pools
    int64_t nC = 0;
    Batch p2m, l2p, p2p;
    std::vector<ClassBatch> m2m, m2l, l2l;      // indexed by level
    int64_t counts[6] = {0, 0, 0, 0, 0, 0};
};

static int box_L(double k, double edge, double d0, int cap) {
    double kD = k * edge;
    int L = (int)std::ceil(kD + 1.8 * std::pow(d0, 2.0 / 3.0) * std::cbrt(kD));
    if (L < 1) L = 1;
    if (cap > 0 && L > cap) L = cap;
    return L;
}

// Appends one class (shared translation vector dv) with its columns (source,
// target offsets).  For an integer box offset o (o != nullptr, dv = o * edge) the
// length and polar angle are taken from |o|^2 and o_z exactly, so that offsets
// of one length (and one o_z) share their tables bit for bit.
static void add_class(ClassBatch &B, const double *dv, double k, int kind, int Ls, int Lt,
                      const std::vector<std::pair<int64_t, int64_t>> &cols, const int *o = nullptr,
                      double edge = 0.0) {
    if (cols.empty()) return;
    Task T = make_task(dv, k, kind, Ls, Lt, 0, 0);
    T.off_d = T.off_s = T.off_ax = T.off_y = T.off_z = 0;
    if (o) {
        double r = std::sqrt((double)(o[0] * o[0] + o[1] * o[1] + o[2] * o[2]));
        T.t = k * (edge * r);
        T.cth = std::max(-1.0, std::min(1.0, (double)o[2] / r));
    }
    for (auto &pr : cols) {
        B.src.push_back(pr.first);
        B.dst.push_back(pr.second);
    }
    B.ncols.push_back((int)cols.size());
    B.gen.push_back(T);
}

static void box_coords(int64_t n, int l, int64_t v[3]) {
    v[0] = v[1] = v[2] = 0;
    for (int i = l - 1; i >= 0; --i) {
        int64_t t = (n >> (3 * i)) & 7;
        v[0] = (v[0] << 1) | ((t >> 2) & 1);
        v[1] = (v[1] << 1) | ((t >> 1) & 1);
        v[2] = (v[2] << 1) | (t & 1);
    }
}

int build(ifs_ctx *c, Fmm **out) {
    *out = nullptr;
    Fmm *f = new Fmm();
    const int ns = c->ns;
    const double k = c->prm.k;
    const int s = c->prm.leaf_size > 0 ? c->prm.leaf_size : 10;
    const double d0 = c->prm.fmm_d0 > 0.0 ? c->prm.fmm_d0 : 6.0;

    // octree (lib_tree_constructor, libtree/src/lib/lib_tree.f90:147-270)
    std::vector<double> scaled;
    tree::Scaling S = tree::scale_points(c->pos.data(), ns, scaled);
    std::vector<int64_t> u(ns);
    for (int i = 0; i < ns; ++i) u[i] = tree::universal_index(&scaled[3 * i], tree::kThresholdLevel);
    std::vector<int64_t> us = u;
    std::sort(us.begin(), us.end());
    f->l_min = tree::level_min(1);
    f->l_max = std::max(tree::level_max(us, s), f->l_min);
    f->l_max = std::min(f->l_max, tree::kThresholdLevel);
    const int lmax = f->l_max, lmin = f->l_min;
    f->leaf_of.resize(ns);
    for (int i = 0; i < ns; ++i) f->leaf_of[i] = u[i] >> (3 * (tree::kThresholdLevel - lmax));

    // hierarchy of non-empty boxes per level (lib_ml_fmm_hf_create_hierarchy,
    // libmlfmm/src/lib/lib_ml_fmm_helper_functions.f90:366-446)
    f->lv.resize(lmax + 1);
    int64_t pool = 0;
    for (int l = lmin; l <= lmax; ++l) {
        Level &L = f->lv[l];
        L.l = l;
        L.edge = S.D / std::ldexp(1.0, l);
        L.L = box_L(k, L.edge, d0, c->prm.fmm_lmax_cap);
        std::vector<int64_t> b(ns);
        for (int i = 0; i < ns; ++i) b[i] = f->leaf_of[i] >> (3 * (lmax - l));
        std::sort(b.begin(), b.end());
        b.erase(std::unique(b.begin(), b.end()), b.end());
        L.boxes = b;
        L.centre.resize(3 * b.size());
        for (size_t q = 0; q < b.size(); ++q) {
            L.index[b[q]] = (int)q;
            double cs[3];
            tree::centre(b[q], l, cs);
            for (int d = 0; d < 3; ++d) L.centre[3 * q + d] = cs[d] * S.D + S.x_min[d];  // lib_tree.f90:376-389
        }
        L.off = pool;
        pool += (int64_t)b.size() * L.per_box();
    }
    f->nC = pool;
    int rc;
    if ((rc = f->C.alloc(16ull * pool)) || (rc = f->D.alloc(16ull * pool))) {
        delete f;
        return rc;
    }
    // translation vector d = dsign * (origin(source) - origin(target))
    const double sg = c->dsign;
    const int isg = sg > 0 ? 1 : -1;
    auto dvec = [sg](const double *src, const double *dst, double *dv) {
        for (int d = 0; d < 3; ++d) dv[d] = sg * (src[d] - dst[d]);
    };
    const Level &LL = f->lv[lmax];

    // P2M: sphere -> leaf box, grouped by target box (ascending sphere index)
    {
        std::vector<std::vector<int>> members(LL.boxes.size());
        for (int i = 0; i < ns; ++i) members[LL.index.at(f->leaf_of[i])].push_back(i);
        for (size_t q = 0; q < LL.boxes.size(); ++q)
            for (int i : members[q]) {
                double dv[3];
                dvec(&c->pos[3 * i], &LL.centre[3 * q], dv);
                f->p2m.tasks.push_back(make_task(dv, k, 1, c->Leff[i], LL.L, c->off[i], LL.off + q * LL.per_box()));
            }
        f->counts[0] = (int64_t)f->p2m.tasks.size();
    }
    // M2M: child (level l+1) -> parent (level l), one class per child octant;
    // d = dsign * (centre(child) - centre(parent)) = dsign * (b - 1/2) * edge(child)
    f->m2m.resize(lmax + 1);
    for (int l = lmax - 1; l >= lmin; --l) {
        const Level &P = f->lv[l], &Ch = f->lv[l + 1];
        std::vector<std::vector<std::pair<int64_t, int64_t>>> oct(8);
        for (size_t q = 0; q < Ch.boxes.size(); ++q) {
            int p = P.index.at(tree::parent(Ch.boxes[q]));
            oct[Ch.boxes[q] & 7].push_back({Ch.off + (int64_t)q * Ch.per_box(), P.off + (int64_t)p * P.per_box()});
        }
        for (int b = 0; b < 8; ++b) {
            double dv[3] = {sg * (((b >> 2) & 1) - 0.5) * Ch.edge, sg * (((b >> 1) & 1) - 0.5) * Ch.edge,
                            sg * ((b & 1) - 0.5) * Ch.edge};
            std::sort(oct[b].begin(), oct[b].end(), [](auto &x, auto &y) { return x.second < y.second; });
            add_class(f->m2m[l], dv, k, 1, Ch.L, P.L, oct[b]);
            f->counts[1] += (int64_t)oct[b].size();
        }
    }
    // M2L: I4 of every box, one class per integer offset o = coords(src) - coords(tgt);
    // d = dsign * o * edge.  Classes ordered by |o|^2, then o_z, so consecutive classes share tables.
    f->m2l.resize(lmax + 1);
    for (int l = lmin; l <= lmax; ++l) {
        const Level &V = f->lv[l];
        std::map<std::array<int, 3>, std::vector<std::pair<int64_t, int64_t>>> cls;
        for (size_t q = 0; q < V.boxes.size(); ++q) {
            int64_t vt[3];
            box_coords(V.boxes[q], l, vt);
            std::vector<int64_t> i4 = tree::domain_i4(V.boxes[q], l, 1);
            for (int64_t src : i4) {
                auto it = V.index.find(src);
                if (it == V.index.end()) continue;
                int p = it->second;
                int64_t vs[3];
                box_coords(src, l, vs);
                std::array<int, 3> o = {(int)(vs[0] - vt[0]), (int)(vs[1] - vt[1]), (int)(vs[2] - vt[2])};
                cls[o].push_back({V.off + (int64_t)p * V.per_box(), V.off + (int64_t)q * V.per_box()});
            }
        }
        std::vector<std::array<int, 3>> keys;
        for (auto &kv : cls) keys.push_back(kv.first);
        std::sort(keys.begin(), keys.end(), [](const std::array<int, 3> &a, const std::array<int, 3> &b) {
            int ra = a[0] * a[0] + a[1] * a[1] + a[2] * a[2], rb = b[0] * b[0] + b[1] * b[1] + b[2] * b[2];
            if (ra != rb) return ra < rb;
            if (a[2] != b[2]) return a[2] < b[2];
            return a < b;
        });
        for (auto &o : keys) {
            auto &cols = cls[o];
            std::sort(cols.begin(), cols.end(), [](auto &x, auto &y) { return x.second < y.second; });
            int os[3] = {isg * o[0], isg * o[1], isg * o[2]};
            double dv[3] = {os[0] * V.edge, os[1] * V.edge, os[2] * V.edge};
            add_class(f->m2l[l], dv, k, c->prm.z_selector, V.L, V.L, cols, os, V.edge);
            f->counts[2] += (int64_t)cols.size();
        }
    }
    // L2L: parent (level l-1) -> child (level l), one class per child octant;
    // d = dsign * (centre(parent) - centre(child)) = -dsign * (b - 1/2) * edge(child)
    f->l2l.resize(lmax + 1);
    for (int l = lmin + 1; l <= lmax; ++l) {
        const Level &P = f->lv[l - 1], &Ch = f->lv[l];
        std::vector<std::vector<std::pair<int64_t, int64_t>>> oct(8);
        for (size_t q = 0; q < Ch.boxes.size(); ++q) {
            int p = P.index.at(tree::parent(Ch.boxes[q]));
            oct[Ch.boxes[q] & 7].push_back({P.off + (int64_t)p * P.per_box(), Ch.off + (int64_t)q * Ch.per_box()});
        }
        for (int b = 0; b < 8; ++b) {
            double dv[3] = {-sg * (((b >> 2) & 1) - 0.5) * Ch.edge, -sg * (((b >> 1) & 1) - 0.5) * Ch.edge,
                            -sg * ((b & 1) - 0.5) * Ch.edge};
            add_class(f->l2l[l], dv, k, 1, P.L, Ch.L, oct[b]);
            f->counts[3] += (int64_t)oct[b].size();
        }
    }
    // L2P: leaf box -> sphere (targets ascending)
    for (int i = 0; i < ns; ++i) {
        int q = LL.index.at(f->leaf_of[i]);
        double dv[3];
        dvec(&LL.centre[3 * q], &c->pos[3 * i], dv);
        f->l2p.tasks.push_back(make_task(dv, k, 1, LL.L, c->Leff[i], LL.off + q * LL.per_box(), c->off[i]));
    }
    f->counts[4] = (int64_t)f->l2p.tasks.size();
    // P2P: spheres of E2 = I2(leaf box), i != j; sources ascending
    {
        std::vector<std::vector<int>> members(LL.boxes.size());
        for (int i = 0; i < ns; ++i) members[LL.index.at(f->leaf_of[i])].push_back(i);
        for (int j = 0; j < ns; ++j) {
            if (!c->active[j]) continue;
            std::vector<int64_t> i2 = tree::domain_i2(f->leaf_of[j], lmax, 1);
            std::vector<int> src;
            for (int64_t b : i2) {
                auto it = LL.index.find(b);
                if (it == LL.index.end()) continue;
                for (int i : members[it->second])
                    if (i != j) src.push_back(i);
            }
            std::sort(src.begin(), src.end());
            for (int i : src) {
                double dv[3];
                dvec(&c->pos[3 * i], &c->pos[3 * j], dv);
                f->p2p.tasks.push_back(
                    make_task(dv, k, c->prm.z_selector, c->Leff[i], c->Leff[j], c->off[i], c->off[j]));
            }
        }
        f->counts[5] = (int64_t)f->p2p.tasks.size();
    }
    // rows of inactive spheres (reference row zeroing) get no far field either
    if (c->prm.ref_row_zeroing) {
        std::vector<Task> keep;
        for (auto &T : f->l2p.tasks) {
            bool act = false;
            for (int i = 0; i < ns; ++i)
                if (c->off[i] == T.dst) act = c->active[i];
            if (act) keep.push_back(T);
        }
        f->l2p.tasks.swap(keep);
    }

    auto fin = [&](Batch &B, const char *name) -> int {
        B.name = name;
        int r = batch_finalize(c, B, std::min<size_t>(c->ws_budget, (size_t)8 << 30), true);
        if (!r) batch_cost(B, c->flops_mv, c->bytes_mv);
        return r;
    };
    auto cfin = [&](ClassBatch &B, const char *name) -> int {
        B.name = name;
        B.src_stride = B.dst_stride = f->nC;
        int r = class_batch_finalize(c, B);
        if (!r) class_batch_cost(B, c->flops_mv, c->bytes_mv);
        return r;
    };
    // class batches first: their tables have the first claim on the cache budget
    {
        std::vector<ClassBatch *> cls;
        for (int l = lmin; l <= lmax; ++l) {
            if (l < lmax) {
                if ((rc = cfin(f->m2m[l], "m2m"))) goto fail;
                cls.push_back(&f->m2m[l]);
            }
            if ((rc = cfin(f->m2l[l], "m2l"))) goto fail;
            cls.push_back(&f->m2l[l]);
            if (l > lmin) {
                if ((rc = cfin(f->l2l[l], "l2l"))) goto fail;
                cls.push_back(&f->l2l[l]);
            }
        }
        if ((rc = class_batches_prepare(c, cls))) goto fail;
    }
    if ((rc = fin(f->p2m, "p2m"))) goto fail;
    if ((rc = fin(f->l2p, "l2p"))) goto fail;
    if ((rc = fin(f->p2p, "p2p"))) goto fail;
    *out = f;
    return IFS_OK;
fail:
    delete f;
    return rc;
}

void destroy(Fmm *f) { delete f; }

int run(ifs_ctx *c, Fmm *f, const double2 *xh, double2 *acc, int nrhs, int64_t ld, int prec) {
    int rc;
    // multipole / local pools hold nrhs copies at stride nC
    if ((rc = f->C.alloc(16ull * f->nC * nrhs)) || (rc = f->D.alloc(16ull * f->nC * nrhs))) return rc;
    double2 *C = f->C.as<double2>(), *D = f->D.as<double2>();
    const int64_t st = f->nC;
    launch_zero(C, f->nC * nrhs, c->stream);
    launch_zero(D, f->nC * nrhs, c->stream);
    if ((rc = batch_run(c, f->p2m, xh, C, nrhs, ld, st))) return rc;
    for (int l = f->l_max - 1; l >= f->l_min; --l)
        if ((rc = class_batch_run(c, f->m2m[l], C, C, nrhs, prec))) return rc;
    for (int l = f->l_min; l <= f->l_max; ++l) {
        if ((rc = class_batch_run(c, f->m2l[l], C, D, nrhs, prec))) return rc;
        if (l > f->l_min && (rc = class_batch_run(c, f->l2l[l], D, D, nrhs, prec))) return rc;
    }
    if ((rc = batch_run(c, f->l2p, D, acc, nrhs, st, ld))) return rc;
    if ((rc = batch_run(c, f->p2p, xh, acc, nrhs, ld, ld))) return rc;
    return IFS_OK;
}

uint64_t buffers_sig(Fmm *f) {
    uint64_t h = 0;
    auto mix = [&](uint64_t v) { h ^= v + 0x9e3779b97f4a7c15ull + (h << 6) + (h >> 2); };
    auto buf = [&](const DevBuf &d) { mix((uint64_t)(uintptr_t)d.p); };
    buf(f->C);
    buf(f->D);
    for (const Batch *B : {&f->p2m, &f->l2p, &f->p2p})
        for (const DevBuf *d : {&B->d_tasks, &B->d_segs, &B->own_d, &B->own_s, &B->own_ax, &B->own_y, &B->own_z})
            buf(*d);
    for (const auto *v : {&f->m2m, &f->m2l, &f->l2l})
        for (const ClassBatch &B : *v) {
            mix((uint64_t)B.plan_nrhs);
            for (const DevBuf *d : {&B.own_d, &B.own_ax, &B.d_dkeys, &B.d_desc, &B.d_gen, &B.d_src, &B.d_pidx,
                                    &B.d_phase, &B.d_scat, &B.d_scat_cols})
                buf(*d);
        }
    return h;
}

int tree_export(Fmm *f, int64_t *uindex_out, int32_t *levels_out) {
    for (size_t i = 0; i < f->leaf_of.size(); ++i) uindex_out[i] = f->leaf_of[i];
    levels_out[0] = f->l_min;
    levels_out[1] = f->l_max;
    return IFS_OK;
}

int stats(Fmm *f, int64_t counts[6], int32_t levels_L[32]) {
    for (int i = 0; i < 6; ++i) counts[i] = f->counts[i];
    for (int l = 0; l < 32; ++l) levels_L[l] = 0;
    for (int l = f->l_min; l <= f->l_max && l < 32; ++l) levels_L[l] = f->lv[l].L;
    return IFS_OK;
}

}  // namespace fmm
}  // namespace ifs
