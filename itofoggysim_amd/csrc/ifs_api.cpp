// Host orchestration and the C ABI (include/ifs.h).
//
// A context owns: the sphere data on the device, the Mie T-matrix, the
// translation plan (tasks grouped by target, split into workspace chunks), the
// workspace, and one HIP stream.  ifs_matvec_device runs
//     pack (Xu basis -> normalised compact) -> translations -> finalize (T-apply + identity)
// which replaces lib_mie_ms_solver_calculate_vector_b
// (FoggySim/src/lib/mie_theory/lib_mie_ms_solver_interface.f90:739-998).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <complex>
#include <cstring>
#include <map>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "../../include/ifs.h"
#include "ifs_engine.h"
#include "ifs_fmm.h"
#include "ifs_internal.h"

using namespace ifs;

#define HIP_OK(expr)                          \
    do {                                      \
        if ((expr) != hipSuccess) return IFS_E_HIP; \
    } while (0)

namespace {

int64_t fidx(int n, int m) { return (int64_t)n * n - 1 + m + n; }

// s(n,m) = sqrt(n(n+1)(2n+1) (n-m)!/(n+m)!) in long double
std::vector<double> make_snm(int N) {
    std::vector<double> s((size_t)N * (N + 2));
    for (int n = 1; n <= N; ++n)
        for (int m = -n; m <= n; ++m) {
            long double r = 1.0L;  // (n-m)!/(n+m)!
            if (m >= 0)
                for (int i = n - m + 1; i <= n + m; ++i) r /= (long double)i;
            else
                for (int i = n + m + 1; i <= n - m; ++i) r *= (long double)i;
            s[fidx(n, m)] = (double)sqrtl((long double)n * (n + 1) * (2 * n + 1) * r);
        }
    return s;
}

}  // namespace

Task ifs::make_task(const double *dv, double k, int kind, int Ls, int Lt, int64_t src, int64_t dst) {
    Task T{};
    double r = std::sqrt(dv[0] * dv[0] + dv[1] * dv[1] + dv[2] * dv[2]);
    T.src = src;
    T.dst = dst;
    T.t = k * r;
    T.cth = r > 0.0 ? std::max(-1.0, std::min(1.0, dv[2] / r)) : 1.0;
    T.phi = (dv[0] == 0.0 && dv[1] == 0.0) ? 0.0 : std::atan2(dv[1], dv[0]);
    T.Ls = Ls;
    T.Lt = Lt;
    T.K = std::min(Ls, Lt);
    T.kind = kind;
    return T;
}

namespace ifs {

// Profiler names per batch kind (static strings).
static void set_prof_names(Batch &B) {
    static const char *p2m[6] = {"p2m_gen_wigner", "p2m_gen_sect", "p2m_gen_axial", "p2m_rot_fwd", "p2m_axial", "p2m_rot_back"};
    static const char *l2p[6] = {"l2p_gen_wigner", "l2p_gen_sect", "l2p_gen_axial", "l2p_rot_fwd", "l2p_axial", "l2p_rot_back"};
    static const char *p2p[6] = {"p2p_gen_wigner", "p2p_gen_sect", "p2p_gen_axial", "p2p_rot_fwd", "p2p_axial", "p2p_rot_back"};
    const char **src = nullptr;
    if (B.name == "p2m") src = p2m;
    if (B.name == "l2p") src = l2p;
    if (B.name == "p2p") src = p2p;
    if (src)
        for (int i = 0; i < 6; ++i) B.pn[i] = src[i];
}

// Algorithmic flops of one generic task: rotation forward, coaxial shift, rotation back.
// The rotations use the folded tables: E (kk+1)x(n+1) and O kk x n per degree.
static void task_flops(const Task &T, double &ffwd, double &fax, double &fback) {
    ffwd = fax = fback = 0.0;
    auto rot = [&](int n) {
        double kk = std::min(n, T.K);
        return 8.0 * ((kk + 1.0) * (n + 1.0) + kk * n);
    };
    for (int n = 1; n <= T.Ls; ++n) ffwd += rot(n);
    for (int n = 1; n <= T.Lt; ++n) fback += rot(n);
    for (int k = -T.K; k <= T.K; ++k) fax += 16.0 * (double)ax_block(k, T.Lt, T.Ls);
}

// Split a target-sorted task list into segments and workspace chunks and upload it.
// The tables are kept across matvecs when they fit the context's cache budget.
int batch_finalize(ifs_ctx *c, Batch &B, size_t budget_bytes, bool allow_cache) {
    B.segs.clear();
    B.chunks.clear();
    set_prof_names(B);
    const size_t n = B.tasks.size();
    B.gq = n > 0;
    for (auto &T : B.tasks) B.gq = B.gq && T.kind == IFS_ZSEL_BESSEL_J;
    const bool gq = B.gq;
    // per-task sizes
    auto sizes = [gq](const Task &T, int64_t &nd, int64_t &nsct, int64_t &nax, int64_t &ny, int64_t &nz) {
        int L = std::max(T.Ls, T.Lt);
        nd = fold_size(L, T.K);
        nsct = gq ? gqb_alpha_size(T.Lt, T.Ls, T.K) : sect_size(T.K, T.Lt + T.Ls + 2);
        nax = ax_size(T.Lt, T.Ls, T.K);
        ny = 2 * rows_size(T.Ls, T.K);
        nz = 2 * rows_size(T.Lt, T.K);
    };
    int64_t tot_bytes = 0;
    for (auto &T : B.tasks) {
        int64_t nd, nsct, nax, ny, nz;
        sizes(T, nd, nsct, nax, ny, nz);
        tot_bytes += nd * 8 + (nsct + nax + ny + nz) * 16;
    }
    B.cached = allow_cache && (size_t)tot_bytes + c->ws_cached <= c->cache_budget;
    size_t chunk_budget = B.cached ? (size_t)tot_bytes + 1 : budget_bytes;

    Chunk cur;
    cur.t_begin = 0;
    cur.s_begin = 0;
    int64_t used = 0;
    auto close_chunk = [&](int t_end) {
        cur.t_end = t_end;
        cur.s_end = (int)B.segs.size();
        B.chunks.push_back(cur);
        cur = Chunk();
        cur.t_begin = t_end;
        cur.s_begin = (int)B.segs.size();
        used = 0;
    };
    for (size_t i = 0; i < n; ++i) {
        Task &T = B.tasks[i];
        int64_t nd, nsct, nax, ny, nz;
        sizes(T, nd, nsct, nax, ny, nz);
        int64_t bytes = nd * 8 + (nsct + nax + ny + nz) * 16;
        if (i > (size_t)cur.t_begin && used + bytes > (int64_t)chunk_budget) {
            // close the open segment at the chunk boundary
            close_chunk((int)i);
        }
        // segment bookkeeping
        if (B.segs.empty() || (int)B.segs.size() == cur.s_begin || B.segs.back().dst != T.dst ||
            B.segs.back().t_end != (int)i) {
            Segment S;
            S.dst = T.dst;
            S.Lt = T.Lt;
            S.t_begin = (int)i;
            S.t_end = (int)i + 1;
            B.segs.push_back(S);
        } else {
            B.segs.back().t_end = (int)i + 1;
        }
        T.off_d = cur.n_d;
        T.off_s = cur.n_s;
        T.off_ax = cur.n_ax;
        T.off_y = cur.n_y;
        T.off_z = cur.n_z;
        cur.n_d += nd;
        cur.n_s += nsct;
        cur.n_ax += nax;
        cur.n_y += ny;
        cur.n_z += nz;
        int L = std::max(T.Ls, T.Lt);
        cur.maxpairs = std::max(cur.maxpairs, (T.K + 1) * (L + 1));
        cur.maxK = std::max(cur.maxK, T.K);
        cur.max_nmax = std::max(cur.max_nmax, T.Lt + T.Ls + 2);
        cur.max_Ls = std::max(cur.max_Ls, T.Ls);
        cur.max_Lt = std::max(cur.max_Lt, T.Lt);
        cur.max_rows_y = std::max(cur.max_rows_y, frow_size(T.Ls, T.K));
        cur.max_rows_z = std::max(cur.max_rows_z, rows_size(T.Lt, T.K));
        if (gq) {
            GqbChunk &g = cur.gqb;
            g.Q = std::max(g.Q, gqb_nodes(T.t, T.Lt, T.Ls));
            g.Nr = std::max(g.Nr, std::max(T.Lt + 1, T.Ls));
            g.maxK = std::max(g.maxK, T.K);
            g.max_big = std::max(g.max_big, std::max(T.Lt + 1, T.Ls));
            g.max_small = std::max(g.max_small, std::min(T.Lt + 1, T.Ls));
            g.max_ent = std::max(g.max_ent, (int64_t)T.Lt * T.Ls);
        }
        int nent = T.Lt * (T.Lt + 3) / 2;  // back-rotation entries (n, m >= 0)
        cur.rowback_blocks = std::max(cur.rowback_blocks, std::min(64, (nent + 255) / 256));
        double f1, f2, f3;
        task_flops(T, f1, f2, f3);
        cur.f_fwd += f1;
        cur.f_ax += f2;
        cur.f_back += f3;
        used += bytes;
    }
    if (n > 0) close_chunk((int)n);
    // segments were appended in task order; chunk s ranges are consistent by construction
    // workspace sizes: max over chunks (reused) or sum (cached: single chunk)
    int64_t md = 0, ms = 0, ma = 0, my = 0, mz = 0;
    for (auto &ch : B.chunks) {
        md = std::max(md, ch.n_d);
        ms = std::max(ms, ch.n_s);
        ma = std::max(ma, ch.n_ax);
        my = std::max(my, ch.n_y);
        mz = std::max(mz, ch.n_z);
    }
    int rc;
    if (gq) {
        // shared Legendre table per chunk and the chunk's node set
        size_t vb = 0;
        for (auto &ch : B.chunks) {
            const double *mu, *w;
            if ((rc = gq_rule(c, ch.gqb.Q, &mu, &w)) != IFS_OK) return rc;
            vb = std::max(vb, (size_t)(ch.gqb.maxK + 1) * (ch.gqb.Nr + 1) * ch.gqb.Q * 8);
        }
        if ((rc = c->ws_gq.alloc(std::max(c->ws_gq.bytes, vb))) != IFS_OK) return rc;
    }
    if (B.cached) {
        if ((rc = B.own_d.alloc(md * 8)) || (rc = B.own_s.alloc(ms * 16)) || (rc = B.own_ax.alloc(ma * 16)) ||
            (rc = B.own_y.alloc(my * 16)) || (rc = B.own_z.alloc(mz * 16)))
            return rc;
        c->ws_cached += (size_t)(md * 8 + (ms + ma + my + mz) * 16);
    } else {
        if ((rc = c->ws_d.alloc(std::max<size_t>(c->ws_d.bytes, md * 8))) != IFS_OK) return rc;
        if ((rc = c->ws_s.alloc(std::max<size_t>(c->ws_s.bytes, ms * 16))) != IFS_OK) return rc;
        if ((rc = c->ws_ax.alloc(std::max<size_t>(c->ws_ax.bytes, ma * 16))) != IFS_OK) return rc;
        if ((rc = c->ws_y.alloc(std::max<size_t>(c->ws_y.bytes, my * 16))) != IFS_OK) return rc;
        if ((rc = c->ws_z.alloc(std::max<size_t>(c->ws_z.bytes, mz * 16))) != IFS_OK) return rc;
    }
    if ((rc = B.d_tasks.alloc(sizeof(Task) * std::max<size_t>(1, n))) != IFS_OK) return rc;
    if ((rc = B.d_segs.alloc(sizeof(Segment) * std::max<size_t>(1, B.segs.size()))) != IFS_OK) return rc;
    if (n) HIP_OK(hipMemcpy(B.d_tasks.p, B.tasks.data(), sizeof(Task) * n, hipMemcpyHostToDevice));
    if (!B.segs.empty())
        HIP_OK(hipMemcpy(B.d_segs.p, B.segs.data(), sizeof(Segment) * B.segs.size(), hipMemcpyHostToDevice));
    B.tables_ready = false;
    return IFS_OK;
}

Scope::Scope(ifs_ctx *c_, const char *n, double algorithmic_flops) : c(c_), name(n), flops(algorithmic_flops) {
    if (c->prof.on) {
        a = c->prof.get();
        (void)hipEventRecord(a, c->stream);
    }
}
Scope::~Scope() {
    if (c->prof.on) {
        b = c->prof.get();
        (void)hipEventRecord(b, c->stream);
        c->prof.recs.push_back({name, a, b, flops});
    }
}

int batch_run(ifs_ctx *c, Batch &B, const double2 *src_pool, double2 *dst_pool, int nrhs, int64_t src_stride,
              int64_t dst_stride) {
    const Task *dt = B.d_tasks.as<Task>();
    double *wd = B.cached ? B.own_d.as<double>() : c->ws_d.as<double>();
    double2 *wsc = B.cached ? B.own_s.as<double2>() : c->ws_s.as<double2>();
    double2 *wax = B.cached ? B.own_ax.as<double2>() : c->ws_ax.as<double2>();
    double2 *wy = B.cached ? B.own_y.as<double2>() : c->ws_y.as<double2>();
    double2 *wz = B.cached ? B.own_z.as<double2>() : c->ws_z.as<double2>();
    const Segment *dsg = B.d_segs.as<Segment>();
    for (auto &ch : B.chunks) {
        int nt = ch.t_end - ch.t_begin;
        int nsg = ch.s_end - ch.s_begin;
        if (nt <= 0) continue;
        const Task *tt = dt + ch.t_begin;
        // one right-hand side and tables not kept: the coaxial coefficients are applied as
        // they are generated and never stored (the table round trip through HBM is what
        // the separate generate/apply pair spends its time on)
        const bool fused = !B.gq && nrhs == 1 && !B.cached &&
                           ch.max_Lt < (ch.max_nmax + 2 <= 128 ? 256 : 1024) && ch.max_nmax <= 380;  // LDS <= 64 KiB
        // sphere <-> box tables not kept: the quadrature is applied to the vectors directly
        // ((Lt + Ls) Q work per order instead of the table's Lt Ls Q), no table is formed
        const bool gq_apply = B.gq && !B.cached && gqb_apply_lds(ch.gqb.Q, ch.gqb.Nr) <= 65536;
        const double *gmu = nullptr, *gw = nullptr;
        if (B.gq) (void)gq_rule(c, ch.gqb.Q, &gmu, &gw);  // created at finalize
        if (gq_apply) {
            {
                Scope s(c, B.pn[0]);
                launch_gen_wigner_fold(tt, nt, ch.maxpairs, wd, c->stream);
            }
            Scope s(c, B.pn[1]);
            launch_gqb_legendre(ch.gqb, gmu, c->ws_gq.as<double>(), c->stream);
        } else if (!(B.cached && B.tables_ready)) {
            {
                Scope s(c, B.pn[0]);
                launch_gen_wigner_fold(tt, nt, ch.maxpairs, wd, c->stream);
            }
            if (B.gq) {
                Scope s(c, B.pn[2]);
                const double *mu = nullptr, *w = nullptr;
                (void)gq_rule(c, ch.gqb.Q, &mu, &w);  // created at finalize
                launch_gen_axial_gqb(tt, nt, ch.gqb, mu, w, c->ws_gq.as<double>(), wsc, wax, c->stream);
            } else {
                Scope s(c, B.pn[1]);
                launch_gen_sect(tt, nt, ch.max_nmax, wsc, c->stream);
            }
            if (!fused && !B.gq) {
                Scope s(c, B.pn[2]);
                launch_gen_axial(tt, nt, ch.maxK, ch.max_nmax, wsc, wax, c->stream);
            }
        }
        // the tables serve every right-hand side
        for (int r = 0; r < nrhs; ++r) {
            const double2 *sp = src_pool + r * src_stride;
            double2 *dp = dst_pool + r * dst_stride;
            {
                Scope s(c, B.pn[3], ch.f_fwd);
                if (ch.max_Ls >= 128)
                    launch_rot_fwd_wide(tt, nt, ch.max_Ls, sp, wd, wy, c->stream);
                else
                    launch_rot_fwd(tt, nt, ch.max_rows_y, ch.max_Ls, sp, wd, wy, c->stream);
            }
            if (fused) {
                Scope s(c, B.pn[4], ch.f_ax);
                launch_axial_fused(tt, nt, ch.maxK, ch.max_nmax, ch.max_Lt, wsc, wy, wz, c->stream);
            } else if (gq_apply) {
                Scope s(c, B.pn[4], ch.f_ax);
                launch_gqb_apply(tt, nt, ch.gqb, gmu, gw, c->ws_gq.as<double>(), wy, wz, c->stream);
            } else {
                Scope s(c, B.pn[4], ch.f_ax);
                launch_axial_apply(tt, nt, ch.max_rows_z, wax, wy, wz, c->stream);
            }
            {
                Scope s(c, B.pn[5], ch.f_back);
                launch_rot_back(dt, dsg + ch.s_begin, nsg, ch.rowback_blocks, wd, wz, dp, c->stream);
            }
        }
        if (hipPeekAtLastError() != hipSuccess) return IFS_E_HIP;
    }
    if (B.cached) B.tables_ready = true;
    return IFS_OK;
}

// Per-column algorithmic flops of the three products (folded rotations, decoupled
// coaxial shift): forward rotation, coaxial shift, back rotation.
static void class_col_flops(const Task &T, double &ffwd, double &fax, double &fback) {
    ffwd = fax = fback = 0.0;
    // folded rotation: E (kk+1)x(n+1) and O kk x n (its zero row and column excluded)
    auto rot = [&](int n) {
        double kk = std::min(n, T.K);
        return 8.0 * ((kk + 1.0) * (n + 1.0) + kk * n);
    };
    for (int n = 1; n <= T.Ls; ++n) ffwd += rot(n);
    for (int n = 1; n <= T.Lt; ++n) fback += rot(n);
    for (int k = -T.K; k <= T.K; ++k) fax += 16.0 * (double)ax_block(k, T.Lt, T.Ls);
}

static bool same_d(const Task &a, const Task &b) { return a.cth == b.cth && a.Ls == b.Ls && a.Lt == b.Lt && a.K == b.K; }
static bool same_ax(const Task &a, const Task &b) {
    return a.t == b.t && a.kind == b.kind && a.Ls == b.Ls && a.Lt == b.Lt && a.K == b.K;
}

static int class_batch_keys(ifs_ctx *c, ClassBatch &B);

int class_batch_finalize(ifs_ctx *c, ClassBatch &B) {
    const int nc = (int)B.gen.size();
    B.chunks.clear();
    B.descs.clear();
    B.scat.clear();
    B.scat_cols.clear();
    if (nc == 0) return IFS_OK;
    const Task &T0 = B.gen[0];
    for (auto &T : B.gen)
        if (T.Ls != T0.Ls || T.Lt != T0.Lt || T.K != T0.K) return IFS_E_ARG;  // one level pair per batch
    // order classes by (radial kind, length, polar angle, azimuth): axial segments and
    // rotation runs become consecutive
    std::vector<int64_t> cb(nc + 1, 0);
    for (int i = 0; i < nc; ++i) cb[i + 1] = cb[i] + B.ncols[i];
    std::vector<int> ord(nc);
    for (int i = 0; i < nc; ++i) ord[i] = i;
    std::stable_sort(ord.begin(), ord.end(), [&](int a, int b) {
        const Task &x = B.gen[a], &y = B.gen[b];
        if (x.kind != y.kind) return x.kind < y.kind;
        if (x.t != y.t) return x.t < y.t;
        // polar angles by |cos|: the runs of theta and pi - theta share one Wigner table
        // and come out adjacent
        if (std::fabs(x.cth) != std::fabs(y.cth)) return std::fabs(x.cth) < std::fabs(y.cth);
        if (x.cth != y.cth) return x.cth < y.cth;
        return x.phi < y.phi;
    });
    std::vector<Task> gen(nc);
    std::vector<int> ncols(nc);
    std::vector<int64_t> src, dst;
    B.pidx.clear();
    for (int i = 0; i < nc; ++i) {
        int o = ord[i];
        gen[i] = B.gen[o];
        ncols[i] = B.ncols[o];
        for (int64_t j = cb[o]; j < cb[o + 1]; ++j) {
            src.push_back(B.src[j]);
            dst.push_back(B.dst[j]);
            B.pidx.push_back(i);
        }
    }
    B.gen.swap(gen);
    B.ncols.swap(ncols);
    B.src.swap(src);
    B.dst.swap(dst);
    // phase rows exp(i m phi), m = -PL..PL, evaluated in long double
    B.PL = std::max(T0.Ls, T0.Lt);
    const int PW = 2 * B.PL + 1;
    B.phase.assign((size_t)nc * PW, make_double2(0.0, 0.0));
    for (int i = 0; i < nc; ++i)
        for (int m = -B.PL; m <= B.PL; ++m) {
            long double a = (long double)m * (long double)B.gen[i].phi;
            B.phase[(size_t)i * PW + m + B.PL] = make_double2((double)cosl(a), (double)sinl(a));
        }
    B.plan_nrhs = 0;
    return class_batch_keys(c, B);
}

// Chunks, runs, segments and reduction lists for nrhs right-hand sides per call.
// Gauss-Legendre nodes and weights on [-1, 1] by Newton's iteration on P_Q in long
// double (Golub-Welsch would do as well; Q <= a few thousand here).
static void gauss_legendre(int Q, std::vector<double> &x, std::vector<double> &w) {
    x.assign(Q, 0.0);
    w.assign(Q, 0.0);
    const long double pi = 3.14159265358979323846264338327950288L;
    for (int i = 0; i < (Q + 1) / 2; ++i) {
        long double z = cosl(pi * (i + 0.75L) / (Q + 0.5L)), dp = 1.0L;
        for (int it = 0; it < 100; ++it) {
            long double p0 = 1.0L, p1 = z;
            for (int k = 2; k <= Q; ++k) {
                long double p2 = ((2 * k - 1) * z * p1 - (k - 1) * p0) / k;
                p0 = p1;
                p1 = p2;
            }
            if (Q == 1) p0 = 1.0L;
            dp = Q * (z * p1 - p0) / (z * z - 1.0L);
            long double dz = p1 / dp;
            z -= dz;
            if (fabsl(dz) < 1e-19L) break;
        }
        long double wt = 2.0L / ((1.0L - z * z) * dp * dp);
        x[i] = (double)-z;
        x[Q - 1 - i] = (double)z;
        w[i] = w[Q - 1 - i] = (double)wt;
    }
}

int gq_rule(ifs_ctx *c, int Q, const double **mu, const double **w) {
    auto it = c->gq_nodes.find(Q);
    if (it == c->gq_nodes.end()) {
        std::vector<double> x, wt;
        gauss_legendre(Q, x, wt);
        DevBuf &d = c->gq_nodes[Q];
        int rc = d.alloc(16 * (size_t)Q);
        if (rc) return rc;
        if (hipMemcpy(d.p, x.data(), 8 * (size_t)Q, hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(d.as<double>() + Q, wt.data(), 8 * (size_t)Q, hipMemcpyHostToDevice) != hipSuccess)
            return IFS_E_HIP;
        it = c->gq_nodes.find(Q);
    }
    *mu = it->second.as<double>();
    *w = it->second.as<double>() + Q;
    return IFS_OK;
}

// Table keys of a finalized class batch: one Wigner table per |cos(theta)|, one
// coaxial table per (length, kind), with the quadrature plan of each coaxial table.
static int class_batch_keys(ifs_ctx *c, ClassBatch &B) {
    const Task &T0 = B.gen[0];
    B.n_d = fold_size(std::max(T0.Ls, T0.Lt), T0.K);
    B.n_s = sect_size(T0.K, T0.Lt + T0.Ls + 2);
    B.n_ax = ax_size(T0.Lt, T0.Ls, T0.K);
    // table keys (Ls, Lt, K are common to the batch)
    {
        std::map<double, int> dk;
        std::map<std::pair<double, int>, int> ak;
        B.dkey.assign(B.gen.size(), 0);
        B.axkey.assign(B.gen.size(), 0);
        B.dkeys.clear();
        B.axkey_cls.clear();
        for (size_t g = 0; g < B.gen.size(); ++g) {
            const Task &T = B.gen[g];
            const double a = std::fabs(T.cth);
            auto it = dk.find(a);
            if (it == dk.end()) {
                it = dk.emplace(a, (int)B.dkeys.size()).first;
                Task K = T;
                K.cth = a;
                B.dkeys.push_back(K);
            }
            B.dkey[g] = it->second;
            auto jt = ak.find({T.t, T.kind});
            if (jt == ak.end()) {
                jt = ak.emplace(std::make_pair(T.t, T.kind), (int)B.axkey_cls.size()).first;
                B.axkey_cls.push_back((int)g);
            }
            B.axkey[g] = jt->second;
        }
    }
    // coaxial tables by quadrature (stable at every order)
    B.gq.clear();
    for (int g : B.axkey_cls) {
        const Task &T = B.gen[g];
        B.gq.push_back(gq_plan(T.kind, T.t, T0.Lt, T0.Ls, T0.K, (size_t)1 << 30));
        const double *mu, *w;
        int r = gq_rule(c, B.gq.back().Q, &mu, &w);
        if (r) return r;
    }
    return IFS_OK;
}

// Tables kept across matvecs, decided over all class batches at once: coaxial
// tables first (largest first: they are the dearest to regenerate), then Wigner
// tables, each batch's set only as a whole and only while the cache budget lasts.
int class_batches_prepare(ifs_ctx *c, const std::vector<ClassBatch *> &bs) {
    std::vector<ClassBatch *> ord;
    for (auto *B : bs)
        if (!B->gen.empty()) ord.push_back(B);
    auto ax_bytes = [](const ClassBatch *B) { return (size_t)B->n_ax * 16 * B->axkey_cls.size(); };
    auto d_bytes = [](const ClassBatch *B) { return (size_t)B->n_d * 8 * B->dkeys.size(); };
    std::stable_sort(ord.begin(), ord.end(), [&](ClassBatch *a, ClassBatch *b) { return ax_bytes(a) > ax_bytes(b); });
    for (auto *B : ord) {
        const size_t need = ax_bytes(B);
        if (B->ax_persist || !need || c->ws_cached + need > c->cache_budget) continue;
        if (B->own_ax.alloc(need) != IFS_OK) {
            (void)hipGetLastError();
            continue;
        }
        c->ws_cached += need;
        B->ax_persist = true;
        B->ax_ready.assign(B->axkey_cls.size(), 0);
    }
    std::stable_sort(ord.begin(), ord.end(), [&](ClassBatch *a, ClassBatch *b) { return d_bytes(a) > d_bytes(b); });
    for (auto *B : ord) {
        const size_t need = d_bytes(B);
        if (B->d_persist || !need || c->ws_cached + need > c->cache_budget) continue;
        if (B->own_d.alloc(need) != IFS_OK) {
            (void)hipGetLastError();
            continue;
        }
        c->ws_cached += need;
        B->d_persist = true;
        B->d_ready.assign(B->dkeys.size(), 0);
    }
    for (auto *B : ord) {
        int rc = class_batch_plan(c, *B, 1);
        if (rc) return rc;
    }
    return IFS_OK;
}

int class_batch_plan(ifs_ctx *c, ClassBatch &B, int nrhs) {
    B.chunks.clear();
    B.descs.clear();
    B.scat.clear();
    B.scat_cols.clear();
    if (B.gen.empty()) return IFS_OK;
    const Task &T0 = B.gen[0];
    // chunk capacity from the workspace budget (Y/W and Z per column and RHS)
    const int64_t ycol = 4 * std::max<int64_t>(rows_size(T0.Ls, T0.K), (int64_t)T0.Lt * (T0.Lt + 2));
    const int64_t zcol = 4 * rows_size(T0.Lt, T0.K);
    const int64_t per_col = 8 * (ycol + zcol) * nrhs;
    const int64_t total = (int64_t)B.src.size();
    int64_t cap = std::max<int64_t>(1, (int64_t)(c->ws_budget / (size_t)per_col));
    cap = std::min<int64_t>(cap, total);
    // the class kernels address one degree's 2K+1 rows of Y and Z with 32-bit offsets
    cap = std::min<int64_t>(cap, std::max<int64_t>(1, (int64_t)INT32_MAX / (4LL * nrhs * (2LL * T0.K + 2))));
    B.n_y = B.n_z = 0;
    // walk the class columns, cutting chunks of at most cap columns
    const int64_t nbox = 2LL * T0.Lt * (T0.Lt + 2);
    int cls = 0;
    int64_t done_in_cls = 0;
    for (int64_t c0 = 0; c0 < total; c0 += cap) {
        ClassBatch::ChunkPlan ch;
        ch.col_begin = (int)c0;
        ch.ncols = (int)std::min<int64_t>(cap, total - c0);
        // pieces: class fragments inside the chunk
        std::vector<std::pair<int, std::pair<int, int>>> frags;  // class, [begin, end) global columns
        int64_t col = c0;
        while (col < c0 + ch.ncols) {
            int64_t left = B.ncols[cls] - done_in_cls;
            int64_t take = std::min<int64_t>(left, c0 + ch.ncols - col);
            frags.push_back({cls, {(int)col, (int)(col + take)}});
            col += take;
            done_in_cls += take;
            if (done_in_cls == B.ncols[cls]) {
                ++cls;
                done_in_cls = 0;
            }
        }
        auto piece = [&](int g, int b, int e) {
            ClassBatch::Piece P;
            P.gen = g;
            GroupDesc &d = P.g;
            d = GroupDesc{};
            d.Ls = T0.Ls;
            d.Lt = T0.Lt;
            d.K = T0.K;
            d.ncols = e - b;
            d.col_begin = b;
            d.ycol0 = b - (int)c0;
            d.ncy = ch.ncols;
            d.PL = B.PL;
            d.nrhs = nrhs;
            d.src_stride = B.src_stride;
            d.dst_stride = B.dst_stride;
            return P;
        };
        // rotation runs: consecutive fragments with one polar angle and one length.  A run
        // never crosses an axial segment: with nrhs > 1 a piece lays out its columns as
        // [right-hand side][translation] from its own first column, so the coaxial shift
        // (which applies one table to every column of its segment, whatever their order)
        // reads and writes exactly the columns the runs inside it wrote and read back only
        // when the runs partition the segment.  (Consecutive classes of one polar angle
        // and two lengths occur where a level holds few offsets.)
        for (size_t i = 0; i < frags.size();) {
            size_t j = i + 1;
            while (j < frags.size() && same_d(B.gen[frags[j].first], B.gen[frags[i].first]) &&
                   same_ax(B.gen[frags[j].first], B.gen[frags[i].first]))
                ++j;
            ch.runs.push_back(piece(frags[i].first, frags[i].second.first, frags[j - 1].second.second));
            i = j;
        }
        // axial segments: consecutive fragments with one length and kind
        for (size_t i = 0; i < frags.size();) {
            size_t j = i + 1;
            while (j < frags.size() && same_ax(B.gen[frags[j].first], B.gen[frags[i].first])) ++j;
            ch.segs.push_back(piece(frags[i].first, frags[i].second.first, frags[j - 1].second.second));
            i = j;
        }
        for (auto &P : ch.runs) {
            P.di = (int)B.descs.size();
            B.descs.push_back(P.g);
        }
        for (auto &P : ch.segs) {
            P.di = (int)B.descs.size();
            B.descs.push_back(P.g);
        }
        // per-target reduction lists (chunk-local columns, ascending per target)
        std::vector<std::pair<int64_t, int32_t>> td;
        for (int j = 0; j < ch.ncols; ++j) td.push_back({B.dst[c0 + j], j});
        std::sort(td.begin(), td.end());
        ch.scat_begin = (int)B.scat.size();
        for (size_t i = 0; i < td.size();) {
            size_t j = i;
            ScatterSeg S;
            S.dst = td[i].first;
            S.beg = (int32_t)B.scat_cols.size();
            while (j < td.size() && td[j].first == td[i].first) B.scat_cols.push_back(td[j++].second);
            S.end = (int32_t)B.scat_cols.size();
            B.scat.push_back(S);
            i = j;
        }
        ch.scat_end = (int)B.scat.size();
        B.chunks.push_back(ch);
        B.n_y = std::max<int64_t>(B.n_y, ycol * ch.ncols * nrhs);
        B.n_z = std::max<int64_t>(B.n_z, zcol * ch.ncols * nrhs);
    }
    (void)nbox;
    const int L = std::max(T0.Ls, T0.Lt);
    B.n_d = fold_size(L, T0.K);
    int rc;
    size_t gq_ws = 0;
    for (auto &G : B.gq) gq_ws = std::max(gq_ws, G.ws_bytes);
    // otherwise Wigner table slots: the tables of a chunk's runs stay for the back pass
    // when a quarter of the workspace budget holds them (run r uses slot r % nslots)
    size_t maxruns = 1;
    for (auto &ch : B.chunks) maxruns = std::max(maxruns, ch.runs.size());
    const size_t dbytes = (size_t)B.n_d * 8;
    B.n_dslots = B.d_persist ? 0
                             : (int)std::max<size_t>(1, std::min(maxruns, (c->ws_budget / 4) / std::max<size_t>(dbytes, 1)));
    for (auto &ch : B.chunks) {
        // consecutive runs with one Wigner key share a slot; a new key takes the next slot
        int slot = -1, prev_key = -1;
        for (size_t r = 0; r < ch.runs.size(); ++r) {
            auto &P = ch.runs[r];
            const int key = B.dkey[P.gen];
            if (key != prev_key) {
                slot = B.n_dslots > 0 ? (slot + 1) % B.n_dslots : 0;
                prev_key = key;
            }
            P.g.off_d = B.d_persist ? (int64_t)key * B.n_d : (int64_t)slot * B.n_d;
            P.g.mirror = B.gen[P.gen].cth < 0.0 ? 1 : 0;
            B.descs[P.di] = P.g;
        }
        for (auto &P : ch.segs) {
            P.g.off_ax = B.ax_persist ? (int64_t)B.axkey[P.gen] * B.n_ax : 0;
            B.descs[P.di] = P.g;
        }
    }
    if ((rc = c->ws_d.alloc(std::max<size_t>(c->ws_d.bytes, B.n_d * 8 * (size_t)B.n_dslots))) ||
        (rc = c->ws_gq.alloc(std::max<size_t>(c->ws_gq.bytes, gq_ws))) ||
        (rc = c->ws_ax.alloc(std::max<size_t>(c->ws_ax.bytes, B.ax_persist ? 16 : B.n_ax * 16))) ||
        (rc = c->ws_y.alloc(std::max<size_t>(c->ws_y.bytes, B.n_y * 8))) ||
        (rc = c->ws_z.alloc(std::max<size_t>(c->ws_z.bytes, B.n_z * 8))))
        return rc;
    auto up = [](DevBuf &d, const void *h, size_t bytes) -> int {
        int r = d.alloc(std::max<size_t>(bytes, 16));
        if (r) return r;
        if (bytes && hipMemcpy(d.p, h, bytes, hipMemcpyHostToDevice) != hipSuccess) return IFS_E_HIP;
        return IFS_OK;
    };
    if ((rc = up(B.d_desc, B.descs.data(), sizeof(GroupDesc) * B.descs.size())) ||
        (rc = up(B.d_gen, B.gen.data(), sizeof(Task) * B.gen.size())) ||
        (rc = up(B.d_dkeys, B.dkeys.data(), sizeof(Task) * B.dkeys.size())) ||
        (rc = up(B.d_src, B.src.data(), 8 * B.src.size())) || (rc = up(B.d_pidx, B.pidx.data(), 4 * B.pidx.size())) ||
        (rc = up(B.d_phase, B.phase.data(), sizeof(double2) * B.phase.size())) ||
        (rc = up(B.d_scat, B.scat.data(), sizeof(ScatterSeg) * B.scat.size())) ||
        (rc = up(B.d_scat_cols, B.scat_cols.data(), 4 * B.scat_cols.size())))
        return rc;
    B.plan_nrhs = nrhs;
    return IFS_OK;
}

// Rotation block heights chosen by timing, per (Ls, Lt, K, precision, right-hand sides,
// log2 of the translation count), shared by every context of the process.
static std::mutex g_rows_mu;
static std::map<std::array<int, 6>, std::array<int, 2>> g_rows;
static int ilog2(int64_t v) {
    int r = 0;
    while (v > 1) {
        v >>= 1;
        ++r;
    }
    return r;
}
static int tuned_rows(const std::array<int, 6> &k, int st) {
    std::lock_guard<std::mutex> lk(g_rows_mu);
    auto it = g_rows.find(k);
    return it == g_rows.end() ? 0 : it->second[st];
}
static void remember_rows(const std::array<int, 6> &k, int st, int r) {
    std::lock_guard<std::mutex> lk(g_rows_mu);
    g_rows[k][st] = r;
}

int class_batch_run(ifs_ctx *c, ClassBatch &B, const double2 *src_pool, double2 *dst_pool, int nrhs, int prec) {
    if (B.gen.empty() || nrhs <= 0) return IFS_OK;
    if (B.plan_nrhs != nrhs) {
        int rc = class_batch_plan(c, B, nrhs);
        if (rc) return rc;
    }
    const GroupDesc *dd = B.d_desc.as<GroupDesc>();
    const Task *dg = B.d_gen.as<Task>();
    const Task *dk = B.d_dkeys.as<Task>();
    const int64_t *dsrc = B.d_src.as<int64_t>();
    const int32_t *dpidx = B.d_pidx.as<int32_t>();
    const double2 *dph = B.d_phase.as<double2>();
    double *wd = B.d_persist ? B.own_d.as<double>() : c->ws_d.as<double>();
    double2 *wsc = c->ws_s.as<double2>();
    double2 *wax = B.ax_persist ? B.own_ax.as<double2>() : c->ws_ax.as<double2>();
    double *wy = c->ws_y.as<double>();
    double *wz = c->ws_z.as<double>();
    const Task &T0 = B.gen[0];
    double ffwd, fax, fback;
    class_col_flops(T0, ffwd, fax, fback);
    const int L = std::max(T0.Ls, T0.Lt);
    const int64_t nbox = 2LL * T0.Lt * (T0.Lt + 2);
    // the FP32 products take Wigner tables regenerated per matvec in FP32 (half the bytes to
    // store and to read); tables kept across matvecs stay FP64 and are converted at the load
    const int ftab = prec == 1 && !B.d_persist ? 1 : 0;
    int cur_ax = -1;                                        // coaxial key in the shared workspace
    std::vector<int> owner(std::max(B.n_dslots, 1), -1);    // Wigner key in each shared slot
    auto need_d = [&](const ClassBatch::Piece &P) {
        const int key = B.dkey[P.gen];
        if (B.d_persist) {
            if (B.d_ready[key]) return;
            B.d_ready[key] = 1;
        } else {
            int slot = (int)(P.g.off_d / std::max<int64_t>(B.n_d, 1));
            if (owner[slot] == key) return;
            owner[slot] = key;
        }
        if (ftab) {
            Scope s(c, "class_gen_wigner_f32");
            launch_gen_wigner_fold_f32(dk + key, 1, (T0.K + 1) * (L + 1), (float *)wd + P.g.off_d, c->stream);
        } else {
            Scope s(c, "class_gen_wigner");
            launch_gen_wigner_fold(dk + key, 1, (T0.K + 1) * (L + 1), wd + P.g.off_d, c->stream);
        }
    };
    auto need_ax = [&](const ClassBatch::Piece &P) {
        const int key = B.axkey[P.gen];
        if (B.ax_persist) {
            if (B.ax_ready[key]) return;
            B.ax_ready[key] = 1;
        } else {
            if (cur_ax == key) return;
            cur_ax = key;
        }
        Scope s(c, "class_gen_axial");
        const GqPlan &G = B.gq[key];
        const double *mu = nullptr, *w = nullptr;
        (void)gq_rule(c, G.Q, &mu, &w);  // created at plan time
        launch_gen_axial_gq(G, mu, w, c->ws_gq.p, wax + P.g.off_ax, c->stream);
    };
    const char *pn_fwd = prec == 1 ? "class_rot_fwd_f32" : "class_rot_fwd";
    const char *pn_ax = prec == 1 ? "class_axial_f32" : "class_axial";
    const char *pn_back = prec == 1 ? "class_rot_back_f32" : "class_rot_back";
    // Rotation block heights (64 or 32 rows of each folded product).  64 halves the
    // column-panel re-reads; 32 wastes less padding where the degrees hold few rows
    // (small orders) and gives more blocks to short runs.  A batch's first product
    // times both on every run launch (they write the same Y or W), keeps the faster per
    // product, and remembers the choice per shape for later contexts.
    int rows[2];
    bool tune[2] = {false, false};
    hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
    (void)hipStreamIsCapturing(c->stream, &cap);
    const bool can_tune = c->autotune && !c->prof.on && cap == hipStreamCaptureStatusNone;
    const std::array<int, 6> tkey = {T0.Ls, T0.Lt, T0.K, prec, nrhs, ilog2((int64_t)B.src.size())};
    for (int st = 0; st < 2; ++st) {
        int r = c->force_rows[st] ? c->force_rows[st] : B.rot_rows[prec & 1][st];
        if (!r) r = tuned_rows(tkey, st);
        if (!r && can_tune) tune[st] = true;
        rows[st] = r ? r : 64;
    }
    std::vector<std::array<hipEvent_t, 3>> tev[2];
    auto launch_rot = [&](const ClassBatch::Piece &P, int st) {
        const int stage = st == 0 ? 0 : 2;
        if (!tune[st]) {
            launch_class(P.g, dd + P.di, dsrc, dpidx, dph, src_pool, wd, wax, wy, wz, stage, prec, ftab, rows[st],
                         c->stream);
            return;
        }
        std::array<hipEvent_t, 3> e{};
        for (auto &x : e) (void)hipEventCreate(&x);
        (void)hipEventRecord(e[0], c->stream);
        launch_class(P.g, dd + P.di, dsrc, dpidx, dph, src_pool, wd, wax, wy, wz, stage, prec, ftab, 64, c->stream);
        (void)hipEventRecord(e[1], c->stream);
        launch_class(P.g, dd + P.di, dsrc, dpidx, dph, src_pool, wd, wax, wy, wz, stage, prec, ftab, 32, c->stream);
        (void)hipEventRecord(e[2], c->stream);
        tev[st].push_back(e);
    };
    for (auto &ch : B.chunks) {
        for (auto &P : ch.runs) {
            need_d(P);
            Scope s(c, pn_fwd, ffwd * P.g.ncols * nrhs);
            launch_rot(P, 0);
        }
        for (auto &P : ch.segs) {
            need_ax(P);
            if (P.g.ncols > 0 && P.g.nrhs > 0) {
                int nfull;
                bool tail;
                axial_tile_split(P.g.ncols * P.g.nrhs, nfull, tail);
                c->axial_tiles[0] += nfull > 0;
                c->axial_tiles[1] += tail;
            }
            Scope s(c, pn_ax, fax * P.g.ncols * nrhs);
            launch_class(P.g, dd + P.di, dsrc, dpidx, dph, src_pool, wd, wax, wy, wz, 1, prec, ftab, 0, c->stream);
        }
        for (size_t r = ch.runs.size(); r-- > 0;) {
            auto &P = ch.runs[r];
            need_d(P);
            Scope s(c, pn_back, fback * P.g.ncols * nrhs);
            launch_rot(P, 1);
        }
        {
            Scope s(c, "class_scatter");
            launch_scatter(B.d_scat.as<ScatterSeg>() + ch.scat_begin, ch.scat_end - ch.scat_begin,
                           B.d_scat_cols.as<int32_t>(), (const double2 *)wy, nbox, dst_pool, ch.ncols, nrhs,
                           B.dst_stride, c->stream);
        }
    }
    if (hipPeekAtLastError() != hipSuccess) return IFS_E_HIP;
    for (int st = 0; st < 2; ++st) {
        if (!tune[st]) continue;
        double t64 = 0.0, t32 = 0.0;
        bool ok = !tev[st].empty() && hipEventSynchronize(tev[st].back()[2]) == hipSuccess;
        for (auto &e : tev[st]) {
            float a = 0.f, b = 0.f;
            ok = ok && hipEventElapsedTime(&a, e[0], e[1]) == hipSuccess &&
                 hipEventElapsedTime(&b, e[1], e[2]) == hipSuccess;
            t64 += a;
            t32 += b;
            for (auto &x : e) (void)hipEventDestroy(x);
        }
        if (!ok) continue;
        // 32 rows only when clearly faster (the two agree to within noise on large orders)
        const int r = t32 < 0.97 * t64 ? 32 : 64;
        B.rot_rows[prec & 1][st] = r;
        remember_rows(tkey, st, r);
    }
    return IFS_OK;
}

void class_batch_cost(const ClassBatch &B, double &flops, double &bytes) {
    if (B.gen.empty()) return;
    const Task &T = B.gen[0];
    double ffwd, fax, fback;
    class_col_flops(T, ffwd, fax, fback);
    const double ncol = (double)B.src.size();
    flops += (ffwd + fax + fback) * ncol;
    bytes += (4.0 * T.Ls * (T.Ls + 2) * 16.0 * 0.5 + 4.0 * T.Lt * (T.Lt + 2) * 16.0 * 0.5) * ncol;
}

void batch_cost(const Batch &B, double &flops, double &bytes) {
    for (auto &T : B.tasks) {
        double f1, f2, f3;
        task_flops(T, f1, f2, f3);
        flops += f1 + f2 + f3;
        bytes += 4.0 * T.Ls * (T.Ls + 2) * 16.0 * 0.5 + 4.0 * T.Lt * (T.Lt + 2) * 16.0 * 0.5;
    }
}

hipStream_t ctx_stream(ifs_ctx *c) { return c->stream; }
int ctx_ns(ifs_ctx *c) { return c->ns; }
int ctx_N(ifs_ctx *c) { return c->N; }
const double *ctx_pos(ifs_ctx *c) { return c->pos.data(); }
double ctx_k(ifs_ctx *c) { return c->prm.k; }

double2 *c_bhat(ifs_ctx *c, int64_t n) { return c->k_bhat.alloc(16ull * n) ? nullptr : c->k_bhat.as<double2>(); }
double2 *c_xhat(ifs_ctx *c, int64_t n) { return c->k_xhat.alloc(16ull * n) ? nullptr : c->k_xhat.as<double2>(); }
double2 *c_stage(ifs_ctx *c, int64_t n) { return c->k_stage.alloc(16ull * n) ? nullptr : c->k_stage.as<double2>(); }

int ctx_krylov_alloc(ifs_ctx *c, int restart, int64_t n, double2 **V, double2 **w, double2 **r, double2 **part,
                     double2 **h, int nblocks) {
    int rc;
    if ((rc = c->k_V.alloc(16ull * (restart + 1) * n)) || (rc = c->k_w.alloc(16ull * n)) ||
        (rc = c->k_r.alloc(16ull * n)) || (rc = c->k_part.alloc(16ull * nblocks * (restart + 1))) ||
        (rc = c->k_h.alloc(16ull * (restart + 1))))
        return rc;
    *V = c->k_V.as<double2>();
    *w = c->k_w.as<double2>();
    *r = c->k_r.as<double2>();
    *part = c->k_part.as<double2>();
    *h = c->k_h.as<double2>();
    return IFS_OK;
}

}  // namespace ifs

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
extern "C" {

static int create_impl(ifs_ctx **out, const ifs_cluster *cl, const ifs_params *p, const int32_t *group,
                       int32_t nclusters);

int ifs_create(ifs_ctx **out, const ifs_cluster *cl, const ifs_params *p) {
    return create_impl(out, cl, p, nullptr, 1);
}

int ifs_create_multi(ifs_ctx **out, int32_t nclusters, const ifs_cluster *cls, const ifs_params *p) {
    if (!out || !cls || !p || nclusters < 1) return IFS_E_ARG;
    *out = nullptr;
    if (p->use_fmm) return IFS_E_ARG;  // independent clusters: direct plans only
    std::vector<double> centres, radii, mrel;
    std::vector<int32_t> lmax, group;
    bool any_l = false, all_l = true;
    for (int g = 0; g < nclusters; ++g) {
        const ifs_cluster &q = cls[g];
        if (q.n_spheres < 1 || !q.centres || !q.radii || !q.m_rel) return IFS_E_ARG;
        any_l = any_l || q.lmax;
        all_l = all_l && q.lmax;
    }
    if (any_l && !all_l) return IFS_E_ARG;  // per-sphere orders for every cluster or for none
    for (int g = 0; g < nclusters; ++g) {
        const ifs_cluster &q = cls[g];
        centres.insert(centres.end(), q.centres, q.centres + 3 * q.n_spheres);
        radii.insert(radii.end(), q.radii, q.radii + q.n_spheres);
        mrel.insert(mrel.end(), q.m_rel, q.m_rel + 2 * q.n_spheres);
        if (all_l) lmax.insert(lmax.end(), q.lmax, q.lmax + q.n_spheres);
        group.insert(group.end(), (size_t)q.n_spheres, g);
    }
    if (group.size() > (size_t)INT32_MAX) return IFS_E_ARG;
    ifs_cluster all{(int32_t)group.size(), centres.data(), radii.data(), mrel.data(),
                    all_l ? lmax.data() : nullptr};
    return create_impl(out, &all, p, group.data(), nclusters);
}

int32_t ifs_cluster_count(const ifs_ctx *c) { return c ? (int32_t)c->grp_first.size() - 1 : -1; }

int ifs_cluster_offsets(const ifs_ctx *c, int64_t *first_sphere, int64_t *first_entry) {
    if (!c) return IFS_E_ARG;
    const int64_t blk = 2LL * c->N * (c->N + 2);
    for (size_t g = 0; g < c->grp_first.size(); ++g) {
        if (first_sphere) first_sphere[g] = c->grp_first[g];
        if (first_entry) first_entry[g] = c->grp_first[g] * blk;
    }
    return IFS_OK;
}

static int create_impl(ifs_ctx **out, const ifs_cluster *cl, const ifs_params *p, const int32_t *group,
                       int32_t nclusters) {
    if (!out || !cl || !p) return IFS_E_ARG;
    *out = nullptr;
    if (cl->n_spheres < 1 || !cl->centres || !cl->radii || !cl->m_rel) return IFS_E_ARG;
    if (!(p->k > 0.0) || p->n_global < 1) return IFS_E_ARG;
    if (p->z_selector != IFS_ZSEL_BESSEL_J && p->z_selector != IFS_ZSEL_HANKEL1) return IFS_E_ARG;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return IFS_E_NODEVICE;

    ifs_ctx *c = new (std::nothrow) ifs_ctx();
    if (!c) return IFS_E_NOMEM;
    auto fail = [&](int rc) {
        ifs_destroy(c);
        return rc;
    };
    if (hipGetDevice(&c->device) != hipSuccess) return fail(IFS_E_HIP);
    if (hipStreamCreateWithFlags(&c->own_stream, hipStreamNonBlocking) != hipSuccess) return fail(IFS_E_HIP);
    c->stream = c->own_stream;
    c->prm = *p;
    c->ns = cl->n_spheres;
    c->N = p->n_global;
    const int ns = c->ns, N = c->N;
    if (group) {
        c->group.assign(group, group + ns);
        c->grp_first.assign(1, 0);
        for (int j = 1; j < ns; ++j)
            if (group[j] != group[j - 1]) c->grp_first.push_back(j);
        c->grp_first.push_back(ns);
        if ((int)c->grp_first.size() != nclusters + 1) return fail(IFS_E_ARG);
    } else {
        c->grp_first = {0, ns};
    }
    c->pos.assign(cl->centres, cl->centres + 3 * ns);
    c->radius.assign(cl->radii, cl->radii + ns);
    c->mrel.resize(ns);
    c->Lj.resize(ns);
    c->Leff.resize(ns);
    c->active.resize(ns);
    c->off.resize(ns);
    for (int j = 0; j < ns; ++j) {
        if (!(c->radius[j] > 0.0)) return fail(IFS_E_ARG);
        c->mrel[j] = {cl->m_rel[2 * j], cl->m_rel[2 * j + 1]};
        int L;
        if (cl->lmax) {
            L = cl->lmax[j];
        } else {
            double x = p->k * c->radius[j];
            L = (int)std::ceil(x + 4.05 * std::cbrt(x) + 2.0);
        }
        if (L < 1) return fail(IFS_E_ARG);
        c->Lj[j] = L;
        c->Leff[j] = std::min(L, N);
        c->active[j] = p->ref_row_zeroing ? (L >= N ? 1 : 0) : 1;
    }
    int64_t acc = 0;
    for (int j = 0; j < ns; ++j) {
        c->off[j] = acc;
        acc += 2LL * c->Leff[j] * (c->Leff[j] + 2);
    }
    c->ncompact = acc;
    c->nflat = 2LL * ns * N * (N + 2);

    int rc;
    if ((rc = c->d_Leff.alloc(4 * ns)) || (rc = c->d_active.alloc(4 * ns)) || (rc = c->d_off.alloc(8 * ns)) ||
        (rc = c->d_snm.alloc(8ull * N * (N + 2))) || (rc = c->d_an.alloc(16ull * ns * N)) ||
        (rc = c->d_bn.alloc(16ull * ns * N)) || (rc = c->d_xh.alloc(16ull * c->ncompact)) ||
        (rc = c->d_acc.alloc(16ull * c->ncompact)) || (rc = c->d_pos.alloc(24ull * ns)))
        return fail(rc);
    if ((rc = far_field_scratch(c->device, &c->ff_scratch)) != IFS_OK) return fail(rc);
    std::vector<double> snm = make_snm(N);
    if (hipMemcpy(c->d_Leff.p, c->Leff.data(), 4 * ns, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(c->d_active.p, c->active.data(), 4 * ns, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(c->d_off.p, c->off.data(), 8 * ns, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(c->d_snm.p, snm.data(), 8ull * N * (N + 2), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(c->d_pos.p, c->pos.data(), 24ull * ns, hipMemcpyHostToDevice) != hipSuccess)
        return fail(IFS_E_HIP);

    // Mie T-matrix on the device (FP64)
    {
        DevBuf dx, dm;
        std::vector<double> xs(ns);
        std::vector<double2> ms(ns);
        for (int j = 0; j < ns; ++j) {
            xs[j] = p->k * c->radius[j];
            ms[j] = make_double2(c->mrel[j].real(), c->mrel[j].imag());
        }
        if ((rc = dx.alloc(8 * ns)) || (rc = dm.alloc(16 * ns))) return fail(rc);
        if (hipMemcpy(dx.p, xs.data(), 8 * ns, hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(dm.p, ms.data(), 16 * ns, hipMemcpyHostToDevice) != hipSuccess)
            return fail(IFS_E_HIP);
        launch_mie(ns, dx.as<double>(), dm.as<double2>(), N, c->d_an.as<double2>(), c->d_bn.as<double2>(),
                   c->stream);
        if (hipStreamSynchronize(c->stream) != hipSuccess) return fail(IFS_E_HIP);
    }

    // workspace budget: a fraction of free memory
    size_t freeb = 0, totb = 0;
    (void)hipMemGetInfo(&freeb, &totb);
    c->ws_budget = std::min<size_t>(freeb / 4, (size_t)48 << 30);
    if (p->workspace_mb > 0) c->ws_budget = (size_t)p->workspace_mb << 20;
    c->cache_budget = freeb / 20 * 13;  // 65%: the translation workspace takes at most 48 GiB of the rest
    if (p->cache_mb > 0) c->cache_budget = (size_t)p->cache_mb << 20;
    if (p->cache_mb < 0) c->cache_budget = 0;
    if (p->d_convention != IFS_D_PHYSICAL && p->d_convention != IFS_D_REFERENCE) return fail(IFS_E_ARG);
    if (p->precision != IFS_PREC_FP64 && p->precision != IFS_PREC_MIXED) return fail(IFS_E_ARG);
    c->dsign = p->d_convention == IFS_D_REFERENCE ? 1.0 : -1.0;

    if (!p->use_fmm) {
        // direct all-pairs plan: for each target j, sources l != j in ascending order
        Batch &B = c->direct;
        B.name = "direct";
        size_t npairs = 0;
        for (size_t g = 0; g + 1 < c->grp_first.size(); ++g) {
            const size_t q = (size_t)(c->grp_first[g + 1] - c->grp_first[g]);
            npairs += q * (q - 1);
        }
        B.tasks.reserve(npairs);
        for (int j = 0; j < ns; ++j) {
            if (!c->active[j]) continue;
            // sources: the spheres of the target's own cluster
            const int lo = group ? c->grp_first[group[j]] : 0, hi = group ? c->grp_first[group[j] + 1] : ns;
            for (int l = lo; l < hi; ++l) {
                if (l == j) continue;
                double dv[3] = {c->dsign * (c->pos[3 * l] - c->pos[3 * j]),
                                c->dsign * (c->pos[3 * l + 1] - c->pos[3 * j + 1]),
                                c->dsign * (c->pos[3 * l + 2] - c->pos[3 * j + 2])};
                B.tasks.push_back(
                    make_task(dv, p->k, p->z_selector, c->Leff[l], c->Leff[j], c->off[l], c->off[j]));
            }
        }
        if ((rc = batch_finalize(c, B, std::min<size_t>(c->ws_budget, (size_t)8 << 30), true)) != IFS_OK)
            return fail(rc);
        batch_cost(B, c->flops_mv, c->bytes_mv);
    } else {
        if ((rc = fmm::build(c, &c->fmm)) != IFS_OK) return fail(rc);
    }
    for (int j = 0; j < ns; ++j) c->flops_mv += 12.0 * c->Leff[j] * (c->Leff[j] + 2);
    c->flops_mv += 4.0 * (double)c->nflat;
    c->bytes_mv += 2.0 * 16.0 * (double)c->nflat;
    *out = c;
    return IFS_OK;
}

int ifs_destroy(ifs_ctx *c) {
    if (!c) return IFS_OK;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    graphs_clear(c);
    if (c->fmm) fmm::destroy(c->fmm);
    if (c->own_stream) (void)hipStreamDestroy(c->own_stream);
    delete c;
    return IFS_OK;
}

int64_t ifs_vector_length(const ifs_ctx *c) { return c ? c->nflat : -1; }

int ifs_set_stream(ifs_ctx *c, void *s) {
    if (!c) return IFS_E_ARG;
    c->stream = s ? (hipStream_t)s : c->own_stream;
    return IFS_OK;
}

void *ifs_get_stream(ifs_ctx *c) { return c ? (void *)c->stream : nullptr; }

int ifs_set_precision(ifs_ctx *c, int32_t precision) {
    if (!c || (precision != IFS_PREC_FP64 && precision != IFS_PREC_MIXED)) return IFS_E_ARG;
    c->prm.precision = precision;
    return IFS_OK;
}

}  // extern "C"

namespace ifs {
// b = M x in the reference basis, or b_hat = S M S^-1 x_hat in the normalised
// basis (normalised = 1), both in the flat layout.
int matvec_impl(ifs_ctx *c, const double2 *x, double2 *b, int normalised, int nrhs, int64_t ld, int64_t ldb,
                int prec) {
    if (nrhs < 1) return IFS_E_ARG;
    if (ldb < 0) ldb = ld;
    if (nrhs > 1 && (ld < c->nflat || ldb < c->nflat)) return IFS_E_ARG;
    const double *snm = normalised ? nullptr : c->d_snm.as<double>();
    const int64_t nc = c->ncompact;
    int rc;
    if ((rc = c->d_xh.alloc(16ull * nc * nrhs)) || (rc = c->d_acc.alloc(16ull * nc * nrhs))) return rc;
    {
        Scope s(c, "pack_in");
        launch_pack_in(c->ns, c->N, c->d_Leff.as<int32_t>(), c->d_off.as<int64_t>(), snm, x, c->d_xh.as<double2>(),
                       nrhs, ld, nc, c->stream);
    }
    {
        Scope s(c, "zero");
        launch_zero(c->d_acc.as<double2>(), nc * nrhs, c->stream);
    }
    if (c->fmm) {
        rc = fmm::run(c, c->fmm, c->d_xh.as<double2>(), c->d_acc.as<double2>(), nrhs, nc, prec);
    } else {
        rc = batch_run(c, c->direct, c->d_xh.as<double2>(), c->d_acc.as<double2>(), nrhs, nc, nc);
    }
    if (rc != IFS_OK) return rc;
    {
        Scope s(c, "finalize");
        launch_finalize(c->ns, c->N, c->d_Leff.as<int32_t>(), c->d_active.as<int32_t>(), c->d_off.as<int64_t>(),
                        snm, c->d_an.as<double2>(), c->d_bn.as<double2>(), x, c->d_acc.as<double2>(), b, nrhs, ld,
                        ldb, nc, c->stream);
    }
    if (hipPeekAtLastError() != hipSuccess) return IFS_E_HIP;
    return IFS_OK;
}

static uint64_t matvec_sig(ifs_ctx *c) {
    uint64_t h = 0;
    auto mix = [&](uint64_t v) { h ^= v + 0x9e3779b97f4a7c15ull + (h << 6) + (h >> 2); };
    for (const DevBuf *d : {&c->d_xh, &c->d_acc, &c->ws_d, &c->ws_s, &c->ws_ax, &c->ws_y, &c->ws_z, &c->ws_gq,
                            &c->direct.d_tasks, &c->direct.d_segs, &c->direct.own_d, &c->direct.own_s,
                            &c->direct.own_ax, &c->direct.own_y, &c->direct.own_z, &c->d_snm, &c->d_an, &c->d_bn})
        mix((uint64_t)(uintptr_t)d->p);
    for (auto &kv : c->gq_nodes) mix((uint64_t)(uintptr_t)kv.second.p);
    if (c->fmm) mix(fmm::buffers_sig(c->fmm));
    return h;
}

void graphs_clear(ifs_ctx *c) {
    for (auto &g : c->graphs)
        if (g.exec) (void)hipGraphExecDestroy(g.exec);
    c->graphs.clear();
}

int matvec_graph(ifs_ctx *c, const double2 *x, double2 *b, int normalised, int nrhs, int64_t ld, int64_t ldb,
                 int prec) {
    if (ldb < 0) ldb = ld;
    if (c->prof.on || !c->stream) return matvec_impl(c, x, b, normalised, nrhs, ld, ldb, prec);
    ifs_ctx::Graph *g = nullptr;
    for (auto &e : c->graphs)
        if (e.x == x && e.b == b && e.normalised == normalised && e.nrhs == nrhs && e.ld == ld && e.ldb == ldb &&
            e.prec == prec)
            g = &e;
    const uint64_t sig = matvec_sig(c);
    if (g && g->exec && g->sig == sig)
        return hipGraphLaunch(g->exec, c->stream) == hipSuccess ? IFS_OK : IFS_E_HIP;
    if (!g) {
        // first call with these arguments: eager (sizes every buffer, builds kept tables);
        // at most kMaxGraphs argument sets are kept, the oldest dropped first
        int rc = matvec_impl(c, x, b, normalised, nrhs, ld, ldb, prec);
        if (rc) return rc;
        constexpr size_t kMaxGraphs = 8;
        if (c->graphs.size() >= kMaxGraphs) {
            if (c->graphs.front().exec) (void)hipGraphExecDestroy(c->graphs.front().exec);
            c->graphs.erase(c->graphs.begin());
        }
        c->graphs.push_back({x, b, normalised, nrhs, prec, ld, ldb, 0, nullptr});
        return IFS_OK;
    }
    if (g->exec) {  // a buffer or plan changed since the capture
        (void)hipGraphExecDestroy(g->exec);
        g->exec = nullptr;
    }
    hipGraph_t graph = nullptr;
    if (hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal) != hipSuccess) {
        (void)hipGetLastError();
        return matvec_impl(c, x, b, normalised, nrhs, ld, ldb, prec);
    }
    int rc = matvec_impl(c, x, b, normalised, nrhs, ld, ldb, prec);
    const hipError_t e = hipStreamEndCapture(c->stream, &graph);
    hipGraphExec_t exec = nullptr;
    const bool ok = rc == IFS_OK && e == hipSuccess && graph && matvec_sig(c) == sig &&
                    hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0) == hipSuccess;
    if (graph) (void)hipGraphDestroy(graph);
    if (!ok) {
        // nothing ran: the launches went into the capture; run them eagerly
        (void)hipGetLastError();
        if (exec) (void)hipGraphExecDestroy(exec);
        return matvec_impl(c, x, b, normalised, nrhs, ld, ldb, prec);
    }
    g->exec = exec;
    g->sig = sig;
    return hipGraphLaunch(exec, c->stream) == hipSuccess ? IFS_OK : IFS_E_HIP;
}

// ||S^-1 r_hat|| / ||S^-1 b_hat|| (the relative residual in the flat basis) from the
// normalised residual and right-hand side on the device, summed on the host in long
// double (flat entries reach 1e+-80 at N = 50, their squares overflow doubles at N ~ 80).
double flat_rel_residual(ifs_ctx *c, const double2 *r_hat, const double2 *b_hat, int64_t n) {
    std::vector<double2> r((size_t)n), b((size_t)n);
    if (hipMemcpyAsync(r.data(), r_hat, 16ull * n, hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
        hipMemcpyAsync(b.data(), b_hat, 16ull * n, hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
        hipStreamSynchronize(c->stream) != hipSuccess)
        return -1.0;
    const std::vector<double> snm = make_snm(c->N);
    const int64_t blk = (int64_t)c->N * (c->N + 2);
    long double rr = 0.0L, bb = 0.0L;
    for (int64_t i = 0; i < n; ++i) {
        const long double is = 1.0L / (long double)snm[(size_t)(i % blk)];
        const long double a = r[i].x * is, d = r[i].y * is, e = b[i].x * is, f = b[i].y * is;
        rr += a * a + d * d;
        bb += e * e + f * f;
    }
    return bb > 0.0L ? (double)sqrtl(rr / bb) : 0.0;
}

int scale_flat(ifs_ctx *c, const double2 *x, double2 *y, int inverse) {
    launch_scale_flat(c->ns, c->N, c->d_snm.as<double>(), x, y, inverse, c->stream);
    return hipPeekAtLastError() == hipSuccess ? IFS_OK : IFS_E_HIP;
}
}  // namespace ifs

extern "C" {

int ifs_matvec_device(ifs_ctx *c, const void *x_dev, void *b_dev, int64_t n) {
    if (!c || !x_dev || !b_dev || n != c->nflat) return IFS_E_ARG;
    return matvec_impl(c, (const double2 *)x_dev, (double2 *)b_dev, 0, 1, n, -1, c->prm.precision);
}

int ifs_set_class_tiles(ifs_ctx *c, int32_t fwd_rows, int32_t back_rows, int32_t autotune) {
    auto ok = [](int32_t r) { return r == 0 || r == 32 || r == 64; };
    if (!c || !ok(fwd_rows) || !ok(back_rows) || (autotune != 0 && autotune != 1)) return IFS_E_ARG;
    c->force_rows[0] = fwd_rows;
    c->force_rows[1] = back_rows;
    c->autotune = autotune;
    graphs_clear(c);  // captured launches carry the old block heights
    return IFS_OK;
}

int ifs_class_tile_stats(ifs_ctx *c, int64_t counts[2]) {
    if (!c || !counts) return IFS_E_ARG;
    counts[0] = c->axial_tiles[0];
    counts[1] = c->axial_tiles[1];
    return IFS_OK;
}

int ifs_matvec_graph_device(ifs_ctx *c, const void *x_dev, void *b_dev, int64_t n) {
    if (!c || !x_dev || !b_dev || n != c->nflat) return IFS_E_ARG;
    return matvec_graph(c, (const double2 *)x_dev, (double2 *)b_dev, 0, 1, n, -1, c->prm.precision);
}

int ifs_matvec_batch_device(ifs_ctx *c, int32_t nrhs, const void *x_dev, void *b_dev, int64_t n, int64_t ld) {
    if (!c || !x_dev || !b_dev || n != c->nflat || nrhs < 1 || ld < n) return IFS_E_ARG;
    return matvec_impl(c, (const double2 *)x_dev, (double2 *)b_dev, 0, nrhs, ld, -1, c->prm.precision);
}

int ifs_matvec_batch(ifs_ctx *c, int32_t nrhs, const double *x, double *b, int64_t n) {
    if (!c || !x || !b || n != c->nflat || nrhs < 1) return IFS_E_ARG;
    int rc;
    if ((rc = c->d_x.alloc(16ull * n * nrhs)) || (rc = c->d_b.alloc(16ull * n * nrhs))) return rc;
    HIP_OK(hipMemcpyAsync(c->d_x.p, x, 16ull * n * nrhs, hipMemcpyHostToDevice, c->stream));
    if ((rc = matvec_impl(c, c->d_x.as<double2>(), c->d_b.as<double2>(), 0, nrhs, n, -1, c->prm.precision))) return rc;
    HIP_OK(hipMemcpyAsync(b, c->d_b.p, 16ull * n * nrhs, hipMemcpyDeviceToHost, c->stream));
    HIP_OK(hipStreamSynchronize(c->stream));
    return IFS_OK;
}

int ifs_matvec(ifs_ctx *c, const double *x, double *b, int64_t n) {
    if (!c || !x || !b || n != c->nflat) return IFS_E_ARG;
    int rc;
    if ((rc = c->d_x.alloc(16ull * n)) || (rc = c->d_b.alloc(16ull * n))) return rc;
    HIP_OK(hipMemcpyAsync(c->d_x.p, x, 16ull * n, hipMemcpyHostToDevice, c->stream));
    if ((rc = ifs_matvec_device(c, c->d_x.p, c->d_b.p, n)) != IFS_OK) return rc;
    HIP_OK(hipMemcpyAsync(b, c->d_b.p, 16ull * n, hipMemcpyDeviceToHost, c->stream));
    HIP_OK(hipStreamSynchronize(c->stream));
    return IFS_OK;
}

int ifs_get_mie(ifs_ctx *c, double *an, double *bn) {
    if (!c || !an || !bn) return IFS_E_ARG;
    HIP_OK(hipStreamSynchronize(c->stream));
    HIP_OK(hipMemcpy(an, c->d_an.p, 16ull * c->ns * c->N, hipMemcpyDeviceToHost));
    HIP_OK(hipMemcpy(bn, c->d_bn.p, 16ull * c->ns * c->N, hipMemcpyDeviceToHost));
    return IFS_OK;
}

int ifs_profile_enable(ifs_ctx *c, int32_t enable) {
    if (!c) return IFS_E_ARG;
    if (!enable) c->prof.flush();
    c->prof.on = enable != 0;
    if (enable) c->prof.totals.clear();
    return IFS_OK;
}

int ifs_profile_read(ifs_ctx *c, const char **names, double *total_ms, int64_t *launches, double *flops,
                     int32_t cap) {
    if (!c || !names || !total_ms || !launches || !flops) return IFS_E_ARG;
    c->prof.flush();
    int i = 0;
    for (auto &kv : c->prof.totals) {
        if (i >= cap) break;
        names[i] = kv.first;
        total_ms[i] = kv.second.ms;
        launches[i] = kv.second.launches;
        flops[i] = kv.second.flops;
        ++i;
    }
    return i;
}

int ifs_tree_export(ifs_ctx *c, int64_t *uindex_out, int32_t *levels_out) {
    if (!c || !uindex_out || !levels_out) return IFS_E_ARG;
    if (!c->fmm) return IFS_E_ARG;
    return fmm::tree_export(c->fmm, uindex_out, levels_out);
}

int ifs_fmm_stats(ifs_ctx *c, int64_t counts[6], int32_t levels_L[32]) {
    if (!c || !counts || !levels_L) return IFS_E_ARG;
    if (!c->fmm) return IFS_E_ARG;
    return fmm::stats(c->fmm, counts, levels_L);
}

int ifs_matvec_cost(ifs_ctx *c, double *flops, double *bytes) {
    if (!c || !flops || !bytes) return IFS_E_ARG;
    *flops = c->flops_mv;
    *bytes = c->bytes_mv;
    return IFS_OK;
}

int ifs_coaxial_block(int32_t kind, double t, int32_t Lt, int32_t Ls, int32_t K, int32_t method, int32_t k,
                      double *out) {
    if ((kind != 1 && kind != 3) || !(t > 0.0) || Lt < 1 || Ls < 1 || K < 0 || K > std::min(Lt, Ls) || k < -K ||
        k > K || (method != 0 && method != 1) || !out)
        return IFS_E_ARG;
    struct Stream {
        hipStream_t s = nullptr;
        ~Stream() {
            if (s) (void)hipStreamDestroy(s);
        }
    } st;
    if (hipStreamCreate(&st.s) != hipSuccess) return IFS_E_HIP;
    DevBuf wax, ws, rule, dtask;
    int rc;
    if ((rc = wax.alloc(16 * (size_t)ax_size(Lt, Ls, K)))) return rc;
    if (method == 0) {
        // the recurrence generator of the sphere-level tables: sectorial columns, then
        // the three-term recurrence in the degrees
        Task T{};
        T.t = t;
        T.Lt = Lt;
        T.Ls = Ls;
        T.K = K;
        T.kind = kind;
        const int nmax = Lt + Ls + 2;
        if ((rc = dtask.alloc(sizeof(Task))) || (rc = ws.alloc(16 * (size_t)sect_size(K, nmax)))) return rc;
        if (hipMemcpy(dtask.p, &T, sizeof(Task), hipMemcpyHostToDevice) != hipSuccess) return IFS_E_HIP;
        launch_gen_sect(dtask.as<Task>(), 1, nmax, ws.as<double2>(), st.s);
        launch_gen_axial(dtask.as<Task>(), 1, K, nmax, ws.as<double2>(), wax.as<double2>(), st.s);
    } else {
        // the Gauss-Legendre generator of the box tables
        GqPlan G = gq_plan(kind, t, Lt, Ls, K, (size_t)1 << 30);
        std::vector<double> x, w;
        gauss_legendre(G.Q, x, w);
        if ((rc = rule.alloc(16 * (size_t)G.Q)) || (rc = ws.alloc(G.ws_bytes))) return rc;
        if (hipMemcpy(rule.p, x.data(), 8 * (size_t)G.Q, hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(rule.as<double>() + G.Q, w.data(), 8 * (size_t)G.Q, hipMemcpyHostToDevice) != hipSuccess)
            return IFS_E_HIP;
        launch_gen_axial_gq(G, rule.as<double>(), rule.as<double>() + G.Q, ws.p, wax.as<double2>(), st.s);
    }
    if (hipPeekAtLastError() != hipSuccess) return IFS_E_HIP;
    const size_t nb = 16 * 2 * (size_t)ax_block(k, Lt, Ls);  // the P block, then the Q block
    if (hipMemcpyAsync(out, wax.as<double2>() + ax_off(k, Lt, Ls, K), nb, hipMemcpyDeviceToHost, st.s) !=
            hipSuccess ||
        hipStreamSynchronize(st.s) != hipSuccess)
        return IFS_E_HIP;
    return IFS_OK;
}

}  // extern "C"
