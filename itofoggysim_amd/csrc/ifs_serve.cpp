// Serving loop: continuous batching of scattering requests on one GPU.
//
// Requests (a cluster, its parameters and 1..n incidence directions) are queued
// as they arrive.  Worker slots take the oldest queued request whose estimated HBM
// footprint fits what the admitted requests leave of the budget, so requests that
// arrive while others run are admitted as soon as a slot and memory free up.
// Every admitted request owns a context with its own HIP stream: its tables and
// Krylov state stay device-resident for the whole request, and the work of
// concurrent requests overlaps on the device.  A request's result does not depend
// on what runs beside it (no shared device state, deterministic kernels).
//
// Per request: right-hand sides T p for every direction, lock-step GMRES over the
// directions (one batched matvec per Arnoldi step), then the cross sections.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <deque>
#include <map>
#include <cstdio>
#include <mutex>
#include <thread>
#include <vector>

#include "../../include/ifs.h"

namespace {

using clk = std::chrono::steady_clock;

struct Req {
    int64_t id = 0;
    std::vector<double> centres, radii, mrel, alpha, beta;
    std::vector<int32_t> lmax;
    ifs_params prm{};
    double tol = 1e-6;
    int restart = 30, max_iter = 1000;
    size_t footprint = 0;
    double est_s = 0.0;  // estimated device time (admission order)
    clk::time_point t_submit, t_start, t_done;
    // result
    int status = 0;
    bool done = false;
    std::vector<double> csec;  // 3 per direction
    int iterations = 0;
};

}  // namespace

struct ifs_server {
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<int64_t> queue;
    std::map<int64_t, Req> reqs;
    std::vector<std::thread> workers;
    size_t budget = 0, admitted = 0, peak_admitted = 0;
    int slots = 1;
    int running = 0, peak_running = 0;
    int64_t next_id = 1;
    bool stop = false;
    int verbose = 0;  // progress lines on stderr: each request's GMRES steps and its completion
    clk::time_point t0;
    std::vector<double> latencies;
    int64_t completed = 0;
};

namespace {

// Device memory of a request, planned on the host before admission, and the limits its
// context is created with so that it stays inside the plan (ifs_create would otherwise
// size its workspace and table cache from the free memory of the whole device):
//   vectors    right-hand sides, iterates, the lock-step GMRES basis, compact copies;
//   pools      MLFMM multipole and local expansions of every box (octree built on the
//              host with the library's own functions: scaling, Morton keys, l_max);
//   workspace  the translation workspace (ifs_params.workspace_mb), at most a quarter
//              of a slot's share of the budget and at most 8 GiB;
//   tables     translation tables kept across matvecs (ifs_params.cache_mb): the
//              direct path's tables or the MLFMM's M2L coaxial tables, when they fit
//              the rest of the slot's share; otherwise regenerated in every matvec;
//   fixed      per-context buffers and the quadrature workspaces.
struct Plan {
    size_t vectors = 0, pools = 0, workspace = 0, tables = 0, fixed = 0;
    double flops = 0.0;  // translation flops of one matvec per right-hand side (about 21 L^3 each)
    size_t total() const { return vectors + pools + workspace + tables + fixed; }
};

int box_L(double k, double edge, double d0, int cap) {
    const double kD = k * edge;
    int L = (int)std::ceil(kD + 1.8 * std::pow(d0, 2.0 / 3.0) * std::cbrt(kD));
    if (L < 1) L = 1;
    if (cap > 0 && L > cap) L = cap;
    return L;
}

// sum_{q=1..K} 2 (L+1-q)^2 + (L)^2 complex per coaxial table of orders (L, L), both P and Q
double coax_table_bytes(int L) {
    double e = (double)L * L;
    for (int q = 1; q <= L; ++q) e += 2.0 * (L + 1.0 - q) * (L + 1.0 - q);
    return 2.0 * 16.0 * e;
}

Plan plan_request(Req &r, size_t budget, int slots) {
    Plan P;
    const int ns = (int)r.radii.size();
    int N = r.prm.n_global;
    const double k = r.prm.k;
    std::vector<int> L(ns);
    for (int j = 0; j < ns; ++j) {
        int lj = r.lmax.empty() ? (int)std::ceil(k * r.radii[j] + 4.05 * std::cbrt(k * r.radii[j]) + 2.0) : r.lmax[j];
        L[j] = N > 0 ? std::min(lj, N) : lj;
    }
    if (N <= 0) N = *std::max_element(L.begin(), L.end());
    const double n = 2.0 * ns * N * (N + 2);
    const int nd = (int)r.alpha.size();
    P.vectors = (size_t)(16.0 * n * ((r.restart + 1) * nd + 6.0 * nd + 4.0));
    const size_t share = budget / (size_t)std::max(slots, 1);
    const size_t ws_cap = std::min<size_t>(std::max<size_t>(share / 4, (size_t)256 << 20), (size_t)8 << 30);
    double table_need = 0.0, ws_need = 0.0;
    if (r.prm.use_fmm) {
        std::vector<double> scaled(3 * ns), xmd(4);
        std::vector<int64_t> u(ns);
        ifs_tree_scale(r.centres.data(), ns, scaled.data(), xmd.data());
        ifs_tree_universal_index(scaled.data(), ns, 21, u.data());
        std::sort(u.begin(), u.end());
        const int s = r.prm.leaf_size > 0 ? r.prm.leaf_size : 10;
        const int lmax = std::max(2, ifs_tree_level_max(u.data(), ns, s));
        const double d0 = r.prm.fmm_d0 > 0.0 ? r.prm.fmm_d0 : 6.0;
        double pools = 0.0;
        for (int l = 2; l <= lmax; ++l) {
            std::vector<int64_t> b(u.size());
            for (int i = 0; i < ns; ++i) b[i] = u[i] >> (3 * (21 - l));
            b.erase(std::unique(b.begin(), b.end()), b.end());
            const int Lb = box_L(k, xmd[3] / std::ldexp(1.0, l), d0, r.prm.fmm_lmax_cap);
            pools += 2.0 * (double)b.size() * 2.0 * Lb * (Lb + 2.0) * 16.0 * nd;
            P.flops += (double)b.size() * 100.0 * 21.0 * Lb * Lb * (double)Lb;  // ~100 M2L per box
            // M2L: at most 15 distinct offset lengths per level, one coaxial table each
            table_need += 15.0 * coax_table_bytes(Lb);
            // class columns: up to 189 M2L per box, Y/W and Z of 4 L(L+2) doubles each
            ws_need += (double)b.size() * 189.0 * 2.0 * 4.0 * Lb * (Lb + 2.0) * 8.0 * nd;
            // sphere <-> box quadrature tables of one chunk
            ws_need = std::max(ws_need, (double)ns * 2.0 * 16.0 * Lb * 51.0 * 51.0);
        }
        P.pools = (size_t)pools;
        double l3 = 0.0;
        for (int j = 0; j < ns; ++j) l3 += (double)L[j] * L[j] * L[j];
        P.flops += 27.0 * std::min(ns, 2 * s) * 21.0 * l3;  // near field: about 27 leaves of up to 2 s spheres
        P.fixed = ((size_t)1 << 30) + ((size_t)256 << 20);  // quadrature workspaces and tables of the box levels
    } else {
        for (int a = 0; a < ns; ++a)
            for (int b = 0; b < ns; ++b) {
                if (a == b) continue;
                double Ls = L[a], Lt = L[b], K = std::min(Ls, Lt), Lm = std::max(Ls, Lt);
                table_need += 8.0 * 2.0 * (K + 1) * (Lm + 1) * (Lm + 2) / 2.0 + 16.0 * 2.0 * (2 * K + 1) * Ls * Lt +
                              16.0 * 2.0 * (2 * K + 1) * (Ls + Lt);
                P.flops += 21.0 * Lm * Lm * Lm;
            }
        ws_need = table_need * nd;
        P.fixed = (size_t)64 << 20;
    }
    P.workspace = (size_t)std::min<double>((double)ws_cap, ws_need + (64 << 20));
    // tables kept only when they fit what the slot's share leaves
    const size_t used = P.vectors + P.pools + P.workspace + P.fixed;
    if (table_need > 0.0 && used + (size_t)table_need <= share) P.tables = (size_t)table_need;
    r.prm.workspace_mb = (int32_t)std::max<size_t>(1, P.workspace >> 20);
    r.prm.cache_mb = P.tables > 0 ? (int32_t)((P.tables >> 20) + 1) : -1;
    if (P.tables > 0) P.tables = (size_t)r.prm.cache_mb << 20;
    return P;
}

void run_request(Req &r, int verbose) {
    ifs_cluster cl{(int32_t)r.radii.size(), r.centres.data(), r.radii.data(), r.mrel.data(),
                   r.lmax.empty() ? nullptr : r.lmax.data()};
    ifs_ctx *c = nullptr;
    int rc = ifs_create(&c, &cl, &r.prm);
    if (rc) {
        r.status = rc;
        return;
    }
    const int nd = (int)r.alpha.size();
    const int64_t n = ifs_vector_length(c);
    std::vector<double> rhs(2 * n * nd), p(2 * n * nd), x(2 * n * nd);
    for (int d = 0; d < nd && !rc; ++d) {
        rc = ifs_plane_wave_rhs_dir(c, r.alpha[d], r.beta[d], rhs.data() + 2 * n * d, n);
        if (!rc) rc = ifs_plane_wave(c, r.alpha[d], r.beta[d], p.data() + 2 * n * d, n);
    }
    std::vector<ifs_solve_info> info(nd);
    if (!rc) {
        ifs_solve_opts op{r.restart, r.max_iter, r.tol, 0, verbose, r.prm.precision, 0};
        rc = ifs_solve_batch(c, nd, rhs.data(), x.data(), n, &op, info.data());
    }
    r.csec.assign(3 * nd, 0.0);
    for (int d = 0; d < nd && (rc == IFS_OK || rc == IFS_E_NOCONV); ++d) {
        int q = ifs_cross_sections(c, x.data() + 2 * n * d, p.data() + 2 * n * d, n, r.csec.data() + 3 * d);
        if (q) rc = q;
        r.iterations = std::max(r.iterations, info[d].iterations);
    }
    r.status = rc;
    ifs_destroy(c);
}

void worker(ifs_server *s) {
    std::unique_lock<std::mutex> lk(s->mu);
    for (;;) {
        // among the queued requests whose footprint fits (any one if nothing runs), the one
        // with the least estimated device time less its waiting time: short requests pass
        // long ones, and a long one's priority grows as it waits (no starvation)
        int64_t pick = 0;
        s->cv_work.wait(lk, [&] {
            if (s->stop && s->queue.empty()) return true;
            const clk::time_point now = clk::now();
            double best = 0.0;
            pick = 0;
            for (int64_t id : s->queue) {
                Req &r = s->reqs[id];
                if (!(s->admitted + r.footprint <= s->budget || s->running == 0)) continue;
                const double prio = r.est_s - std::chrono::duration<double>(now - r.t_submit).count();
                if (!pick || prio < best) {
                    pick = id;
                    best = prio;
                }
            }
            return pick != 0;
        });
        if (!pick) {
            if (s->stop && s->queue.empty()) return;
            continue;
        }
        s->queue.erase(std::find(s->queue.begin(), s->queue.end(), pick));
        Req &r = s->reqs[pick];
        s->admitted += r.footprint;
        s->peak_admitted = std::max(s->peak_admitted, s->admitted);
        s->running += 1;
        s->peak_running = std::max(s->peak_running, s->running);
        r.t_start = clk::now();
        const int verbose = s->verbose;
        lk.unlock();
        run_request(r, verbose);  // r stays valid: map nodes are stable and erased only by the waiter
        lk.lock();
        r.t_done = clk::now();
        if (verbose) {
            std::fprintf(stderr, "serve: request %lld done: %zu spheres, %zu directions, %s, %.1f s (queued %.1f s), "
                                 "status %d\n",
                         (long long)r.id, r.radii.size(), r.alpha.size(),
                         r.prm.precision == IFS_PREC_MIXED ? "mixed" : "fp64",
                         std::chrono::duration<double>(r.t_done - r.t_submit).count(),
                         std::chrono::duration<double>(r.t_start - r.t_submit).count(), r.status);
            std::fflush(stderr);
        }
        r.done = true;
        s->admitted -= r.footprint;
        s->running -= 1;
        s->latencies.push_back(std::chrono::duration<double>(r.t_done - r.t_submit).count());
        s->completed += 1;
        s->cv_done.notify_all();
        s->cv_work.notify_all();
    }
}

}  // namespace

extern "C" {

int ifs_server_set_verbose(ifs_server *s, int32_t verbose) {
    if (!s || (verbose != 0 && verbose != 1)) return IFS_E_ARG;
    std::lock_guard<std::mutex> lk(s->mu);
    s->verbose = verbose;
    return IFS_OK;
}

int ifs_server_create(ifs_server **out, int32_t slots, double hbm_budget_gb) {
    if (!out || slots < 1 || slots > 64 || !(hbm_budget_gb > 0.0)) return IFS_E_ARG;
    ifs_server *s = new ifs_server();
    s->budget = (size_t)(hbm_budget_gb * 1073741824.0);
    s->slots = slots;
    s->t0 = clk::now();
    for (int i = 0; i < slots; ++i) s->workers.emplace_back(worker, s);
    *out = s;
    return IFS_OK;
}

int ifs_server_submit(ifs_server *s, const ifs_cluster *cl, const ifs_params *prm, int32_t ndir, const double *alpha,
                      const double *beta, double tol, int32_t restart, int64_t *id_out) {
    if (!s || !cl || !prm || ndir < 1 || !alpha || !beta || !id_out || cl->n_spheres < 1 || !cl->centres ||
        !cl->radii || !cl->m_rel || !(tol > 0.0) || restart < 1)
        return IFS_E_ARG;
    Req r;
    const int ns = cl->n_spheres;
    r.centres.assign(cl->centres, cl->centres + 3 * ns);
    r.radii.assign(cl->radii, cl->radii + ns);
    r.mrel.assign(cl->m_rel, cl->m_rel + 2 * ns);
    if (cl->lmax) r.lmax.assign(cl->lmax, cl->lmax + ns);
    r.alpha.assign(alpha, alpha + ndir);
    r.beta.assign(beta, beta + ndir);
    r.prm = *prm;
    r.tol = tol;
    r.restart = restart;
    if (r.prm.precision != IFS_PREC_FP64 && r.prm.precision != IFS_PREC_MIXED) return IFS_E_ARG;
    // the context's workspace and table cache come from the plan, not from free memory
    {
        const Plan P = plan_request(r, s->budget, s->slots);
        r.footprint = P.total();
        // ~20 matvecs per solve at ~10 TFLOP/s effective, per right-hand side
        r.est_s = 20.0 * P.flops * (double)r.alpha.size() / 1e13;
    }
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->stop) return IFS_E_ARG;
    r.id = s->next_id++;
    r.t_submit = clk::now();
    *id_out = r.id;
    s->reqs.emplace(r.id, std::move(r));
    s->queue.push_back(*id_out);
    s->cv_work.notify_all();
    return IFS_OK;
}

int ifs_server_plan(const ifs_cluster *cl, const ifs_params *prm, int32_t ndir, int32_t restart,
                    double hbm_budget_gb, int32_t slots, double out[8]) {
    if (!cl || !prm || !out || ndir < 1 || cl->n_spheres < 1 || !cl->centres || !cl->radii || restart < 1 ||
        slots < 1 || !(hbm_budget_gb > 0.0))
        return IFS_E_ARG;
    Req r;
    const int ns = cl->n_spheres;
    r.centres.assign(cl->centres, cl->centres + 3 * ns);
    r.radii.assign(cl->radii, cl->radii + ns);
    if (cl->lmax) r.lmax.assign(cl->lmax, cl->lmax + ns);
    r.alpha.assign((size_t)ndir, 0.0);
    r.beta.assign((size_t)ndir, 0.0);
    r.prm = *prm;
    r.restart = restart;
    const Plan P = plan_request(r, (size_t)(hbm_budget_gb * 1073741824.0), slots);
    out[0] = (double)P.vectors;
    out[1] = (double)P.pools;
    out[2] = (double)P.workspace;
    out[3] = (double)P.tables;
    out[4] = (double)P.fixed;
    out[5] = (double)P.total();
    out[6] = (double)r.prm.workspace_mb;
    out[7] = (double)r.prm.cache_mb;
    return IFS_OK;
}

int ifs_server_wait(ifs_server *s, int64_t id, double *csec, ifs_request_info *info) {
    if (!s) return IFS_E_ARG;
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->reqs.find(id);
    if (it == s->reqs.end()) return IFS_E_ARG;
    s->cv_done.wait(lk, [&] { return it->second.done; });
    Req &r = it->second;
    if (csec) std::copy(r.csec.begin(), r.csec.end(), csec);
    if (info) {
        info->status = r.status;
        info->iterations = r.iterations;
        info->latency_s = std::chrono::duration<double>(r.t_done - r.t_submit).count();
        info->queue_s = std::chrono::duration<double>(r.t_start - r.t_submit).count();
        info->footprint_bytes = (int64_t)r.footprint;
    }
    int st = r.status;
    s->reqs.erase(it);
    return st;
}

int ifs_server_stats(ifs_server *s, double out[6]) {
    if (!s || !out) return IFS_E_ARG;
    std::lock_guard<std::mutex> lk(s->mu);
    std::vector<double> l = s->latencies;
    std::sort(l.begin(), l.end());
    auto pct = [&](double q) {
        if (l.empty()) return 0.0;
        size_t i = (size_t)std::ceil(q * l.size()) - 1;
        return l[std::min(i, l.size() - 1)];
    };
    double el = std::chrono::duration<double>(clk::now() - s->t0).count();
    out[0] = (double)s->completed;
    out[1] = pct(0.50);
    out[2] = pct(0.99);
    out[3] = el > 0.0 ? s->completed / el : 0.0;
    out[4] = (double)s->peak_running;
    out[5] = (double)s->peak_admitted;
    return IFS_OK;
}

int ifs_server_destroy(ifs_server *s) {
    if (!s) return IFS_OK;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
        s->cv_work.notify_all();
    }
    for (auto &t : s->workers) t.join();
    delete s;
    return IFS_OK;
}

}  // extern "C"
