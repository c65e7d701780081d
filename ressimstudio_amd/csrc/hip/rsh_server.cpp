// Serving loop for ensemble solves (include/ressim_hip.h, rsh_server_*).
//
// The reference's ensemble drivers (runBatch / runESMDA, study.cpp:326-364, 1254) run many Simulators in threads,
// each calling the global solver hook once per Newton iteration.  The server is that hook's device side for many
// callers at once: models (one sparsity pattern each) stay resident with their slice layouts, ILU schedule and AMG
// hierarchy, so a request is only new values + numeric setup + solve; requests queue in arrival order and are
// admitted as soon as a worker slot is free, the request's model is idle and the model fits the HBM budget (idle
// models are evicted least-recently-used to make room); each worker owns a HIP stream, so admitted requests run
// concurrently on the device.  Every kernel is deterministic, so a request's result does not depend on what runs
// beside it.
#include <atomic>
#include <condition_variable>
#include <deque>
#include <list>
#include <thread>
#include <unordered_map>

#include "rsh_solver.hpp"

using namespace rsh;


struct rsh_server {
    struct Model {
        int n = 0, nb = 1, nnzb = 0;
        std::vector<int> rp, ci;
        std::vector<double> baseHost;        // kept on the host so an evicted model can come back
        rsh_options opt{};
        rsh_system* S = nullptr;             // resident system (null when evicted)
        DBuf<double> base, x, b;             // resident base values, right-hand side and solution
        size_t footprint = 0;                // HBM bytes while resident (estimated until first residency)
        bool busy = false;
        double lastUse = 0;
        int loads = 0;
    };
    struct Req {
        long long id = 0;
        rsh_request r{};
        double tSubmit = 0, tStart = 0, tEnd = 0;
        rsh_result res{};
        std::string msg;                     // the worker's error message, handed to the waiting thread
        bool done = false;
    };
    // batched rounds (rsh_server_create_batched): the models are grouped once, at the first request, into rsh_batch
    // groups (same block size and options, ascending size, at most groupCells cells or one model, at most
    // groupSystems models); a round serves every queued request of one idle group (the earliest per model) with one
    // numeric set-up and one batched BiCGSTAB of the group, the group's other members solving a zero right-hand side
    struct Group {
        std::vector<int> models;             // model ids, batch member order
        rsh_context* ctx = nullptr;
        rsh_batch* B = nullptr;
        DBuf<double> base, b, x, tol;        // base values in the batch's group layout; rhs / solution in input order
        std::vector<double> tolH;
        size_t footprint = 0;
        bool busy = false;
        long long rounds = 0;
    };
    bool batched = false, grouped = false;
    long long groupCells = 0;
    int groupSystems = 0;
    std::vector<std::unique_ptr<Group>> groups;
    std::vector<int> groupOf;                // model -> group
    long long rounds = 0;
    int device = 0;
    size_t budget = 0, used = 0;
    std::vector<std::unique_ptr<Model>> models;
    std::vector<rsh_context*> ctxs;
    std::vector<std::thread> workers;
    std::mutex mu;
    std::condition_variable cvWork, cvDone;
    std::deque<long long> queue;
    std::unordered_map<long long, std::unique_ptr<Req>> reqs;
    long long nextId = 1;
    bool stop = false;
    int running = 0, peakRunning = 0;
    long long completed = 0, evictions = 0;

    static size_t estimate(const Model& m) {
        const size_t vals = size_t(m.nnzb) * m.nb * m.nb * 8;
        const size_t vec = size_t(m.n) * m.nb * 8;
        // natural values, SpMV slices, ILU factors + sweep slices, pressure column, base values; ~20 vectors
        return vals * 6 + vec * 20 + size_t(m.nnzb) * 16;
    }
    size_t measure(const Model& m) const { return m.S->deviceBytes() + m.base.bytes() + m.x.bytes() + m.b.bytes(); }
    // Make room for `need` bytes (lock held).  Nothing is evicted unless the eviction of idle resident models
    // (least recently used first) is enough; a model larger than the whole budget (its measured footprint, which
    // can exceed the estimate it was admitted with) is loaded once nothing else is resident, so its requests never
    // wait forever.
    bool makeRoom(size_t need, int except) {
        size_t idle = 0;
        for (size_t i = 0; i < models.size(); ++i) {
            const Model& m = *models[i];
            if (int(i) != except && m.S && !m.busy) idle += m.footprint;
        }
        const size_t pinned = used - std::min(used, idle);    // busy models stay
        const bool alone = need > budget && pinned == 0;
        if (pinned + need > budget && !alone) return false;
        while (used + need > budget && used > 0) {
            int victim = -1;
            for (size_t i = 0; i < models.size(); ++i) {
                const Model& m = *models[i];
                if (int(i) == except || !m.S || m.busy) continue;
                if (victim < 0 || m.lastUse < models[size_t(victim)]->lastUse) victim = int(i);
            }
            if (victim < 0) break;
            Model& v = *models[size_t(victim)];
            rsh_system_destroy(v.S);
            v.S = nullptr;
            v.base.release();
            v.x.release();
            v.b.release();
            used -= std::min(used, v.footprint);
            ++evictions;
        }
        return true;
    }
    // next admissible request in arrival order (lock held): its model idle and resident or fitting the budget.
    // Once a request waits for memory, later requests that would also load a model are held back behind it (and
    // behind a model larger than the budget, every later request), so the waiting one is not starved.
    long long pick() {
        bool loadBlocked = false, allBlocked = false;
        for (auto it = queue.begin(); it != queue.end(); ++it) {
            Req& q = *reqs.at(*it);
            Model& m = *models[size_t(q.r.model)];
            if (m.busy) continue;
            if (!m.S) {
                if (allBlocked || loadBlocked) continue;
                if (!makeRoom(m.footprint, q.r.model)) {
                    loadBlocked = true;
                    allBlocked = m.footprint > budget;
                    continue;
                }
            } else if (allBlocked) {
                continue;
            }
            const long long id = *it;
            queue.erase(it);
            return id;
        }
        return 0;
    }
    void serve(Req& q, Model& m, rsh_context* ctx) {
        hipStream_t st = ctx->c.stream;
        RSH_HIP(hipSetDevice(device));
        if (!m.S) {
            // (re)load: pattern, base values, symbolic set-up on the first setup below
            rsh_system* S = nullptr;
            const int rc = rsh_system_create(ctx, m.n, m.nb, m.rp.data(), m.ci.data(), &S);
            if (rc != RSH_OK) throw Error(rc, g_err);
            m.S = S;
            m.base.upload(m.baseHost, st);
            m.x.alloc(size_t(m.n) * m.nb);
            m.b.alloc(size_t(m.n) * m.nb);
            ++m.loads;
        }
        rsh_system& S = *m.S;
        S.ctx = ctx;                          // this request runs on the worker's stream
        const long long nv = (long long)m.nnzb * m.nb * m.nb;
        const double* src = q.r.values ? q.r.values : m.base.p;
        if (q.r.values && !q.r.values_on_device) {
            RSH_HIP(hipMemcpyAsync(S.val.p, q.r.values, size_t(nv) * sizeof(double), hipMemcpyHostToDevice, st));
            src = S.val.p;
        }
        if (q.r.perturb > 0) {
            launchPerturbRows(st, m.n, S.rp.p, m.nb, src, S.val.p, q.r.perturb, q.r.perturb_seed);
        } else if (src != S.val.p) {
            RSH_HIP(hipMemcpyAsync(S.val.p, src, size_t(nv) * sizeof(double), hipMemcpyDeviceToDevice, st));
        }
        S.valuesChanged();
        int rc = rsh_system_setup(&S, &m.opt);
        if (rc != RSH_OK) throw Error(rc, g_err);
        const size_t N = size_t(m.n) * m.nb;
        const double* bd = m.b.p;
        if (q.r.rhs_on_device) bd = q.r.rhs;
        else RSH_HIP(hipMemcpyAsync(m.b.p, q.r.rhs, N * sizeof(double), hipMemcpyHostToDevice, st));
        double* xd = q.r.x_on_device ? q.r.x : m.x.p;
        int its = 0;
        double rel = 0;
        rc = rsh_solve_device(&S, bd, xd, q.r.tol, q.r.max_it, &its, &rel);
        if (!q.r.x_on_device && q.r.x) RSH_HIP(hipMemcpyAsync(q.r.x, m.x.p, N * sizeof(double), hipMemcpyDeviceToHost, st));
        RSH_HIP(hipStreamSynchronize(st));
        q.res.status = rc;
        if (rc != RSH_OK) q.msg = g_err;
        q.res.iterations = its;
        q.res.rel_res = rel;
    }
    // models -> groups (lock held; at the first request)
    void formGroups() {
        std::vector<int> ord(models.size());
        for (size_t i = 0; i < ord.size(); ++i) ord[i] = int(i);
        auto same = [&](int a, int b) {
            const Model& A = *models[size_t(a)];
            const Model& B = *models[size_t(b)];
            return A.nb == B.nb && std::memcmp(&A.opt, &B.opt, sizeof A.opt) == 0;
        };
        std::stable_sort(ord.begin(), ord.end(), [&](int a, int b) {
            const Model& A = *models[size_t(a)];
            const Model& B = *models[size_t(b)];
            if (A.nb != B.nb) return A.nb < B.nb;
            const int c = std::memcmp(&A.opt, &B.opt, sizeof A.opt);
            if (c != 0) return c < 0;
            return A.n < B.n;
        });
        groupOf.assign(models.size(), -1);
        for (size_t k = 0; k < ord.size();) {
            auto g = std::make_unique<Group>();
            long long cells = 0;
            size_t e = k;
            while (e < ord.size() && same(ord[k], ord[e]) && int(g->models.size()) < groupSystems &&
                   (g->models.empty() || cells + models[size_t(ord[e])]->n <= groupCells)) {
                cells += models[size_t(ord[e])]->n;
                g->models.push_back(ord[e]);
                ++e;
            }
            for (int m : g->models) {
                groupOf[size_t(m)] = int(groups.size());
                g->footprint += models[size_t(m)]->footprint;
            }
            groups.push_back(std::move(g));
            k = e;
        }
        grouped = true;
    }
    // next round (lock held): the first queued request whose group is idle, with every queued request of that group
    // (the earliest per model).  Groups are not evicted: a group that is not resident and does not fit the budget left
    // while others are resident gets its requests back as refused (refuse = true) instead of waiting for room that
    // never comes
    int pickRound(std::vector<long long>& ids, bool& refuse) {
        ids.clear();
        refuse = false;
        for (long long id : queue) {
            const int g = groupOf[size_t(reqs.at(id)->r.model)];
            Group& G = *groups[size_t(g)];
            if (G.busy) continue;
            const bool resident = G.models.size() == 1 ? models[size_t(G.models[0])]->S != nullptr : G.B != nullptr;
            refuse = !resident && used + G.footprint > budget && used > 0;
            std::vector<char> taken(models.size(), 0);
            for (auto it = queue.begin(); it != queue.end();) {
                const int m = reqs.at(*it)->r.model;
                if (groupOf[size_t(m)] == g && !taken[size_t(m)]) {
                    taken[size_t(m)] = 1;
                    ids.push_back(*it);
                    it = queue.erase(it);
                } else {
                    ++it;
                }
            }
            return g;
        }
        return -1;
    }
    void serveRound(Group& G, const std::vector<Req*>& qs) {
        RSH_HIP(hipSetDevice(device));
        const int cnt = int(G.models.size());
        if (!G.B) {
            int rc = G.ctx ? RSH_OK : rsh_context_create(device, &G.ctx);
            if (rc != RSH_OK) {
                G.ctx = nullptr;
                throw Error(rc, g_err);
            }
            const size_t nc = static_cast<size_t>(cnt);
            std::vector<int> n(nc, 0), nb(nc, 0);
            std::vector<const int*> rp(nc, nullptr), ci(nc, nullptr);
            for (int j = 0; j < cnt; ++j) {
                const Model& m = *models[size_t(G.models[size_t(j)])];
                n[size_t(j)] = m.n;
                nb[size_t(j)] = m.nb;
                rp[size_t(j)] = m.rp.data();
                ci[size_t(j)] = m.ci.data();
            }
            rc = rsh_batch_create(G.ctx, cnt, n.data(), nb.data(), rp.data(), ci.data(), &G.B);
            if (rc != RSH_OK) throw Error(rc, g_err);
            hipStream_t st = batchStream(G.B);
            // base values in the group layout, rhs / solution in input order
            long long vals = 0, vec = 0;
            for (int j = 0; j < cnt; ++j) {
                const BatchMember q = batchMember(G.B, j);
                vals = std::max(vals, q.valOff + (long long)q.nnzb * q.nb * q.nb);
                vec = std::max(vec, q.inOff + (long long)q.n * q.nb);
            }
            G.base.alloc(size_t(vals));
            for (int j = 0; j < cnt; ++j) {
                const BatchMember q = batchMember(G.B, j);
                const Model& m = *models[size_t(G.models[size_t(j)])];
                RSH_HIP(hipMemcpyAsync(G.base.p + q.valOff, m.baseHost.data(), m.baseHost.size() * sizeof(double), hipMemcpyHostToDevice, st));
                // every member starts from its base values (its factors must exist even when it has no request)
                RSH_HIP(hipMemcpyAsync(q.groupVal + q.valOff, G.base.p + q.valOff, m.baseHost.size() * sizeof(double), hipMemcpyDeviceToDevice, st));
            }
            G.b.alloc(size_t(vec));
            G.x.alloc(size_t(vec));
            G.tolH.assign(size_t(cnt), 1.0);
            RSH_HIP(hipStreamSynchronize(st));
        }
        hipStream_t st = batchStream(G.B);
        launchFill(st, G.b.p, 0.0, (long long)G.b.n);
        std::fill(G.tolH.begin(), G.tolH.end(), 1.0);
        int maxIt = 1;
        std::vector<int> member(qs.size());
        for (size_t k = 0; k < qs.size(); ++k) {
            const rsh_request& r = qs[k]->r;
            int j = 0;
            while (G.models[size_t(j)] != r.model) ++j;
            member[k] = j;
            const BatchMember q = batchMember(G.B, j);
            const long long nv = (long long)q.nnzb * q.nb * q.nb;
            double* dst = q.groupVal + q.valOff;
            const double* src = G.base.p + q.valOff;
            if (r.values) {
                RSH_HIP(hipMemcpyAsync(dst, r.values, size_t(nv) * sizeof(double), r.values_on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, st));
                if (r.perturb > 0) launchPerturbRows(st, q.n, q.rp, q.nb, q.groupVal, q.groupVal, r.perturb, r.perturb_seed);
            } else if (r.perturb > 0) {
                launchPerturbRows(st, q.n, q.rp, q.nb, G.base.p, q.groupVal, r.perturb, r.perturb_seed);
            } else {
                RSH_HIP(hipMemcpyAsync(dst, src, size_t(nv) * sizeof(double), hipMemcpyDeviceToDevice, st));
            }
            RSH_HIP(hipMemcpyAsync(G.b.p + q.inOff, r.rhs, sizeof(double) * size_t(q.n) * q.nb, r.rhs_on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, st));
            G.tolH[size_t(j)] = r.tol;
            maxIt = std::max(maxIt, r.max_it);
        }
        batchValuesChanged(G.B);
        const rsh_options& opt = models[size_t(G.models[0])]->opt;
        int rc = rsh_batch_setup(G.B, &opt);
        if (rc != RSH_OK) throw Error(rc, g_err);
        const size_t nc = static_cast<size_t>(cnt);
        std::vector<int> its(nc, 0), status(nc, 0);
        std::vector<double> rel(nc, 0.0);
        rc = rsh_batch_solve_device(G.B, G.b.p, G.x.p, G.tolH.data(), maxIt, its.data(), rel.data(), status.data());
        if (rc != RSH_OK && rc != RSH_ERR_NOT_CONVERGED && rc != RSH_ERR_BREAKDOWN && rc != RSH_ERR_NONFINITE) throw Error(rc, g_err);
        for (size_t k = 0; k < qs.size(); ++k) {
            const int j = member[k];
            const BatchMember q = batchMember(G.B, j);
            const rsh_request& r = qs[k]->r;
            if (r.x)
                RSH_HIP(hipMemcpyAsync(r.x, G.x.p + q.inOff, sizeof(double) * size_t(q.n) * q.nb, r.x_on_device ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost, st));
            qs[k]->res.status = status[size_t(j)];
            qs[k]->res.iterations = its[size_t(j)];
            qs[k]->res.rel_res = rel[size_t(j)];
            if (status[size_t(j)] != RSH_OK) qs[k]->msg = "the request's system did not converge in its batched round";
        }
        RSH_HIP(hipStreamSynchronize(st));
    }
    void batchedLoop(int w) {
        for (;;) {
            std::vector<long long> ids;
            std::vector<Req*> qs;
            int g = -1;
            bool wasResident = false, refuse = false;
            {
                std::unique_lock<std::mutex> lk(mu);
                cvWork.wait(lk, [&] {
                    if (stop) return true;
                    if (!grouped && !queue.empty()) formGroups();
                    return grouped && (g = pickRound(ids, refuse)) >= 0;
                });
                if (g < 0) return;                   // stop with nothing admissible
                Group& G = *groups[size_t(g)];
                G.busy = true;
                wasResident = G.models.size() == 1 ? models[size_t(G.models[0])]->S != nullptr : G.B != nullptr;
                if (!wasResident && !refuse) used += G.footprint;
                const double now = nowMs();
                for (long long id : ids) {
                    Req* q = reqs.at(id).get();
                    q->tStart = now;
                    qs.push_back(q);
                    models[size_t(q->r.model)]->busy = true;
                }
                running += int(qs.size());
                peakRunning = std::max(peakRunning, running);
                ++rounds;
            }
            Group& G = *groups[size_t(g)];
            const bool single = G.models.size() == 1;   // a model alone: the resident single-system path
            int code = RSH_OK;
            std::string msg;
            try {
                if (refuse) throw Error(RSH_ERR_UNSUPPORTED, "batched server: the request's model group does not fit the HBM budget left (groups are "
                                                             "not evicted; size the budget for every group)");
                if (single) serve(*qs[0], *models[size_t(G.models[0])], ctxs[size_t(w)]);
                else serveRound(G, qs);
            } catch (const Error& e) {
                code = e.code;
                msg = e.what();
            } catch (const std::exception& e) {
                code = RSH_ERR_DEVICE;
                msg = e.what();
            }
            {
                std::lock_guard<std::mutex> lk(mu);
                const Model* sm = single ? models[size_t(G.models[0])].get() : nullptr;
                const bool loaded = single ? sm->S != nullptr : G.B != nullptr;
                if (loaded && !wasResident) {
                    used -= std::min(used, G.footprint);
                    G.footprint = single ? measure(*sm) : batchDeviceBytes(G.B) + G.base.bytes() + G.b.bytes() + G.x.bytes();
                    used += G.footprint;
                } else if (!loaded && !wasResident && !refuse) {
                    used -= std::min(used, G.footprint);
                }
                ++G.rounds;
                const double now = nowMs();
                for (Req* q : qs) {
                    if (code != RSH_OK) {
                        q->res.status = code;
                        q->msg = msg;
                    }
                    q->tEnd = now;
                    q->res.queue_ms = q->tStart - q->tSubmit;
                    q->res.service_ms = q->tEnd - q->tStart;
                    q->res.latency_ms = q->tEnd - q->tSubmit;
                    q->done = true;
                    Model& m = *models[size_t(q->r.model)];
                    m.busy = false;
                    m.lastUse = now;
                }
                G.busy = false;
                running -= int(qs.size());
                completed += (long long)qs.size();
            }
            cvDone.notify_all();
            cvWork.notify_all();
        }
    }
    void workerLoop(int w) {
        if (batched) { batchedLoop(w); return; }
        rsh_context* ctx = ctxs[size_t(w)];
        for (;;) {
            long long id = 0;
            Req* qp = nullptr;
            Model* mp = nullptr;
            bool wasResident = false;
            {
                std::unique_lock<std::mutex> lk(mu);
                cvWork.wait(lk, [&] { return stop || (id = pick()) != 0; });
                if (!id) return;                  // stop with nothing admissible
                qp = reqs.at(id).get();
                mp = models[size_t(qp->r.model)].get();
                mp->busy = true;
                wasResident = mp->S != nullptr;
                if (!wasResident) used += mp->footprint;    // reserved now, corrected after the load
                qp->tStart = nowMs();
                peakRunning = std::max(peakRunning, ++running);
            }
            Req& q = *qp;
            Model& m = *mp;
            try {
                serve(q, m, ctx);
            } catch (const Error& e) {
                q.res.status = e.code;
                q.msg = e.what();
            } catch (const std::exception& e) {
                q.res.status = RSH_ERR_DEVICE;
                q.msg = e.what();
            }
            {
                std::lock_guard<std::mutex> lk(mu);
                if (m.S) {
                    const size_t real = measure(m);
                    if (!wasResident || real != m.footprint) {
                        used -= std::min(used, m.footprint);
                        m.footprint = real;
                        used += real;
                    }
                } else if (!wasResident) {
                    used -= std::min(used, m.footprint);   // the load failed: release the reservation
                }
                q.tEnd = nowMs();
                q.res.queue_ms = q.tStart - q.tSubmit;
                q.res.service_ms = q.tEnd - q.tStart;
                q.res.latency_ms = q.tEnd - q.tSubmit;
                q.done = true;
                m.busy = false;
                m.lastUse = q.tEnd;
                --running;
                ++completed;
            }
            cvDone.notify_all();
            cvWork.notify_all();
        }
    }
};

extern "C" {

static int serverCreate(int device, int slots, double hbm_budget_bytes, bool batched, long long groupCells, int groupSystems,
                        rsh_server** out) {
    return guarded([&] {
        if (!out) return fail(RSH_ERR_ARG, "null output pointer");
        *out = nullptr;
        if (slots < 1 || slots > 16) return fail(RSH_ERR_ARG, "slots must be 1..16");
        if (!(hbm_budget_bytes > 0)) return fail(RSH_ERR_ARG, "the HBM budget must be positive");
        auto S = std::make_unique<rsh_server>();
        S->device = device;
        S->budget = size_t(hbm_budget_bytes);
        S->batched = batched;
        S->groupCells = groupCells;
        S->groupSystems = groupSystems;
        for (int w = 0; w < slots; ++w) {
            rsh_context* c = nullptr;
            const int rc = rsh_context_create(device, &c);
            if (rc != RSH_OK) {
                for (auto* x : S->ctxs) rsh_context_destroy(x);
                return rc;
            }
            S->ctxs.push_back(c);
        }
        rsh_server* raw = S.get();
        for (int w = 0; w < slots; ++w) S->workers.emplace_back([raw, w] { raw->workerLoop(w); });
        *out = S.release();
        return RSH_OK;
    });
}

int rsh_server_create(int device, int slots, double hbm_budget_bytes, rsh_server** out) {
    return serverCreate(device, slots, hbm_budget_bytes, false, 0, 0, out);
}

int rsh_server_create_batched(int device, int slots, double hbm_budget_bytes, long long group_cells, int group_systems, rsh_server** out) {
    if (group_cells < 1 || group_systems < 1 || group_systems > 1024)
        return fail(RSH_ERR_ARG, "group_cells must be positive and group_systems 1..1024");
    return serverCreate(device, slots, hbm_budget_bytes, true, group_cells, group_systems, out);
}

int rsh_server_destroy(rsh_server* srv) {
    if (!srv) return RSH_OK;
    return guarded([&] {
        {
            std::unique_lock<std::mutex> lk(srv->mu);
            srv->cvDone.wait(lk, [&] { return srv->queue.empty() && srv->running == 0; });
            srv->stop = true;
        }
        srv->cvWork.notify_all();
        for (auto& t : srv->workers) t.join();
        for (auto& m : srv->models)
            if (m->S) rsh_system_destroy(m->S);
        for (auto& g : srv->groups) {
            if (g->B) rsh_batch_destroy(g->B);
            g->base.release();
            g->b.release();
            g->x.release();
            if (g->ctx) rsh_context_destroy(g->ctx);
        }
        srv->groups.clear();
        srv->models.clear();
        for (auto* c : srv->ctxs) rsh_context_destroy(c);
        delete srv;
        return RSH_OK;
    });
}

int rsh_server_add_model(rsh_server* srv, int n, int nb, const int* row_ptr, const int* col, const double* values, const rsh_options* opt,
                         int* model) {
    return guarded([&] {
        if (!srv || !row_ptr || !col || !values || !opt || !model) return fail(RSH_ERR_ARG, "null argument");
        if (n <= 0 || nb < 1 || nb > kMaxNb) return fail(RSH_ERR_ARG, "bad system size");
        {
            std::lock_guard<std::mutex> lk(srv->mu);
            if (srv->grouped) return fail(RSH_ERR_ARG, "batched server: models are grouped at the first request; add every model before it");
        }
        auto m = std::make_unique<rsh_server::Model>();
        m->n = n;
        m->nb = nb;
        m->nnzb = row_ptr[n];
        m->rp.assign(row_ptr, row_ptr + n + 1);
        m->ci.assign(col, col + m->nnzb);
        m->baseHost.assign(values, values + size_t(m->nnzb) * nb * nb);
        m->opt = *opt;
        m->footprint = rsh_server::estimate(*m);
        // a model larger than the budget is accepted: it is served alone (makeRoom), so its requests cannot starve
        std::lock_guard<std::mutex> lk(srv->mu);
        *model = int(srv->models.size());
        srv->models.push_back(std::move(m));
        return RSH_OK;
    });
}

int rsh_server_submit(rsh_server* srv, const rsh_request* req, long long* ticket) {
    return guarded([&] {
        if (!srv || !req || !ticket) return fail(RSH_ERR_ARG, "null argument");
        if (!req->rhs) return fail(RSH_ERR_ARG, "null right-hand side");
        std::lock_guard<std::mutex> lk(srv->mu);
        if (req->model < 0 || req->model >= int(srv->models.size())) return fail(RSH_ERR_ARG, "unknown model");
        auto q = std::make_unique<rsh_server::Req>();
        q->id = srv->nextId++;
        q->r = *req;
        q->tSubmit = nowMs();
        *ticket = q->id;
        srv->queue.push_back(q->id);
        srv->reqs.emplace(q->id, std::move(q));
        srv->cvWork.notify_one();
        return RSH_OK;
    });
}

// blocks until the request is done; the request's record is released.  The record is looked up again by its
// ticket after the wait: other threads submit (and insert into the map) while this one waits.
int rsh_server_wait(rsh_server* srv, long long ticket, rsh_result* res) {
    return guarded([&] {
        if (!srv || !res) return fail(RSH_ERR_ARG, "null argument");
        std::unique_lock<std::mutex> lk(srv->mu);
        if (srv->reqs.find(ticket) == srv->reqs.end()) return fail(RSH_ERR_ARG, "unknown ticket");
        rsh_server::Req* q = srv->reqs.at(ticket).get();   // the Req object itself never moves
        srv->cvDone.wait(lk, [&] { return q->done; });
        *res = q->res;
        const std::string msg = q->msg;
        srv->reqs.erase(ticket);
        if (res->status != RSH_OK && !msg.empty()) g_err = msg;   // rsh_last_error() in the waiting thread
        return RSH_OK;
    });
}

int rsh_server_stats(rsh_server* srv, char* json, int cap) {
    return guarded([&] {
        if (!srv || !json || cap <= 0) return fail(RSH_ERR_ARG, "null argument");
        std::lock_guard<std::mutex> lk(srv->mu);
        int resident = 0, loads = 0;
        for (auto& m : srv->models) {
            resident += m->S != nullptr;
            loads += m->loads;
        }
        std::ostringstream o;
        o << "{\"models\":" << srv->models.size() << ",\"resident\":" << resident << ",\"loads\":" << loads << ",\"evictions\":" << srv->evictions
          << ",\"hbm_used\":" << srv->used << ",\"hbm_budget\":" << srv->budget << ",\"slots\":" << srv->ctxs.size() << ",\"queued\":"
          << srv->queue.size() << ",\"running\":" << srv->running << ",\"peak_running\":" << srv->peakRunning << ",\"completed\":" << srv->completed
          << ",\"batched\":" << (srv->batched ? 1 : 0) << ",\"groups\":" << srv->groups.size() << ",\"rounds\":" << srv->rounds
          << ",\"footprints\":[";
        for (size_t i = 0; i < srv->models.size(); ++i) o << (i ? "," : "") << srv->models[i]->footprint;
        o << "],\"group_models\":[";
        for (size_t i = 0; i < srv->groups.size(); ++i) o << (i ? "," : "") << srv->groups[i]->models.size();
        o << "],\"group_rounds\":[";
        for (size_t i = 0; i < srv->groups.size(); ++i) o << (i ? "," : "") << srv->groups[i]->rounds;
        o << "]}";
        const std::string s = o.str();
        if (int(s.size()) + 1 > cap) return fail(RSH_ERR_ARG, "buffer too small");
        std::memcpy(json, s.c_str(), s.size() + 1);
        return RSH_OK;
    });
}

}  // extern "C"
