// Serving loop: continuous batching of solve requests with admission by HBM
// footprint and device-resident problem state.
//
// Requests queue on the host.  Each step() (1) retires in-flight groups whose
// solve finished (non-blocking event polls), (2) starts the solve of every
// group whose batch has been built, on its own stream, (3) admits queued
// requests of one class (element type, right-hand-side count, load kind) as
// a new group while solve slots and the HBM budget allow, and hands the
// build of its batch (pattern, layout, assembly on the device) to a host
// worker thread, and (4) enqueues one more chunk of iterations for every
// in-flight group.  Builds and the release of finished batches (hipFree
// synchronises the device) run off the serving thread, so a step never
// waits for them and in-flight groups keep their chunks coming.  Groups
// admitted at different times run concurrently on the device, each problem
// with its own scalars and reductions: a problem's result does not depend
// on which other problems share its group.
// Requests with a persistent key keep their batch (matrix, damage state and
// solution) resident between requests.  Each such request is one load step
// in the order of the reference's FeatureTree::stepInternal
// (features/features.cpp:6150-6373: boundary conditions, assembly, solve,
// then the element state update): the request's Dirichlet values become the
// resident problem's, the matrix is re-assembled with the damage so far,
// the solve starts from the previous solution, and after it the damage law
// (the request's parameters) updates the damage for the next step.  A
// request that cannot be run completes with a negative status instead of
// stopping the server: a mesh or Dirichlet DOFs that differ from the
// resident problem's, a batch build that fails, or a footprint larger than
// what the budget leaves next to idle resident problems (release keys with
// amie_gpu_server_release).
#include <algorithm>
#include <chrono>
#include <cstring>
#include <deque>
#include <future>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "batch.hpp"

namespace amie {

// deep copy of a problem description
struct OwnedProblem {
    amie_problem_desc d{};
    std::vector<double> coords, cmat, damage, bc_vals, forces, macro;
    std::vector<int32_t> conn, dofmap, phase;
    std::vector<int64_t> bc_dofs;
    int64_t footprint = 0;

    void copy(const amie_problem_desc &s)
    {
        const EType et = etype_info(s.elem_type);
        d = s;
        const int S = et.s;
        coords.assign(s.coords, s.coords + s.nnodes * et.dim);
        conn.assign(s.conn, s.conn + s.nelem * et.nper);
        cmat.assign(s.cmat, s.cmat + (size_t)s.nphase * S * S);
        if (s.dofmap) dofmap.assign(s.dofmap, s.dofmap + s.nnodes);
        if (s.phase) phase.assign(s.phase, s.phase + s.nelem);
        if (s.damage) damage.assign(s.damage, s.damage + s.nelem);
        if (s.nbc) {
            bc_dofs.assign(s.bc_dofs, s.bc_dofs + s.nbc);
            bc_vals.assign(s.bc_vals, s.bc_vals + s.nbc);
        }
        const int64_t nd = (s.dofmap ? s.ndofnodes : s.nnodes) * et.b;
        if (s.forces) forces.assign(s.forces, s.forces + nd * s.nrhs);
        if (s.macro_strain) macro.assign(s.macro_strain, s.macro_strain + (size_t)s.nrhs * S);
        d.coords = coords.data();
        d.conn = conn.data();
        d.cmat = cmat.data();
        d.dofmap = s.dofmap ? dofmap.data() : nullptr;
        d.phase = s.phase ? phase.data() : nullptr;
        d.damage = s.damage ? damage.data() : nullptr;
        d.bc_dofs = s.nbc ? bc_dofs.data() : nullptr;
        d.bc_vals = s.nbc ? bc_vals.data() : nullptr;
        d.forces = s.forces ? forces.data() : nullptr;
        d.macro_strain = s.macro_strain ? macro.data() : nullptr;
        // HBM estimate: blocks per row of the element type, FP64 + FP32
        // values, column index, 11 vectors, element data, the gather
        // assembly's lists (per slot, per element block, per row)
        const int64_t rows = s.dofmap ? s.ndofnodes : s.nnodes;
        const int bpr = et.et == 1 ? 7 : (et.et == 2 ? 12 : 15);
        const int64_t slots = rows * bpr * 5 / 4;
        footprint = slots * (int64_t)(8 * et.b * et.b + 4 * et.b * et.b + 4 + 4) + rows * et.b * s.nrhs * 8 * 11 +
                    s.nelem * (int64_t)(2 * et.nper * et.nper * 4 + et.nper * 8 + 40) + s.nnodes * et.dim * 8 +
                    (s.macro_strain ? rows * 4 + s.nelem * et.nper * 4 : 0) + (1 << 20);
    }
    int cls() const { return d.elem_type * 64 + d.nrhs * 2 + (d.macro_strain ? 1 : 0); }
};

struct Request {
    uint64_t ticket = 0;
    uint64_t key = 0;
    OwnedProblem prob;
    amie_damage_params prm{1e-4, 1e-3, 1. - 1e-6};   // the damage law of the incremental-loading case
    double submit_ms = 0., start_ms = 0.;
};

struct Group {
    std::unique_ptr<Batch> batch;
    std::vector<Request> reqs;
    hipStream_t stream = nullptr;
    bool persistent = false;
    size_t bytes = 0;
    // persistent problems: what later requests on the key must share
    int32_t elem_type = 0, nrhs = 0;
    int64_t nnodes = 0, nelem = 0;
    std::vector<int64_t> bc_dofs;
};

// A batch built on a worker thread.
struct Built {
    Batch *batch = nullptr;
    int rc = AMIE_OK;
    std::string err;
};

struct Building {
    std::unique_ptr<Group> g;
    std::future<Built> fut;
};

template <typename T> static bool ready(std::future<T> &f)
{
    return f.wait_for(std::chrono::seconds(0)) == std::future_status::ready;
}

struct Server {
    amie_gpu_ctx *ctx = nullptr;
    size_t budget = 0, in_use = 0;
    int max_slots = 1, used_slots = 0;
    amie_solve_opts opts{};
    uint64_t next_ticket = 1;
    std::deque<Request> queue;
    std::vector<std::unique_ptr<Group>> inflight;
    std::vector<Building> building;              // batches being built on worker threads
    std::vector<std::future<void>> releasing;    // finished batches being freed on worker threads
    std::map<uint64_t, std::unique_ptr<Group>> resident;   // persistent key -> idle resident batch
    std::vector<amie_request_result> done;
    std::map<uint64_t, std::vector<double>> solutions;
    std::vector<hipStream_t> streams;
    size_t next_stream = 0;

    ~Server()
    {
        for (auto &b : building) delete b.fut.get().batch;
        building.clear();
        releasing.clear();   // each future waits for its release
        for (auto &g : inflight)
            if (g->batch) (void)hipStreamSynchronize(g->batch->stream);
        inflight.clear();
        resident.clear();
        for (auto s : streams) (void)hipStreamDestroy(s);
    }

    hipStream_t pick_stream()
    {
        hipStream_t s = streams[next_stream % streams.size()];
        next_stream++;
        return s;
    }

    int start_group(std::unique_ptr<Group> g, double now)
    {
        for (auto &r : g->reqs) r.start_ms = now;
        int rc;
        amie_solve_opts o = opts;
        if (g->persistent) o.use_x0 = 1;
        if ((rc = g->batch->solve_begin(o))) return rc;
        if ((rc = g->batch->solve_advance())) return rc;
        inflight.push_back(std::move(g));
        return AMIE_OK;
    }

    // a request that is not run: its result carries the error status
    void fail(const Request &r, int status, double now)
    {
        amie_request_result o{};
        o.ticket = r.ticket;
        o.status = status;
        o.iterations = 0;
        o.relres = 0.;
        o.submit_ms = r.submit_ms;
        o.start_ms = now;
        o.done_ms = now;
        done.push_back(o);
    }

    void release_group(std::unique_ptr<Group> g)
    {
        in_use -= g->bytes;
        const int dev = ctx->device;
        Group *gp = g.release();
        releasing.push_back(std::async(std::launch::async, [gp, dev]() {
            (void)hipSetDevice(dev);
            delete gp;
        }));
    }

    static bool same_problem(const Group &g, const Request &r)
    {
        const amie_problem_desc &d = r.prob.d;
        return d.elem_type == g.elem_type && d.nnodes == g.nnodes && d.nelem == g.nelem && d.nrhs == g.nrhs &&
               r.prob.bc_dofs == g.bc_dofs;
    }

    int retire(std::unique_ptr<Group> &g, double now)
    {
        amie_solve_stats st{};
        int rc = g->batch->solve_end(&st);
        if (rc < 0 && rc != AMIE_ERR_NAN) return rc;
        Batch &B = *g->batch;
        for (size_t i = 0; i < g->reqs.size(); i++) {
            const ProbDev &P = B.h_prob[i];
            amie_request_result r{};
            r.ticket = g->reqs[i].ticket;
            r.status = P.status == ST_CONVERGED ? AMIE_OK : (P.status == ST_NAN ? AMIE_ERR_NAN : AMIE_NOT_CONVERGED);
            r.iterations = (int32_t)P.iters;
            double rel = 0.;
            for (int k = 0; k < B.nrhs; k++) rel = std::max(rel, P.relres_out[k]);
            r.relres = rel;
            r.submit_ms = g->reqs[i].submit_ms;
            r.start_ms = g->reqs[i].start_ms;
            r.done_ms = now;
            done.push_back(r);
            // solution of the request (original order, nrhs columns)
            const ProbHost &H = B.ph[i];
            std::vector<double> x((size_t)H.nrows * B.et.b * B.nrhs);
            if (int rc = download_vec(B, (int)i, B.x64, x.data())) return rc;
            solutions[r.ticket] = std::move(x);
        }
        used_slots -= (int)g->reqs.size();
        if (g->persistent) {
            // the element state update closes the load step; the matrix is
            // re-assembled when the next step's loads are known
            const uint64_t key = g->reqs[0].key;
            int64_t grown = 0;
            if (int rc = g->batch->damage_update(g->reqs[0].prm, &grown)) return rc;
            g->reqs.clear();
            resident[key] = std::move(g);   // stays in HBM, still counted in in_use
        } else {
            release_group(std::move(g));
        }
        return AMIE_OK;
    }

    // (2) groups whose batch is built start their solve
    int start_built(double now)
    {
        for (size_t i = 0; i < building.size();) {
            if (!ready(building[i].fut)) {
                i++;
                continue;
            }
            Built b = building[i].fut.get();
            std::unique_ptr<Group> g = std::move(building[i].g);
            building.erase(building.begin() + i);
            if (b.rc) {
                // the group's requests complete with the build's error
                in_use -= g->bytes;
                used_slots -= (int)g->reqs.size();
                set_error("%s", b.err.c_str());
                for (auto &r : g->reqs) fail(r, b.rc, now);
                continue;
            }
            g->batch.reset(b.batch);
            b.batch->stream = pick_stream();
            in_use -= g->bytes;   // the estimate made at admission
            g->bytes = b.batch->bytes();
            in_use += g->bytes;
            if (int rc = start_group(std::move(g), now)) return rc;
        }
        for (size_t i = 0; i < releasing.size();) {
            if (ready(releasing[i])) releasing.erase(releasing.begin() + i);
            else i++;
        }
        return AMIE_OK;
    }

    bool key_busy(uint64_t key) const
    {
        if (resident.count(key)) return true;
        for (auto &g : inflight)
            if (g->persistent && !g->reqs.empty() && g->reqs[0].key == key) return true;
        for (auto &b : building)
            if (b.g->persistent && !b.g->reqs.empty() && b.g->reqs[0].key == key) return true;
        return false;
    }

    int step(double now, int *n_active, int *n_queued)
    {
        int rc;
        // (1) retire finished groups
        for (size_t i = 0; i < inflight.size();) {
            bool fin = false;
            if ((rc = inflight[i]->batch->solve_poll(false, &fin))) return rc;
            if (fin) {
                if ((rc = retire(inflight[i], now))) return rc;
                inflight.erase(inflight.begin() + i);
            } else {
                i++;
            }
        }
        if ((rc = start_built(now))) return rc;
        // (3) admit: persistent requests whose key is resident and idle, then
        // one group of the class at the head of the queue
        for (size_t i = 0; i < queue.size();) {
            Request &r = queue[i];
            auto it = r.key ? resident.find(r.key) : resident.end();
            if (it != resident.end() && used_slots < max_slots) {
                std::unique_ptr<Group> g = std::move(it->second);
                resident.erase(it);
                Request req = std::move(r);
                queue.erase(queue.begin() + i);
                if (!same_problem(*g, req)) {
                    set_error("request %llu: its mesh or Dirichlet DOFs differ from the resident problem of key %llu",
                              (unsigned long long)req.ticket, (unsigned long long)req.key);
                    fail(req, AMIE_ERR_ARG, now);
                    resident[req.key] = std::move(g);
                    continue;
                }
                // this step's loads, then the matrix with the damage so far
                amie_gpu_batch wrap{g->batch.get()};
                const amie_problem_desc &d = req.prob.d;
                if ((rc = amie_gpu_batch_set_bc_values(&wrap, 0, d.nbc, d.bc_dofs, d.bc_vals))) return rc;
                if ((rc = g->batch->assemble())) return rc;
                g->reqs.push_back(std::move(req));
                used_slots += 1;
                if ((rc = start_group(std::move(g), now))) return rc;
            } else {
                i++;
            }
        }
        if (!queue.empty() && used_slots < max_slots) {
            const int cls = queue.front().prob.cls();
            const bool pers = queue.front().key != 0;
            std::vector<Request> pick;
            size_t bytes = 0;
            for (size_t i = 0; i < queue.size() && used_slots + (int)pick.size() < max_slots;) {
                Request &r = queue[i];
                const bool resident_key = r.key && key_busy(r.key);
                if (r.prob.cls() != cls || (r.key != 0) != pers || resident_key) { i++; continue; }
                if (in_use + bytes + (size_t)r.prob.footprint > budget) {
                    // nothing in flight or being built will free memory: what
                    // is in use belongs to idle resident problems, and this
                    // request cannot run until keys are released
                    if (pick.empty() && inflight.empty() && building.empty()) {
                        set_error("request %llu needs %lld bytes; idle resident problems hold %zu of the %zu budget",
                                  (unsigned long long)r.ticket, (long long)r.prob.footprint, in_use, budget);
                        fail(r, AMIE_ERR_NOMEM, now);
                        queue.erase(queue.begin() + i);
                        continue;
                    }
                    break;
                }
                bytes += (size_t)r.prob.footprint;
                pick.push_back(std::move(r));
                queue.erase(queue.begin() + i);
                if (pers) break;   // a persistent problem keeps a batch of its own
            }
            if (!pick.empty()) {
                std::unique_ptr<Group> g(new Group());
                g->persistent = pers;
                g->bytes = bytes;
                in_use += bytes;
                used_slots += (int)pick.size();
                g->reqs = std::move(pick);
                if (pers) {
                    const Request &r = g->reqs[0];
                    g->elem_type = r.prob.d.elem_type;
                    g->nrhs = r.prob.d.nrhs;
                    g->nnodes = r.prob.d.nnodes;
                    g->nelem = r.prob.d.nelem;
                    g->bc_dofs = r.prob.bc_dofs;
                }
                // the descriptions point into the requests' own arrays, which
                // stay in place while the group waits for its batch
                std::vector<amie_problem_desc> descs;
                for (auto &r : g->reqs) descs.push_back(r.prob.d);
                amie_gpu_ctx *cx = ctx;
                const amie_solve_opts so = opts;
                Building b;
                b.fut = std::async(std::launch::async, [cx, descs, so]() {
                    Built o;
                    if (hipSetDevice(cx->device) != hipSuccess) {
                        o.rc = AMIE_ERR_HIP;
                        o.err = "hipSetDevice failed on the build thread";
                        return o;
                    }
                    o.rc = batch_create(cx, (int)descs.size(), descs.data(), &o.batch);
                    // the SpMV candidates are timed here too, off the serving thread
                    // (under the load of the groups in flight)
                    const bool tunable = so.precision == AMIE_FP64 || so.precision == AMIE_MIXED ||
                                         so.precision == AMIE_MIXED_FIX48;
                    if (o.rc == AMIE_OK && so.autotune && tunable)
                        o.rc = o.batch->autotune(so.precision, so.precond == AMIE_BLOCK_JACOBI, 3);
                    if (o.rc) {
                        o.err = amie_gpu_last_error();
                        delete o.batch;
                        o.batch = nullptr;
                    }
                    return o;
                });
                b.g = std::move(g);
                building.push_back(std::move(b));
            }
        }
        // (4) one more chunk for every in-flight group
        for (auto &g : inflight)
            if ((rc = g->batch->solve_advance())) return rc;
        if (n_active) *n_active = used_slots;
        if (n_queued) *n_queued = (int)queue.size();
        return AMIE_OK;
    }
};

} // namespace amie

using namespace amie;

struct amie_gpu_server {
    amie::Server *s;
};

extern "C" {

int amie_gpu_server_create(amie_gpu_ctx *ctx, size_t hbm_budget, int32_t max_slots, const amie_solve_opts *opts,
                           amie_gpu_server **out)
{
    if (!ctx || !out || max_slots <= 0 || hbm_budget == 0) { set_error("bad arguments"); return AMIE_ERR_ARG; }
    AMIE_HIP(hipSetDevice(ctx->device));
    std::unique_ptr<Server> S(new Server());
    S->ctx = ctx;
    S->budget = hbm_budget;
    S->max_slots = max_slots;
    amie_gpu_default_opts(&S->opts);
    if (opts) S->opts = *opts;
    S->opts.time_kernels = 0;
    // default-priority streams, one per hardware queue
    for (int i = 0; i < 4; i++) {
        hipStream_t st;
        AMIE_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
        S->streams.push_back(st);
    }
    ctx_retain(ctx);
    *out = new amie_gpu_server{S.release()};
    return AMIE_OK;
}

void amie_gpu_server_destroy(amie_gpu_server *s)
{
    if (!s) return;
    amie_gpu_ctx *ctx = s->s->ctx;
    delete s->s;
    delete s;
    ctx_release(ctx);
}

static int submit(amie_gpu_server *s, const amie_problem_desc *desc, uint64_t key, const amie_damage_params *prm,
                  double submit_ms, uint64_t *ticket)
{
    if (!s || !desc) { set_error("bad arguments"); return AMIE_ERR_ARG; }
    const EType et = etype_info(desc->elem_type);
    if (!et.et || !desc->coords || !desc->conn || !desc->cmat) { set_error("bad problem"); return AMIE_ERR_ARG; }
    if (prm && !(prm->kappa0 > 0. && prm->kappaf > 0. && prm->dmax >= 0. && prm->dmax < 1.)) {
        set_error("damage parameters need kappa0 > 0, kappaf > 0, 0 <= dmax < 1");
        return AMIE_ERR_ARG;
    }
    Request r;
    r.key = key;
    r.submit_ms = submit_ms;
    if (prm) r.prm = *prm;
    r.prob.copy(*desc);
    if ((size_t)r.prob.footprint > s->s->budget) { set_error("request larger than the HBM budget"); return AMIE_ERR_NOMEM; }
    r.ticket = s->s->next_ticket++;
    if (ticket) *ticket = r.ticket;
    s->s->queue.push_back(std::move(r));
    return AMIE_OK;
}

int amie_gpu_server_submit(amie_gpu_server *s, const amie_problem_desc *desc, uint64_t key, double submit_ms,
                           uint64_t *ticket)
{
    return submit(s, desc, key, nullptr, submit_ms, ticket);
}

int amie_gpu_server_submit_step(amie_gpu_server *s, const amie_problem_desc *desc, uint64_t key,
                                const amie_damage_params *prm, double submit_ms, uint64_t *ticket)
{
    if (key == 0 || !prm) { set_error("a load step needs a persistent key and damage parameters"); return AMIE_ERR_ARG; }
    return submit(s, desc, key, prm, submit_ms, ticket);
}

int amie_gpu_server_request_bytes(const amie_problem_desc *desc, int64_t *bytes)
{
    if (!desc || !bytes) { set_error("bad arguments"); return AMIE_ERR_ARG; }
    const EType et = etype_info(desc->elem_type);
    if (!et.et || !desc->coords || !desc->conn || !desc->cmat) { set_error("bad problem"); return AMIE_ERR_ARG; }
    OwnedProblem o;
    o.copy(*desc);
    *bytes = o.footprint;
    return AMIE_OK;
}

int amie_gpu_server_release(amie_gpu_server *s, uint64_t key)
{
    if (!s || key == 0) { set_error("bad arguments"); return AMIE_ERR_ARG; }
    Server &S = *s->s;
    for (auto &r : S.queue)
        if (r.key == key) { set_error("key %llu has queued requests", (unsigned long long)key); return AMIE_ERR_BUSY; }
    auto it = S.resident.find(key);
    if (it == S.resident.end()) {
        if (S.key_busy(key)) { set_error("key %llu is in flight", (unsigned long long)key); return AMIE_ERR_BUSY; }
        set_error("no resident problem for key %llu", (unsigned long long)key);
        return AMIE_ERR_ARG;
    }
    std::unique_ptr<Group> g = std::move(it->second);
    S.resident.erase(it);
    S.release_group(std::move(g));
    return AMIE_OK;
}

int amie_gpu_server_get_damage(amie_gpu_server *s, uint64_t key, double *d, double *kappa, int64_t nelem)
{
    if (!s || key == 0) { set_error("bad arguments"); return AMIE_ERR_ARG; }
    Server &S = *s->s;
    auto it = S.resident.find(key);
    if (it == S.resident.end()) {
        if (S.key_busy(key)) { set_error("key %llu is in flight", (unsigned long long)key); return AMIE_ERR_BUSY; }
        set_error("no resident problem for key %llu", (unsigned long long)key);
        return AMIE_ERR_ARG;
    }
    if (nelem != it->second->nelem) { set_error("the resident problem has %lld elements", (long long)it->second->nelem); return AMIE_ERR_ARG; }
    amie_gpu_batch wrap{it->second->batch.get()};
    return amie_gpu_batch_get_damage(&wrap, 0, d, kappa);
}

int amie_gpu_server_step(amie_gpu_server *s, double now_ms, int32_t *n_active, int32_t *n_queued)
{
    if (!s) { set_error("server is NULL"); return AMIE_ERR_ARG; }
    int a = 0, q = 0;
    int rc = s->s->step(now_ms, &a, &q);
    if (n_active) *n_active = a;
    if (n_queued) *n_queued = q;
    return rc;
}

int amie_gpu_server_poll(amie_gpu_server *s, amie_request_result *out, int32_t max, int32_t *n)
{
    if (!s || !n || (max > 0 && !out)) { set_error("bad arguments"); return AMIE_ERR_ARG; }
    Server &S = *s->s;
    const int k = std::min<int>(max, (int)S.done.size());
    for (int i = 0; i < k; i++) out[i] = S.done[i];
    // solutions of results handed out stay until the next poll
    std::map<uint64_t, std::vector<double>> keep;
    for (int i = 0; i < k; i++) {
        auto it = S.solutions.find(out[i].ticket);
        if (it != S.solutions.end()) keep[out[i].ticket] = std::move(it->second);
    }
    for (int i = k; i < (int)S.done.size(); i++) {
        auto it = S.solutions.find(S.done[i].ticket);
        if (it != S.solutions.end()) keep[S.done[i].ticket] = std::move(it->second);
    }
    S.solutions.swap(keep);
    S.done.erase(S.done.begin(), S.done.begin() + k);
    *n = k;
    return AMIE_OK;
}

int amie_gpu_server_get_solution(amie_gpu_server *s, uint64_t ticket, double *x, int64_t n)
{
    if (!s || !x) { set_error("bad arguments"); return AMIE_ERR_ARG; }
    auto it = s->s->solutions.find(ticket);
    if (it == s->s->solutions.end()) { set_error("no solution for ticket %llu", (unsigned long long)ticket); return AMIE_ERR_ARG; }
    if ((int64_t)it->second.size() != n) { set_error("solution has %zu entries", it->second.size()); return AMIE_ERR_ARG; }
    std::memcpy(x, it->second.data(), n * 8);
    return AMIE_OK;
}

} // extern "C"
