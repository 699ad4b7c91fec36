// Serving: a request loop over resident systems with continuous batching.
//
// The reference solves one right-hand side per call, synchronously
// (AFIIM/example/example.cpp:100-118 and the PDFIIM/ACIFIM drivers).  The
// server keeps each registered system's A and factors resident on the device
// and admits solve requests as they arrive:
//  * slots: `slots` worker threads, each with its own HIP stream.  A free
//    worker takes the oldest pending request whose system is not already
//    being solved, and with it every other pending request of that system
//    (same tol / NMAX) up to 16 right-hand sides: one block solve (SpMM for
//    all columns, each column its own BiCGSTAB).  Requests that arrive while
//    others are on the GPU join the next batch of their system; batches of
//    different systems run at the same time on different streams.
//  * admission by HBM footprint: registered systems count their resident
//    bytes; a batch's block workspace (per system and block width, kept for
//    reuse) is reserved before the batch starts.  When it does not fit the
//    budget, idle workspaces of other systems are released (least recently
//    used first); when it still does not fit, the batch waits for running
//    batches to finish, and a request that can never fit is failed.
//  * precision: 64 = FP64 block BiCGSTAB; 32 / 16 = FP64 refinement around
//    the FP32 inner block solve (FP32 or FP16-with-row-scale factors).
//  * small systems (n <= small_n, when set): every pending request of any
//    small system goes into one launch of the persistent batched kernel, one
//    workgroup per right-hand side (FP64; under refinement precision the
//    items restart until their true residual is below tol).
// Results do not depend on which other requests share a batch beyond the
// summation order of the column reductions (block width): every column runs
// its own recurrence and exit tests.
//
// Times are host steady_clock: arrival = submit call, start = the batch's
// first device call, done = the solution copied back to the host.
#pragma once
#include <chrono>
#include <condition_variable>
#include <deque>
#include <mutex>
#include <thread>
#include <unordered_map>

namespace cm {

using Clock = std::chrono::steady_clock;

struct Request {
    long long id = 0;
    int sys = 0, k = 0, nmax = 0;
    double tol = 0.0;
    std::vector<double> B, X, resval;
    std::vector<int> info;
    int outer = 0, inner = 0, batch_cols = 0;
    Clock::time_point t_arrive, t_start, t_done;
    bool done = false;
    int rc = CM_OK;
    std::string err;
};

struct ServedSystem {
    System *sys = nullptr;  // registered, not owned
    bool busy = false;
    size_t resident = 0;    // matrices and solve vectors (FP64 and inner copies)
    size_t ws = 0;          // block workspaces, as of the last batch
    long long last_use = 0;
};

static size_t workspace_bytes(const System &S)
{
    size_t b = 0;
    for (const auto &q : S.blks)
        if (q) b += q->bytes();
    return b;
}

struct Server {
    int device = 0, precision = 32, max_cols = 16, max_outer = 30;
    int small_n = 0;         // systems with n <= small_n go to the persistent batched kernel (0: never)
    int max_items = 4096;    // right-hand sides per persistent launch
    int cus = 256;
    std::vector<std::unique_ptr<BatchRunner>> runners;  // one per worker
    std::vector<size_t> runner_bytes;
    // per-worker pinned staging for a batch's right-hand sides and solutions
    // (true asynchronous copies), grown on demand
    std::vector<double *> pin;
    std::vector<size_t> pin_cap;
    long long pbatches = 0, pcolumns = 0;
    double inner_tol = 1e-4;
    size_t budget = 0;
    size_t reserved = 0;  // workspace bytes reserved by running batches, not yet created
    std::vector<ServedSystem> systems;
    std::deque<Request *> pending;
    std::unordered_map<long long, std::unique_ptr<Request>> reqs;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::vector<std::thread> workers;
    std::vector<hipStream_t> streams;
    bool stop = false;
    long long next_id = 1, tick = 0;
    int running = 0, max_running = 0;
    long long batches = 0, columns = 0, evictions = 0;

    // Device bytes accounted: resident systems, their workspaces and the
    // reservations of running batches (caller holds mu).
    size_t accounted() const
    {
        size_t b = reserved;
        for (const auto &s : systems) b += s.resident + s.ws;
        for (size_t q : runner_bytes) b += q;
        return b;
    }

    // Bytes a batch of width kb on idle system si would add (0 when its
    // workspace exists).
    size_t workspace_need(int si, int kb)
    {
        System &S = *systems[si].sys;
        if (S.blks[block_index(kb)]) return 0;
        return block_bytes_estimate(*S.k64, kb, precision);
    }

    // Frees idle block workspaces of other systems, least recently used
    // first, until `need` more bytes fit.  Caller holds mu; no system being
    // freed is busy, so no stream of theirs has work in flight.
    void evict_for(size_t need, int keep)
    {
        while (accounted() + need > budget) {
            int best = -1, bi = -1;
            for (int s = 0; s < (int)systems.size(); s++) {
                if (s == keep || systems[s].busy) continue;
                for (int q = 0; q < 5; q++)
                    if (systems[s].sys->blks[q] && (best < 0 || systems[s].last_use < systems[best].last_use)) {
                        best = s;
                        bi = q;
                    }
            }
            if (best < 0) return;
            System &S = *systems[best].sys;
            if (S.blk == S.blks[bi].get()) S.blk = nullptr;
            S.blks[bi].reset();
            systems[best].ws = workspace_bytes(S);
            evictions++;
        }
    }

    // Device bytes worker w's batch runner would add for these requests.
    size_t persistent_need(int w, const std::vector<Request *> &reqs_) const
    {
        size_t cols = 0, vec = 0;
        int maxn = 0;
        for (const Request *r : reqs_) {
            const int n = systems[r->sys].sys->n;
            cols += r->k;
            vec += 2 * (size_t)n * r->k;
            maxn = std::max(maxn, n);
        }
        const size_t arena = std::min(cols, (size_t)cus) * ((9 * (size_t)maxn + 63) / 64 * 64);
        const size_t want = (arena + vec) * sizeof(double) + cols * sizeof(BatchItem) + systems.size() * sizeof(BatchSys);
        return want > runner_bytes[w] ? want - runner_bytes[w] : 0;
    }

    bool is_small(const Request *r) const { return small_n > 0 && systems[r->sys].sys->n <= small_n; }

    // Admission of `need` more bytes (caller holds mu): evicts idle
    // workspaces when that makes room.  1: admitted; 0: wait; -1: never fits.
    int admit(size_t need, int keep)
    {
        if (accounted() + need > budget) evict_for(need, keep);
        if (accounted() + need <= budget) return 1;
        return running > 0 ? 0 : -1;
    }

    void reject(size_t i)
    {
        Request *r = pending[i];
        r->rc = CM_ERR_ARG;
        r->err = "request does not fit the HBM budget";
        r->done = true;
        r->t_done = Clock::now();
        pending.erase(pending.begin() + i);
        cv_done.notify_all();
    }

    // Picks the next batch for worker w (caller holds mu).  Returns 1 for a
    // block batch of system si, 2 for a persistent batch of small systems,
    // 0 when nothing can start now.
    int pick(int w, std::vector<Request *> &batch, int &si, size_t &res)
    {
        for (size_t i = 0; i < pending.size(); i++) {
            Request *r = pending[i];
            if (is_small(r)) {
                // every pending small-system request, oldest first
                std::vector<size_t> take;
                std::vector<Request *> sel;
                size_t cols = 0;
                for (size_t q = i; q < pending.size(); q++) {
                    Request *c = pending[q];
                    if (!is_small(c) || (q != i && cols + c->k > (size_t)max_items)) continue;
                    take.push_back(q);
                    sel.push_back(c);
                    cols += c->k;
                }
                const size_t need = persistent_need(w, sel);
                const int a = admit(need, -1);
                if (a == 0) return 0;
                if (a < 0) {
                    reject(i);
                    i--;
                    continue;
                }
                reserved += need;
                res = need;
                for (auto it = take.rbegin(); it != take.rend(); ++it) pending.erase(pending.begin() + *it);
                batch = sel;
                si = -1;
                return 2;
            }
            if (systems[r->sys].busy) continue;
            const int s = r->sys;
            std::vector<size_t> take;
            int cols = 0;
            for (size_t q = i; q < pending.size(); q++) {
                Request *c = pending[q];
                if (c->sys != s || c->tol != r->tol || c->nmax != r->nmax) continue;
                if (q != i && cols + c->k > max_cols) continue;  // the oldest one always goes
                take.push_back(q);
                cols += c->k;
            }
            const int kb = block_width(cols);
            size_t need = workspace_need(s, kb);
            const int a = admit(need, s);
            if (a == 0) return 0;  // wait for memory to free up
            if (a < 0) {           // nothing running and still too big: never fits
                reject(i);
                i--;
                continue;
            }
            reserved += need;
            res = need;
            for (auto it = take.rbegin(); it != take.rend(); ++it) {
                batch.insert(batch.begin(), pending[*it]);
                pending.erase(pending.begin() + *it);
            }
            si = s;
            return 1;
        }
        return 0;
    }

    int pinned(int w, size_t count)
    {
        if (count <= pin_cap[w]) return CM_OK;
        if (pin[w]) (void)hipHostFree(pin[w]);
        pin[w] = nullptr;
        pin_cap[w] = 0;
        CM_HIP(hipHostMalloc((void **)&pin[w], count * sizeof(double), hipHostMallocDefault));
        pin_cap[w] = count;
        return CM_OK;
    }

    int run_batch(int w, int si, std::vector<Request *> &batch)
    {
        System &S = *systems[si].sys;
        const int n = S.n;
        int k = 0;
        for (Request *r : batch) k += r->k;
        std::vector<double> rv(k);
        std::vector<int> it(k, 0), info(k, 0);
        CM_TRY(pinned(w, (size_t)n * k));
        double *B = pin[w], *X = pin[w];  // the solution overwrites B once it is on the device
        size_t off = 0;
        for (Request *r : batch) {
            std::memcpy(B + off, r->B.data(), r->B.size() * sizeof(double));
            off += r->B.size();
        }
        const auto t0 = Clock::now();
        for (Request *r : batch) r->t_start = t0;
        CM_TRY(prepare_block_dia(S));
        CM_TRY(make_block(S.blks, S.blk, k, *S.k64));
        BlockBase &blk = *S.blk;
        blk.k32 = S.k32.get();
        blk.k16 = S.k16.get();
        blk.set_graphs(S.block_graphs);
        blk.set_stream(streams[w]);
        int outer = 0, inner = 0;
        int rc = blk.set(k, B, nullptr);
        const Request &r0 = *batch[0];
        if (rc == CM_OK) {
            if (precision == 64) {
                rc = blk.solve64(r0.tol, r0.nmax, rv.data(), it.data(), info.data());
                for (int j = 0; j < k; j++) inner = std::max(inner, it[j]);
            } else {
                rc = blk.solve_mixed(precision, r0.tol, inner_tol, r0.nmax, max_outer, rv.data(), &outer, &inner,
                                     info.data());
            }
        }
        if (rc == CM_OK) rc = blk.get(X);
        blk.set_stream(S.stream);
        if (rc != CM_OK) return rc;
        const auto t1 = Clock::now();
        off = 0;
        int c = 0;
        for (Request *r : batch) {
            r->X.assign(X + off, X + off + r->B.size());
            r->resval.assign(rv.begin() + c, rv.begin() + c + r->k);
            r->info.assign(info.begin() + c, info.begin() + c + r->k);
            r->outer = outer;
            r->inner = inner;
            r->batch_cols = k;
            r->t_done = t1;
            off += r->B.size();
            c += r->k;
        }
        return CM_OK;
    }

    // Small systems: every column of the batch is one item of the persistent
    // kernel (FP64; with refinement precision the items restart until their
    // true residual is below tol).
    int run_persistent(int w, std::vector<Request *> &batch)
    {
        std::vector<const System *> sl;
        std::vector<int> local(systems.size(), -1);
        std::vector<int> sys_of;
        std::vector<const double *> B;
        std::vector<double *> X;
        std::vector<double> tol;
        std::vector<int> nmax;
        for (Request *r : batch) {
            if (local[r->sys] < 0) {
                local[r->sys] = (int)sl.size();
                sl.push_back(systems[r->sys].sys);
            }
            const int n = systems[r->sys].sys->n;
            r->X.assign((size_t)n * r->k, 0.0);
            for (int j = 0; j < r->k; j++) {
                sys_of.push_back(local[r->sys]);
                B.push_back(r->B.data() + (size_t)j * n);
                X.push_back(r->X.data() + (size_t)j * n);
                tol.push_back(r->tol);
                nmax.push_back(r->nmax);
            }
        }
        const int m = (int)B.size();
        std::vector<double> rv(m);
        std::vector<int> it(m), info(m), cyc(m);
        const auto t0 = Clock::now();
        for (Request *r : batch) r->t_start = t0;
        CM_TRY(runners[w]->run(sl, m, sys_of.data(), B.data(), X.data(), tol.data(), nmax.data(),
                               precision != 64 ? 1 : 0, rv.data(), it.data(), info.data(), cyc.data(), streams[w]));
        const auto t1 = Clock::now();
        int c = 0;
        for (Request *r : batch) {
            r->resval.assign(rv.begin() + c, rv.begin() + c + r->k);
            r->info.assign(info.begin() + c, info.begin() + c + r->k);
            r->outer = *std::max_element(cyc.begin() + c, cyc.begin() + c + r->k);
            r->inner = *std::max_element(it.begin() + c, it.begin() + c + r->k);
            r->batch_cols = m;
            r->t_done = t1;
            c += r->k;
        }
        return CM_OK;
    }

    void worker(int w)
    {
        (void)hipSetDevice(device);
        for (;;) {
            std::vector<Request *> batch;
            int si = -1, kind = 0;
            size_t res = 0;
            {
                std::unique_lock<std::mutex> lk(mu);
                for (;;) {
                    if (stop) return;
                    if ((kind = pick(w, batch, si, res)) != 0) break;
                    cv_work.wait(lk);
                }
                if (kind == 1) {
                    systems[si].busy = true;
                    systems[si].last_use = ++tick;
                }
                running++;
                max_running = std::max(max_running, running);
            }
            const int rc = kind == 1 ? run_batch(w, si, batch) : run_persistent(w, batch);
            const std::string err = rc == CM_OK ? std::string() : std::string(cm_last_error());
            {
                std::lock_guard<std::mutex> lk(mu);
                // the reservation becomes the workspace's own size
                reserved -= std::min(reserved, res);
                if (kind == 1) {
                    systems[si].ws = workspace_bytes(*systems[si].sys);
                    systems[si].busy = false;
                    batches++;
                } else {
                    runner_bytes[w] = runners[w]->bytes();
                    pbatches++;
                    for (Request *r : batch) pcolumns += r->k;
                }
                running--;
                for (Request *r : batch) {
                    columns += r->k;
                    if (rc != CM_OK) {
                        r->rc = rc;
                        r->err = err;
                        r->t_done = Clock::now();
                    }
                    r->done = true;
                }
            }
            cv_done.notify_all();
            cv_work.notify_all();
        }
    }
};

}  // namespace cm

struct cm_server_s {
    cm::Server srv;
};

extern "C" {

int cm_server_create(cm_server *out, int slots, int precision, long long hbm_budget)
{
    using namespace cm;
    if (!out) return fail(CM_ERR_ARG, "null handle");
    *out = nullptr;
    if (cm_device_count() <= 0) return fail(CM_ERR_NODEV, "no GPU visible");
    if (slots < 1 || slots > 16) return fail(CM_ERR_ARG, "slots must be 1..16");
    if (precision != 64 && precision != 32 && precision != 16) return fail(CM_ERR_ARG, "precision must be 64, 32 or 16");
    auto h = std::make_unique<cm_server_s>();
    Server &S = h->srv;
    CM_HIP(hipGetDevice(&S.device));
    S.precision = precision;
    if (hbm_budget > 0) {
        S.budget = (size_t)hbm_budget;
    } else {
        size_t fr = 0, tot = 0;
        CM_HIP(hipMemGetInfo(&fr, &tot));
        S.budget = tot / 10 * 9;
    }
    CM_HIP(hipDeviceGetAttribute(&S.cus, hipDeviceAttributeMultiprocessorCount, S.device));
    for (int w = 0; w < slots; w++) {
        hipStream_t s;
        CM_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
        S.streams.push_back(s);
        S.runners.push_back(std::make_unique<BatchRunner>());
        S.runner_bytes.push_back(0);
        S.pin.push_back(nullptr);
        S.pin_cap.push_back(0);
    }
    for (int w = 0; w < slots; w++) S.workers.emplace_back([&S, w] { S.worker(w); });
    *out = h.release();
    return CM_OK;
}

int cm_server_configure(cm_server h, int max_cols, double inner_tol, int max_outer, int small_n)
{
    using namespace cm;
    if (!h) return fail(CM_ERR_ARG, "null handle");
    if (max_cols < 1 || max_cols > 16) return fail(CM_ERR_ARG, "max_cols must be 1..16");
    if (!(inner_tol > 0.0) || max_outer < 0 || small_n < 0) return fail(CM_ERR_ARG, "bad settings");
    Server &S = h->srv;
    std::lock_guard<std::mutex> lk(S.mu);
    S.max_cols = max_cols;
    S.inner_tol = inner_tol;
    S.max_outer = max_outer;
    S.small_n = small_n;
    return CM_OK;
}

int cm_server_destroy(cm_server h)
{
    using namespace cm;
    if (!h) return CM_OK;
    Server &S = h->srv;
    {
        std::lock_guard<std::mutex> lk(S.mu);
        S.stop = true;
        for (Request *r : S.pending) {
            r->rc = CM_ERR_ARG;
            r->err = "server stopped";
            r->done = true;
        }
        S.pending.clear();
    }
    S.cv_work.notify_all();
    S.cv_done.notify_all();
    for (auto &t : S.workers) t.join();
    for (auto s : S.streams) (void)hipStreamDestroy(s);
    for (double *p : S.pin)
        if (p) (void)hipHostFree(p);
    delete h;
    return CM_OK;
}

int cm_server_add_system(cm_server h, cm_system sys, int *sys_id)
{
    using namespace cm;
    if (!h || !sys || !sys_id) return fail(CM_ERR_ARG, "null argument");
    Server &S = h->srv;
    System &Y = sys->sys;
    if (Y.variant == 2 && Y.k64->mode != 'E') return fail(CM_ERR_UNSUPPORTED, "served systems use apply mode E");
    if (S.precision != 64) CM_TRY(cm_system_enable_mixed(sys, S.precision));
    CM_TRY(prepare_block_dia(Y));  // diagonal block products: index and values built before serving
    std::lock_guard<std::mutex> lk(S.mu);
    for (const auto &q : S.systems)
        if (q.sys == &Y) return fail(CM_ERR_ARG, "system already registered");
    ServedSystem ss;
    ss.sys = &Y;
    ss.ws = workspace_bytes(Y);
    ss.resident = Y.bytes() - ss.ws;
    if (S.accounted() + ss.resident + ss.ws > S.budget) return fail(CM_ERR_ARG, "system does not fit the HBM budget");
    *sys_id = (int)S.systems.size();
    S.systems.push_back(ss);
    return CM_OK;
}

int cm_server_submit(cm_server h, int sys_id, int k, const double *B, double tol, int NMAX, long long *req_id)
{
    using namespace cm;
    if (!h || !B || !req_id) return fail(CM_ERR_ARG, "null argument");
    Server &S = h->srv;
    auto r = std::make_unique<Request>();
    r->t_arrive = Clock::now();
    {
        std::lock_guard<std::mutex> lk(S.mu);
        if (sys_id < 0 || sys_id >= (int)S.systems.size()) return fail(CM_ERR_ARG, "unknown system id");
        if (k < 1 || k > 16) return fail(CM_ERR_ARG, "request must have 1..16 right-hand sides");
        if (S.stop) return fail(CM_ERR_ARG, "server stopped");
    }
    const int n = S.systems[sys_id].sys->n;
    r->sys = sys_id;
    r->k = k;
    r->tol = tol;
    r->nmax = NMAX;
    r->B.assign(B, B + (size_t)n * k);
    {
        std::lock_guard<std::mutex> lk(S.mu);
        r->id = S.next_id++;
        *req_id = r->id;
        S.pending.push_back(r.get());
        S.reqs.emplace(r->id, std::move(r));
    }
    S.cv_work.notify_one();
    return CM_OK;
}

// Blocks until request req_id is done, then copies its results out and
// forgets it.  X: n x k column-major; resval, info: k; counts: outer steps,
// inner passes, columns in its batch; times_ms: queue (arrival to start),
// service (start to done), latency (arrival to done).
int cm_server_wait(cm_server h, long long req_id, double *X, double *resval, int *info, int *counts,
                   double *times_ms)
{
    using namespace cm;
    if (!h) return fail(CM_ERR_ARG, "null handle");
    Server &S = h->srv;
    std::unique_ptr<Request> r;
    {
        std::unique_lock<std::mutex> lk(S.mu);
        auto itr = S.reqs.find(req_id);
        if (itr == S.reqs.end()) return fail(CM_ERR_ARG, "unknown request id");
        Request *q = itr->second.get();
        S.cv_done.wait(lk, [&] { return q->done; });
        r = std::move(itr->second);
        S.reqs.erase(itr);
    }
    if (r->rc != CM_OK) return fail(r->rc, r->err);
    if (X) std::memcpy(X, r->X.data(), r->X.size() * sizeof(double));
    if (resval) std::memcpy(resval, r->resval.data(), r->k * sizeof(double));
    if (info) std::memcpy(info, r->info.data(), r->k * sizeof(int));
    if (counts) {
        counts[0] = r->outer;
        counts[1] = r->inner;
        counts[2] = r->batch_cols;
    }
    if (times_ms) {
        auto ms = [](Clock::duration d) { return std::chrono::duration<double, std::milli>(d).count(); };
        times_ms[0] = ms(r->t_start - r->t_arrive);
        times_ms[1] = ms(r->t_done - r->t_start);
        times_ms[2] = ms(r->t_done - r->t_arrive);
    }
    return CM_OK;
}

// stats: block batches, columns solved, most batches running at once,
// workspace evictions, bytes accounted, budget, persistent batches, their
// columns.
int cm_server_stats(cm_server h, long long *stats)
{
    using namespace cm;
    if (!h || !stats) return fail(CM_ERR_ARG, "null argument");
    Server &S = h->srv;
    std::lock_guard<std::mutex> lk(S.mu);
    stats[0] = S.batches;
    stats[1] = S.columns;
    stats[2] = S.max_running;
    stats[3] = S.evictions;
    stats[4] = (long long)S.accounted();
    stats[5] = (long long)S.budget;
    stats[6] = S.pbatches;
    stats[7] = S.pcolumns;
    return CM_OK;
}

}  // extern "C"
