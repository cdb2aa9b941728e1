// serve.cpp -- the persistent request loop of many Poisson solves and short
// Navier-Stokes runs on one GPU.
//
// Requests arrive at any time (saras_server_submit) and wait in a queue.
// `slots` worker threads each own the solves they run; every solver handle
// has its own HIP stream, so solves admitted together overlap on the device
// (continuous batching: a request is admitted as soon as a slot and enough
// device memory are free, while others are still running).  Admission is by
// device-memory footprint: hierarchies stay resident after a solve and are
// reused by later requests of the same grid and parameters; idle ones are
// evicted least-recently-used when a new request needs their memory.  A
// request whose footprint does not fit waits, letting later requests that
// fit pass it, and is run alone if it can only ever fit on an idle device.
//
// The solve of every request is the plain saras_mg_solve path on its own
// handle, so a result does not depend on what else is in flight.  A short
// Navier-Stokes request runs saras_nse_step on a handle of its own the same
// way, starting from a zero run time and pressure correction whether the
// handle is new or resident.
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <list>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "saras_hip.h"

namespace {

using Clock = std::chrono::steady_clock;

template <typename V>
void put(std::string &k, const V &v) {
    k.append(reinterpret_cast<const char *>(&v), sizeof(v));
}

// field-wise key of (grid, parameters): struct padding never enters it
std::string key_of(const saras_grid_desc &g, const saras_mg_params &p) {
    std::string k;
    put(k, g.dim);
    for (int a = 0; a < 3; ++a) {
        put(k, g.n[a]);
        put(k, g.len[a]);
        put(k, g.periodic[a]);
        put(k, g.gtype[a]);
        put(k, g.beta[a]);
    }
    put(k, g.pad);
    put(k, p.vc_count);
    put(k, p.vc_tol);
    put(k, p.solve_flag);
    put(k, p.mg_tol);
    put(k, p.sor);
    put(k, p.pre_smooth);
    put(k, p.post_smooth);
    put(k, p.res_type);
    put(k, p.bc);
    put(k, p.precision);
    put(k, p.inner_cycles);
    put(k, p.max_refine);
    put(k, p.top_dirichlet);
    return k;
}

std::string key_of(const saras_grid_desc &g, const saras_mg_params &m, const saras_nse_params &q) {
    std::string k = "nse";
    k += key_of(g, m);
    put(k, q.dscheme);
    put(k, q.nlscheme);
    put(k, q.has_T);
    put(k, q.nu);
    put(k, q.kappa);
    put(k, q.dt);
    put(k, q.cn_tol);
    for (int a = 0; a < 3; ++a) put(k, q.buoy[a]);
    for (int w = 0; w < 6; ++w) {
        for (int c = 0; c < 3; ++c) {
            put(k, q.vbc_type[c][w]);
            put(k, q.vbc_val[c][w]);
        }
        put(k, q.tbc_type[w]);
        put(k, q.tbc_val[w]);
        put(k, q.pbc_type[w]);
        put(k, q.pbc_val[w]);
    }
    put(k, q.scheme);
    put(k, q.pe_limit);
    put(k, q.central_bias);
    put(k, q.les);
    put(k, q.mean_pgrad);
    put(k, q.plate_radius);
    put(k, q.les_sf32);
    return k;
}

enum Kind { kPoisson = 0, kNse = 1 };

// A handle of either kind.
struct Handle {
    Kind kind = kPoisson;
    void *h = nullptr;
    void destroy() {
        if (!h) return;
        if (kind == kPoisson) saras_mg_destroy(static_cast<saras_mg *>(h));
        else saras_nse_destroy(static_cast<saras_nse *>(h));
        h = nullptr;
    }
};

// saras_nse_request: reset the run, `steps` time steps, fields back to the caller
saras_status run_nse(saras_nse *h, const saras_nse_request &q, saras_completion &c) {
    const saras_grid_desc &g = q.grid;
    const long p = g.pad, yp = g.dim == 2 ? 0 : p, ny = g.dim == 2 ? 1 : g.n[1];
    std::vector<double> zero((size_t)(g.n[0] + 2 * p) * (ny + 2 * yp) * (g.n[2] + 2 * p), 0.0);
    const double *t = q.nse.has_T ? q.t : nullptr;
    saras_status st = saras_nse_set_fields(h, q.vx, q.vy, q.vz, q.p, t, zero.data(), SARAS_PTR_HOST);
    if (st == SARAS_OK) st = saras_nse_set_time(h, 0.0);
    double ms = 0.0;
    for (int k = 0; k < q.steps && st == SARAS_OK; ++k) {
        saras_nse_stats ns{};
        st = saras_nse_step(h, &ns);
        c.stats = ns.mg;
        ms += ns.device_ms;
    }
    if (st == SARAS_OK) st = saras_nse_get_fields(h, q.vx, q.vy, q.vz, q.p, q.nse.has_T ? q.t : nullptr, nullptr,
                                                  SARAS_PTR_HOST);
    c.stats.device_ms = ms;
    c.stats.converged = st == SARAS_OK;
    return st;
}

// a request blocked on memory stops others from passing it after this long
constexpr double kMaxBypassSeconds = 0.25;

struct Pending {
    Kind kind = kPoisson;
    saras_request r{};
    saras_nse_request q{};
    double t_submit = 0.0;
    size_t bytes = 0;
    std::string key;
    long long tag() const { return kind == kPoisson ? r.tag : q.tag; }
};

struct Idle {
    std::string key;
    Handle h;
    size_t bytes;
};

}  // namespace

struct saras_server {
    saras_server_params p{};
    Clock::time_point t0;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<Pending> queue;
    std::list<Idle> idle;                  // front = most recently used
    std::vector<Handle> evicted;           // taken off `idle`, destroyed by a worker outside the lock
    std::vector<saras_completion> done;
    std::vector<std::thread> workers;
    size_t mem_used = 0;                   // idle handles + running solves
    int running = 0;
    long long submitted = 0, finished = 0;
    bool stop = false;

    double now() const { return std::chrono::duration<double>(Clock::now() - t0).count(); }

    // Pick the next admissible request (caller holds mu).  Returns false when
    // none can run now.  On success the request's memory is accounted and, if
    // a resident handle fits it, that handle is taken from the idle list.
    bool take(Pending &out, Handle &h) {
        if (p.largest_first && queue.size() > 1) {
            // visit the waiting requests largest footprint first (stable for
            // equal footprints, so arrival order breaks ties)
            std::stable_sort(queue.begin(), queue.end(),
                             [](const Pending &a, const Pending &b) { return a.bytes > b.bytes; });
        }
        for (auto it = queue.begin(); it != queue.end(); ++it) {
            // a resident hierarchy of the same grid and parameters
            for (auto jt = idle.begin(); jt != idle.end(); ++jt)
                if (jt->key == it->key) {
                    h = jt->h;
                    it->bytes = jt->bytes;   // already accounted
                    idle.erase(jt);
                    out = *it;
                    queue.erase(it);
                    return true;
                }
            const size_t need = it->bytes;
            size_t idle_bytes = 0;
            for (const Idle &v : idle) idle_bytes += v.bytes;
            // Admit when it fits now, when evicting idle hierarchies makes it
            // fit, or when nothing runs (a request larger than the budget then
            // runs alone).  Eviction happens only for an admission, so memory
            // held by running solves never empties the cache for nothing.
            if (mem_used + need <= p.mem_budget || mem_used - idle_bytes + need <= p.mem_budget || running == 0) {
                // evict least recently used first, until it fits
                while (mem_used + need > p.mem_budget && !idle.empty()) {
                    evicted.push_back(idle.back().h);
                    mem_used -= idle.back().bytes;
                    idle.pop_back();
                }
                mem_used += need;
                h = Handle{it->kind, nullptr};
                out = *it;
                queue.erase(it);
                return true;
            }
            // this one waits for memory; later requests may pass it unless it
            // has already waited too long
            if (now() - it->t_submit > kMaxBypassSeconds) return false;
        }
        return false;
    }

    void worker() {
        for (;;) {
            Pending job;
            Handle h;
            std::vector<Handle> victims;
            {
                std::unique_lock<std::mutex> lk(mu);
                bool got = false;
                cv_work.wait(lk, [&] {
                    if (stop) return true;
                    got = !queue.empty() && take(job, h);
                    return got;
                });
                victims.swap(evicted);
                if (!got) {
                    lk.unlock();
                    for (Handle &v : victims) v.destroy();
                    return;
                }
                ++running;
            }
            // hipFree synchronises the device: never under the server lock
            for (Handle &v : victims) v.destroy();
            saras_completion c{};
            c.tag = job.tag();
            c.t_submit = job.t_submit;
            c.t_admit = now();
            saras_status st = SARAS_OK;
            if (job.kind == kPoisson) {
                saras_mg *mg = static_cast<saras_mg *>(h.h);
                if (!mg) {
                    st = saras_mg_create(&job.r.grid, &job.r.mg, &mg);
                    if (st == SARAS_OK) h.h = mg;
                    if (st == SARAS_OK && p.use_graphs) st = saras_mg_set_graphs(mg, 1);
                }
                if (st == SARAS_OK) st = saras_mg_solve(mg, job.r.x, job.r.b, SARAS_PTR_HOST, nullptr, &c.stats);
            } else {
                saras_nse *ns = static_cast<saras_nse *>(h.h);
                if (!ns) {
                    st = saras_nse_create(&job.q.grid, &job.q.mg, &job.q.nse, &ns);
                    if (st == SARAS_OK) h.h = ns;
                    if (st == SARAS_OK && p.use_graphs) {
                        // the time step's pressure solves replay their V-cycles too
                        st = saras_nse_set_graphs(ns, 1);
                    }
                }
                if (st == SARAS_OK) st = run_nse(ns, job.q, c);
            }
            c.status = st;
            c.t_done = now();
            if (st != SARAS_OK && st != SARAS_NOT_CONVERGED)
                std::strncpy(c.message, saras_last_error(), sizeof(c.message) - 1);
            Handle drop;
            {
                std::lock_guard<std::mutex> lk(mu);
                --running;
                if (h.h && p.keep_resident && (st == SARAS_OK || st == SARAS_NOT_CONVERGED)) {
                    idle.push_front({job.key, h, job.bytes});
                } else {
                    drop = h;
                    mem_used -= job.bytes;
                }
                done.push_back(c);
                ++finished;
            }
            drop.destroy();
            cv_done.notify_all();
            cv_work.notify_all();
        }
    }
};

extern "C" {

saras_status saras_server_create(const saras_server_params *params, saras_server **out) {
    if (!params || !out || params->slots < 1 || params->mem_budget == 0) return SARAS_BAD_ARG;
    saras_server *s = new saras_server();
    s->p = *params;
    s->t0 = Clock::now();
    for (int w = 0; w < params->slots; ++w) s->workers.emplace_back([s] { s->worker(); });
    *out = s;
    return SARAS_OK;
}

saras_status saras_server_submit(saras_server *s, const saras_request *req) {
    if (!s || !req || !req->b || !req->x) return SARAS_BAD_ARG;
    const size_t bytes = saras_mg_footprint(&req->grid, &req->mg);
    if (bytes == 0) return SARAS_BAD_SHAPE;
    Pending pd;
    pd.kind = kPoisson;
    pd.r = *req;
    pd.t_submit = s->now();
    pd.bytes = bytes;
    pd.key = key_of(req->grid, req->mg);
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->queue.push_back(pd);
        ++s->submitted;
    }
    s->cv_work.notify_one();
    return SARAS_OK;
}

saras_status saras_server_submit_nse(saras_server *s, const saras_nse_request *req) {
    if (!s || !req || !req->vx || !req->vy || !req->vz || !req->p || (req->nse.has_T && !req->t) || req->steps < 0)
        return SARAS_BAD_ARG;
    const size_t bytes = saras_nse_footprint(&req->grid, &req->mg, &req->nse);
    if (bytes == 0) return SARAS_BAD_SHAPE;
    Pending pd;
    pd.kind = kNse;
    pd.q = *req;
    pd.t_submit = s->now();
    pd.bytes = bytes;
    pd.key = key_of(req->grid, req->mg, req->nse);
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->queue.push_back(pd);
        ++s->submitted;
    }
    s->cv_work.notify_one();
    return SARAS_OK;
}

int saras_server_poll(saras_server *s, saras_completion *out, int max, int wait_ms) {
    if (!s || (!out && max > 0) || max < 0) return -1;
    std::unique_lock<std::mutex> lk(s->mu);
    if (s->done.empty() && wait_ms > 0)
        s->cv_done.wait_for(lk, std::chrono::milliseconds(wait_ms), [&] { return !s->done.empty(); });
    const int n = std::min<int>(max, (int)s->done.size());
    for (int q = 0; q < n; ++q) out[q] = s->done[q];
    s->done.erase(s->done.begin(), s->done.begin() + n);
    return n;
}

saras_status saras_server_drain(saras_server *s) {
    if (!s) return SARAS_BAD_ARG;
    std::unique_lock<std::mutex> lk(s->mu);
    s->cv_done.wait(lk, [&] { return s->finished == s->submitted; });
    return SARAS_OK;
}

double saras_server_clock(const saras_server *s) { return s ? s->now() : 0.0; }

size_t saras_server_mem_used(const saras_server *s) {
    if (!s) return 0;
    std::lock_guard<std::mutex> lk(const_cast<saras_server *>(s)->mu);
    return s->mem_used;
}

saras_status saras_server_destroy(saras_server *s) {
    if (!s) return SARAS_BAD_ARG;
    (void)saras_server_drain(s);
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
    }
    s->cv_work.notify_all();
    for (auto &t : s->workers) t.join();
    for (auto &v : s->idle) v.h.destroy();
    for (Handle &v : s->evicted) v.destroy();
    delete s;
    return SARAS_OK;
}

}  // extern "C"
