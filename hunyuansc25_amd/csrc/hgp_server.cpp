// hgp_server.cpp -- the serving loop: a FIFO of partition jobs admitted by
// free slot and device-memory budget, run concurrently on per-slot streams.
// Graphs may be registered once and stay resident in HBM; partitions of a
// resident graph keep their coarsening hierarchies on the device (under the
// same budget, least recently used dropped first), so a later request for
// the same graph, seed and coarsening threshold starts at the initial
// partition.
#include "hgp_host.hpp"
#include "hgp_hostcheck.hpp"
#include "hgp_tickets.hpp"

#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <thread>

using Clock = std::chrono::steady_clock;

struct hgp_server {
    struct Job {
        hgp::HostJob j;
        int64_t graph = -1;     // resident graph (j holds k, part, stats only), or -1
        hgp_options o;
        hgp_stats stats;
        size_t bytes = 0;
        int status = 0;
        std::string msg;
        Clock::time_point t_submit, t_done;
    };

    int device = 0;
    hgp_options base;
    size_t budget = 0;
    std::vector<std::unique_ptr<hgp::Ctx>> ctx;
    std::vector<char> busy;
    std::deque<int64_t> queue;
    std::mutex mu;
    hgp::TicketBook<Job> jobs{mu};   // jobs by ticket, each waited for once
    std::condition_variable cv_work;
    bool stop = false;
    int64_t completed = 0;
    size_t peak = 0;
    std::vector<std::thread> workers;

    std::vector<size_t> claim;   // bytes a busy slot may hold (its arena never exceeds it)

    // resident graphs
    struct Resident {
        hgp::DevGraph g;         // CSR arrays in buf
        void *buf = nullptr;
        size_t bytes = 0;
        int64_t jobs = 0;        // queued or running partitions of it
        bool dropping = false;
    };
    std::map<int64_t, Resident> graphs;
    int64_t next_graph = 0;
    size_t graph_bytes = 0;

    // kept hierarchies of resident graphs
    struct Kept {
        hgp::CachedHierarchy h;
        int64_t refs = 0;
        uint64_t last_use = 0;
    };
    std::map<hgp::HierKey, Kept> kept;
    size_t kept_bytes = 0;       // kept plus reserved for snapshots in progress
    uint64_t use_clock = 0;
    int64_t hits = 0, misses = 0;

    size_t held(size_t i) const { return busy[i] ? claim[i] : ctx[i]->arena.capacity(); }

    size_t reserved() const
    {
        size_t r = graph_bytes + kept_bytes;
        for (size_t i = 0; i < ctx.size(); i++) r += held(i);
        return r;
    }

    // Drops unreferenced kept hierarchies, least recently used first, until
    // `need` more bytes fit the budget or none is left.  Called with mu held.
    void evict_for(size_t need)
    {
        while (reserved() + need > budget) {
            auto victim = kept.end();
            for (auto it = kept.begin(); it != kept.end(); ++it)
                if (it->second.refs == 0 && (victim == kept.end() || it->second.last_use < victim->second.last_use))
                    victim = it;
            if (victim == kept.end()) return;
            kept_bytes -= victim->second.h.bytes;
            hgp::free_hierarchy(victim->second.h);
            kept.erase(victim);
        }
    }

    // The HierarchyStore of the slots' runs: the kept map under mu.
    struct Store : hgp::HierarchyStore {
        hgp_server *s;
        explicit Store(hgp_server *srv) : s(srv) {}
        const hgp::CachedHierarchy *acquire(const hgp::HierKey &key) override
        {
            std::lock_guard<std::mutex> lk(s->mu);
            auto it = s->kept.find(key);
            if (it == s->kept.end()) {
                s->misses++;
                return nullptr;
            }
            s->hits++;
            it->second.refs++;
            it->second.last_use = ++s->use_clock;
            return &it->second.h;
        }
        void release(const hgp::CachedHierarchy *h) override
        {
            std::lock_guard<std::mutex> lk(s->mu);
            for (auto &kv : s->kept)
                if (&kv.second.h == h) kv.second.refs--;
        }
        bool reserve(size_t bytes) override
        {
            std::lock_guard<std::mutex> lk(s->mu);
            s->evict_for(bytes);
            if (s->reserved() + bytes > s->budget) return false;
            s->kept_bytes += bytes;
            return true;
        }
        void unreserve(size_t bytes) override
        {
            std::lock_guard<std::mutex> lk(s->mu);
            s->kept_bytes -= bytes;
        }
        const hgp::CachedHierarchy *insert(const hgp::HierKey &key, hgp::CachedHierarchy &&h) override
        {
            std::lock_guard<std::mutex> lk(s->mu);
            auto it = s->kept.find(key);
            if (it != s->kept.end()) {   // another slot kept the same hierarchy first
                s->kept_bytes -= h.bytes;
                hgp::free_hierarchy(h);
            } else {
                it = s->kept.emplace(key, Kept{}).first;
                it->second.h = std::move(h);
            }
            it->second.refs++;
            it->second.last_use = ++s->use_clock;
            return &it->second.h;
        }
    } store{this};

    // Whether slot w may start a job of `need` bytes; may release the arenas
    // of idle slots (and its own) to make room.  Called with mu held.
    bool admit(int w, size_t need)
    {
        auto others = [&] {
            size_t r = 0;
            for (size_t i = 0; i < ctx.size(); i++)
                if ((int)i != w) r += held(i);
            return r;
        };
        size_t own = ctx[w]->arena.capacity();
        if (others() + std::max(own, need) <= budget) return true;
        for (size_t i = 0; i < ctx.size(); i++)
            if ((int)i != w && !busy[i] && ctx[i]->arena.capacity() > 0) ctx[i]->arena.free_block();
        if (others() + std::max(own, need) <= budget) return true;
        evict_for(std::max(own, need) > held((size_t)w) ? std::max(own, need) - held((size_t)w) : 0);
        if (others() + std::max(own, need) <= budget) return true;
        if (own > need && others() + need <= budget) {
            ctx[w]->arena.free_block();
            return true;
        }
        return false;
    }

    void worker(int w)
    {
        std::unique_lock<std::mutex> lk(mu);
        for (;;) {
            cv_work.wait(lk, [&] {
                return (stop && queue.empty()) || (!queue.empty() && admit(w, jobs.get(queue.front())->bytes));
            });
            if (queue.empty()) break;   // stopping and drained
            int64_t t = queue.front();
            queue.pop_front();
            Job *job = jobs.get(t);
            size_t need = job->bytes;
            busy[w] = 1;
            claim[w] = std::max(ctx[w]->arena.capacity(), need);
            peak = std::max(peak, reserved());
            // a resident graph stays until its jobs are done (unregister waits)
            const hgp::DevGraph *rg = job->graph >= 0 ? &graphs.at(job->graph).g : nullptr;
            lk.unlock();
            // The arena is capped at the claim; a run that needs more goes back
            // to the head of the queue with twice the claim, or 5/4 of what the
            // failed request needed if that is more (up to the budget).
            bool again = false;
            size_t regrow = 0;
            int rc = hgp::guarded([&] {
                try {
                    if (rg)
                        hgp::exec_resident(*ctx[w], job->o, *rg, job->j.k, job->j.part, job->j.stats, &store,
                                           job->graph, need, claim[w]);
                    else
                        hgp::exec_job(*ctx[w], job->o, job->j, need, claim[w]);
                } catch (const hgp::Error &e) {
                    if (e.overflow && need < budget && e.need <= budget) {
                        again = true;
                        regrow = std::max(2 * need, e.need / 4 * 5 + ((size_t)64 << 20));
                        return;
                    }
                    throw;
                }
            });
            std::string msg = rc ? hgp::last_error_text() : "";
            lk.lock();
            busy[w] = 0;
            if (again) {
                job->bytes = std::min(regrow, budget);
                queue.push_front(t);
            } else {
                job->status = rc;
                job->msg = msg;
                job->t_done = Clock::now();
                completed++;
                if (job->graph >= 0) graphs.at(job->graph).jobs--;
                jobs.finish(t);
            }
            cv_work.notify_all();
        }
    }
};

extern "C" {

int hgp_server_create(const hgp_options *opt, int32_t slots, int64_t budget_bytes, hgp_server **out)
{
    return hgp::guarded([&] {
        if (!out) throw hgp::Error{-1, "out is NULL"};
        *out = nullptr;
        if (slots < 1 || slots > 64) throw hgp::Error{-1, "slots must be 1..64"};
        if (budget_bytes <= 0) throw hgp::Error{-1, "budget must be positive"};
        auto s = std::make_unique<hgp_server>();
        if (opt) s->base = *opt; else hgp_default_options(&s->base);
        s->device = s->base.device;
        s->budget = (size_t)budget_bytes;
        for (int i = 0; i < slots; i++) s->ctx.push_back(hgp::create_context(s->device));
        s->busy.assign((size_t)slots, 0);
        s->claim.assign((size_t)slots, 0);
        for (int i = 0; i < slots; i++) s->workers.emplace_back([p = s.get(), i] {
            (void)hipSetDevice(p->device);
            p->worker(i);
        });
        *out = s.release();
    });
}

int64_t hgp_server_submit(hgp_server *s, int32_t n, const int64_t *xadj, const int32_t *adjncy, const int32_t *vwgt,
                          const int32_t *adjwgt, int32_t k, uint64_t seed, int32_t *part)
{
    int64_t ticket = -1;
    int rc = hgp::guarded([&] {
        if (!s) throw hgp::Error{-1, "server is NULL"};
        auto job = std::make_unique<hgp_server::Job>();
        job->o = s->base;
        job->o.seed = seed;
        job->j = hgp::HostJob{n, xadj, adjncy, vwgt, adjwgt, k, part, &job->stats};
        hgp::check_job(job->o, job->j);
        job->bytes = hgp::job_bytes(job->j);
        if (n > 0 && job->bytes > s->budget) throw hgp::Error{-3, "job needs more device memory than the budget"};
        job->t_submit = Clock::now();
        std::lock_guard<std::mutex> lk(s->mu);
        if (s->stop) throw hgp::Error{-1, "server is stopping"};
        ticket = s->jobs.add(std::move(job));
        s->queue.push_back(ticket);
        s->cv_work.notify_all();
    });
    return rc == HGP_OK ? ticket : rc;
}

int64_t hgp_server_register(hgp_server *s, int32_t n, const int64_t *xadj, const int32_t *adjncy,
                            const int32_t *vwgt, const int32_t *adjwgt)
{
    int64_t id = -1;
    int rc = hgp::guarded([&] {
        if (!s) throw hgp::Error{-1, "server is NULL"};
        hgp::validate_host(n, xadj, adjncy, vwgt, adjwgt, s->base.validate != 0);
        if (n == 0) throw hgp::Error{-1, "the graph has no vertices"};
        const int64_t m = xadj[n];
        const size_t bytes = 8 * ((size_t)n + 1) + 8 * (size_t)m + 4 * (size_t)n;
        {
            std::lock_guard<std::mutex> lk(s->mu);
            if (s->stop) throw hgp::Error{-1, "server is stopping"};
            for (size_t i = 0; i < s->ctx.size(); i++)
                if (!s->busy[i] && s->reserved() + bytes > s->budget && s->ctx[i]->arena.capacity() > 0)
                    s->ctx[i]->arena.free_block();
            s->evict_for(bytes);
            if (s->reserved() + bytes > s->budget) throw hgp::Error{-3, "the graph does not fit the HBM budget"};
            s->graph_bytes += bytes;   // held while uploading
        }
        void *buf = nullptr;
        auto undo = [&] {
            if (buf) (void)hipFree(buf);
            std::lock_guard<std::mutex> lk(s->mu);
            s->graph_bytes -= bytes;
        };
        try {
            HGP_CHECK_HIP(hipSetDevice(s->device));
            if (hipMalloc(&buf, bytes) != hipSuccess) {
                (void)hipGetLastError();
                buf = nullptr;
                throw hgp::Error{-3, "device allocation of the resident graph failed"};
            }
            char *b = (char *)buf;
            hgp::DevGraph g;
            g.n = n;
            g.m = m;
            g.xadj = (int64_t *)b;
            g.adjncy = (int32_t *)(b + 8 * ((size_t)n + 1));
            g.adjwgt = (int32_t *)(b + 8 * ((size_t)n + 1) + 4 * (size_t)m);
            g.vwgt = (int32_t *)(b + 8 * ((size_t)n + 1) + 8 * (size_t)m);
            HGP_CHECK_HIP(hipMemcpy(g.xadj, xadj, 8 * ((size_t)n + 1), hipMemcpyHostToDevice));
            HGP_CHECK_HIP(hipMemcpy(g.adjncy, adjncy, 4 * (size_t)m, hipMemcpyHostToDevice));
            if (adjwgt) {
                HGP_CHECK_HIP(hipMemcpy(g.adjwgt, adjwgt, 4 * (size_t)m, hipMemcpyHostToDevice));
            } else {
                std::vector<int32_t> ones((size_t)m, 1);
                HGP_CHECK_HIP(hipMemcpy(g.adjwgt, ones.data(), 4 * (size_t)m, hipMemcpyHostToDevice));
            }
            if (vwgt) {
                HGP_CHECK_HIP(hipMemcpy(g.vwgt, vwgt, 4 * (size_t)n, hipMemcpyHostToDevice));
            } else {
                std::vector<int32_t> ones((size_t)n, 1);
                HGP_CHECK_HIP(hipMemcpy(g.vwgt, ones.data(), 4 * (size_t)n, hipMemcpyHostToDevice));
            }
            std::lock_guard<std::mutex> lk(s->mu);
            id = s->next_graph++;
            hgp_server::Resident &r = s->graphs[id];
            r.g = g;
            r.buf = buf;
            r.bytes = bytes;
        } catch (...) {
            undo();
            throw;
        }
    });
    return rc == HGP_OK ? id : rc;
}

int64_t hgp_server_submit_graph(hgp_server *s, int64_t graph, int32_t k, uint64_t seed, int32_t *part)
{
    int64_t ticket = -1;
    int rc = hgp::guarded([&] {
        if (!s) throw hgp::Error{-1, "server is NULL"};
        if (!part) throw hgp::Error{-1, "part is NULL"};
        auto job = std::make_unique<hgp_server::Job>();
        job->o = s->base;
        job->o.seed = seed;
        job->graph = graph;
        hgp::check_options(job->o, k);
        std::lock_guard<std::mutex> lk(s->mu);
        if (s->stop) throw hgp::Error{-1, "server is stopping"};
        auto it = s->graphs.find(graph);
        if (it == s->graphs.end() || it->second.dropping) throw hgp::Error{-1, "unknown graph"};
        hgp_server::Resident &r = it->second;
        job->j = hgp::HostJob{r.g.n, nullptr, nullptr, nullptr, nullptr, k, part, &job->stats};
        job->bytes = hgp::resident_job_bytes(r.g.n, r.g.m, k);
        if (job->bytes + r.bytes > s->budget) throw hgp::Error{-3, "job needs more device memory than the budget"};
        job->t_submit = Clock::now();
        r.jobs++;
        ticket = s->jobs.add(std::move(job));
        s->queue.push_back(ticket);
        s->cv_work.notify_all();
    });
    return rc == HGP_OK ? ticket : rc;
}

int hgp_server_unregister(hgp_server *s, int64_t graph)
{
    return hgp::guarded([&] {
        if (!s) throw hgp::Error{-1, "server is NULL"};
        std::unique_lock<std::mutex> lk(s->mu);
        auto it = s->graphs.find(graph);
        if (it == s->graphs.end() || it->second.dropping) throw hgp::Error{-1, "unknown graph"};
        it->second.dropping = true;
        // the graph's queued and running partitions finish first
        s->cv_work.wait(lk, [&] { return it->second.jobs == 0; });
        for (auto k = s->kept.begin(); k != s->kept.end();) {
            if (k->first.graph == graph) {
                s->kept_bytes -= k->second.h.bytes;
                hgp::free_hierarchy(k->second.h);
                k = s->kept.erase(k);
            } else {
                ++k;
            }
        }
        (void)hipSetDevice(s->device);
        (void)hipFree(it->second.buf);
        s->graph_bytes -= it->second.bytes;
        s->graphs.erase(it);
    });
}

int hgp_server_kept(hgp_server *s, int64_t *hits, int64_t *misses, int64_t *kept, int64_t *kept_bytes)
{
    return hgp::guarded([&] {
        if (!s) throw hgp::Error{-1, "server is NULL"};
        std::lock_guard<std::mutex> lk(s->mu);
        if (hits) *hits = s->hits;
        if (misses) *misses = s->misses;
        if (kept) *kept = (int64_t)s->kept.size();
        if (kept_bytes) *kept_bytes = (int64_t)s->kept_bytes;
    });
}

int hgp_server_wait(hgp_server *s, int64_t ticket, hgp_stats *stats, double *latency_ms)
{
    int status = HGP_OK;
    std::string msg;
    int rc = hgp::guarded([&] {
        if (!s) throw hgp::Error{-1, "server is NULL"};
        int w = s->jobs.wait(ticket, [&](hgp_server::Job &job) {
            if (stats) *stats = job.stats;
            if (latency_ms) *latency_ms = std::chrono::duration<double, std::milli>(job.t_done - job.t_submit).count();
            status = job.status;
            msg = job.msg;
        });
        if (w == hgp::TicketBook<hgp_server::Job>::kUnknown) throw hgp::Error{-1, "unknown ticket"};
        if (w == hgp::TicketBook<hgp_server::Job>::kClaimed)
            throw hgp::Error{-1, "ticket is already being waited for"};
    });
    if (rc != HGP_OK) return rc;
    if (status != HGP_OK) hgp::set_error(status, msg);
    return status;
}

int hgp_server_info(hgp_server *s, int64_t *reserved, int64_t *peak_reserved, int64_t *completed)
{
    return hgp::guarded([&] {
        if (!s) throw hgp::Error{-1, "server is NULL"};
        std::lock_guard<std::mutex> lk(s->mu);
        if (reserved) *reserved = (int64_t)s->reserved();
        if (peak_reserved) *peak_reserved = (int64_t)std::max(s->peak, s->reserved());
        if (completed) *completed = s->completed;
    });
}

int hgp_server_destroy(hgp_server *s)
{
    return hgp::guarded([&] {
        if (!s) return;
        {
            std::lock_guard<std::mutex> lk(s->mu);
            s->stop = true;
            s->cv_work.notify_all();
        }
        for (auto &t : s->workers) t.join();
        // every job has finished: the waits in flight return; the server
        // outlives them
        s->jobs.drain();
        (void)hipSetDevice(s->device);
        for (auto &kv : s->kept) hgp::free_hierarchy(kv.second.h);
        for (auto &kv : s->graphs) (void)hipFree(kv.second.buf);
        for (auto &c : s->ctx) hgp::destroy_context(std::move(c));
        delete s;
    });
}

}  // extern "C"
