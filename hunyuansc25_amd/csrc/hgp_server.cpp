// hgp_server.cpp -- the serving loop: a FIFO of partition jobs admitted by
// free slot and device-memory budget, run concurrently on per-slot streams.
#include "hgp_host.hpp"
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

    size_t held(size_t i) const { return busy[i] ? claim[i] : ctx[i]->arena.capacity(); }

    size_t reserved() const
    {
        size_t r = 0;
        for (size_t i = 0; i < ctx.size(); i++) r += held(i);
        return r;
    }

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
            lk.unlock();
            // The arena is capped at the claim; a run that needs more goes back
            // to the head of the queue with twice the claim, or 5/4 of what the
            // failed request needed if that is more (up to the budget).
            bool again = false;
            size_t regrow = 0;
            int rc = hgp::guarded([&] {
                try {
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
        for (auto &c : s->ctx) hgp::destroy_context(std::move(c));
        delete s;
    });
}

}  // extern "C"
