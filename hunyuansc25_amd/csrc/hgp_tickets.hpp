// hgp_tickets.hpp -- the ticket book of the serving loop: queued and running
// jobs by ticket, each waited for by at most one caller, and a drain that
// lets the server be destroyed only after every wait in flight has returned.
// Plain C++ (no HIP), so its threading can be tested on the host alone.
#pragma once

#include <condition_variable>
#include <cstdint>
#include <functional>
#include <map>
#include <memory>
#include <mutex>

namespace hgp {

template <class Job>
class TicketBook {
  public:
    enum { kOk = 0, kUnknown = -1, kClaimed = -2 };

    // All members are guarded by `mu`, which the owner shares (its own state
    // lives under the same lock).
    explicit TicketBook(std::mutex &mu) : mu_(mu) {}

    // Adds a job; returns its ticket.  Caller holds the lock.
    int64_t add(std::unique_ptr<Job> job)
    {
        int64_t t = next_++;
        Entry &e = book_[t];
        e.job = std::move(job);
        return t;
    }

    // The job of ticket t, or nullptr.  Caller holds the lock.
    Job *get(int64_t t)
    {
        auto it = book_.find(t);
        return it == book_.end() ? nullptr : it->second.job.get();
    }

    // Marks ticket t finished and wakes its waiter.  Caller holds the lock.
    void finish(int64_t t)
    {
        auto it = book_.find(t);
        if (it == book_.end()) return;
        it->second.done = true;
        done_.notify_all();
    }

    // Blocks until ticket t is finished, hands its job to `take` (under the
    // lock) and removes it.  A ticket is waited for once: a second wait on
    // it, concurrent or later, returns kClaimed (later: kUnknown once removed)
    // and never touches the job.  Takes the lock itself.
    int wait(int64_t t, const std::function<void(Job &)> &take)
    {
        std::unique_lock<std::mutex> lk(mu_);
        auto it = book_.find(t);
        if (it == book_.end()) return kUnknown;
        if (it->second.claimed) return kClaimed;
        it->second.claimed = true;
        waiters_++;
        done_.wait(lk, [&] { return book_.at(t).done; });
        take(*book_.at(t).job);
        book_.erase(t);
        if (--waiters_ == 0) idle_.notify_all();
        return kOk;
    }

    // Blocks until no wait is in flight (called by the owner's teardown once
    // every job has finished, so every waiter is about to return).
    void drain()
    {
        std::unique_lock<std::mutex> lk(mu_);
        idle_.wait(lk, [&] { return waiters_ == 0; });
    }

    size_t size() const { return book_.size(); }   // caller holds the lock
    int waiters() const { return waiters_; }       // caller holds the lock

  private:
    struct Entry {
        std::unique_ptr<Job> job;
        bool claimed = false;
        bool done = false;
    };
    std::mutex &mu_;
    std::condition_variable done_, idle_;
    std::map<int64_t, Entry> book_;
    int64_t next_ = 0;
    int waiters_ = 0;
};

}  // namespace hgp
