// Threading checks of the serving loop's ticket book (hgp_tickets.hpp), on
// the host alone: one waiter per ticket, a second waiter refused without
// touching the job, and a drain that returns only after every wait in
// flight.  Prints "ok" and exits 0 on success.
#include "hgp_tickets.hpp"

#include <atomic>
#include <chrono>
#include <cstdio>
#include <thread>
#include <vector>

struct Job {
    int value = 0;
};

#define CHECK(c)                                                        \
    do {                                                                \
        if (!(c)) {                                                     \
            std::fprintf(stderr, "check failed at line %d: %s\n", __LINE__, #c); \
            return 1;                                                   \
        }                                                               \
    } while (0)

int main()
{
    using Book = hgp::TicketBook<Job>;
    std::mutex mu;
    Book book(mu);
    int64_t t0, t1;
    {
        std::lock_guard<std::mutex> lk(mu);
        auto j = std::make_unique<Job>();
        j->value = 7;
        t0 = book.add(std::move(j));
        t1 = book.add(std::make_unique<Job>());
    }
    CHECK(t0 == 0 && t1 == 1);

    // two waiters race for t0: exactly one gets the job, the other is refused
    std::atomic<int> got{0}, refused{0}, seen{0};
    auto waiter = [&] {
        int r = book.wait(t0, [&](Job &j) { seen = j.value; });
        if (r == Book::kOk) got++;
        else if (r == Book::kClaimed || r == Book::kUnknown) refused++;
    };
    std::thread a(waiter), b(waiter);
    // let both reach the book before the job finishes
    for (int i = 0; i < 200; i++) {
        {
            std::lock_guard<std::mutex> lk(mu);
            if (book.waiters() == 1 && refused.load() == 1) break;
        }
        std::this_thread::sleep_for(std::chrono::milliseconds(1));
    }
    {
        std::lock_guard<std::mutex> lk(mu);
        book.finish(t0);
    }
    a.join();
    b.join();
    CHECK(got == 1 && refused == 1 && seen == 7);
    CHECK(book.wait(t0, [](Job &) {}) == Book::kUnknown);   // removed after its wait

    // a drain started while a wait is in flight returns only after it
    std::atomic<bool> waited{false}, drained{false};
    std::thread w([&] {
        book.wait(t1, [](Job &) {});
        waited = true;
    });
    for (int i = 0; i < 200; i++) {
        {
            std::lock_guard<std::mutex> lk(mu);
            if (book.waiters() == 1) break;
        }
        std::this_thread::sleep_for(std::chrono::milliseconds(1));
    }
    std::thread d([&] {
        book.drain();
        drained = true;
    });
    std::this_thread::sleep_for(std::chrono::milliseconds(20));
    CHECK(!drained);
    {
        std::lock_guard<std::mutex> lk(mu);
        book.finish(t1);
    }
    w.join();
    d.join();
    CHECK(waited && drained);
    {
        std::lock_guard<std::mutex> lk(mu);
        CHECK(book.size() == 0 && book.waiters() == 0);
    }
    // many tickets, one waiter each, finished from another thread
    std::vector<int64_t> ts;
    {
        std::lock_guard<std::mutex> lk(mu);
        for (int i = 0; i < 64; i++) {
            auto j = std::make_unique<Job>();
            j->value = i;
            ts.push_back(book.add(std::move(j)));
        }
    }
    std::atomic<int> sum{0};
    std::vector<std::thread> ws;
    for (int i = 0; i < 64; i++)
        ws.emplace_back([&, i] { book.wait(ts[i], [&](Job &j) { sum += j.value; }); });
    std::thread fin([&] {
        for (int i = 63; i >= 0; i--) {
            std::lock_guard<std::mutex> lk(mu);
            book.finish(ts[i]);
        }
    });
    fin.join();
    for (auto &t : ws) t.join();
    book.drain();
    CHECK(sum == 64 * 63 / 2);
    std::printf("ok\n");
    return 0;
}
