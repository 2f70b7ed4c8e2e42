// hgp_runtime.cpp -- error state, device contexts and the two-ended arena.
#include "hgp_common.hpp"

#include <algorithm>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>

namespace hgp {

namespace {
thread_local int g_err_code = 0;
thread_local std::string g_err_msg;
std::mutex g_ctx_mu;
std::map<int, std::unique_ptr<Ctx>> g_ctx;
}  // namespace

void set_error(int code, const std::string &msg)
{
    g_err_code = code;
    g_err_msg = msg;
}

void clear_error()
{
    g_err_code = 0;
    g_err_msg.clear();
}

const char *last_error_text() { return g_err_msg.c_str(); }

// ------------------------------------------------------------------ arena

static size_t round_up(size_t b) { return (b + 255) & ~(size_t)255; }

Arena::~Arena()
{
    if (base_) (void)hipFree(base_);
}

void Arena::reserve(size_t bytes)
{
    bytes = round_up(bytes);
    if (bytes <= cap_) return;
    if (lo_ != 0 || hi_ != cap_) throw Error{-3, "arena: reserve while allocations are live"};
    if (base_) HGP_CHECK_HIP(hipFree(base_));
    base_ = nullptr;
    cap_ = lo_ = hi_ = 0;
    hipError_t e = hipMalloc((void **)&base_, bytes);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        base_ = nullptr;
        throw Error{-3, "arena: hipMalloc of " + std::to_string(bytes) + " bytes failed"};
    }
    cap_ = bytes;
    hi_ = cap_;
}

// An allocation that does not fit throws Error{-3}; the caller resets the
// arena, reserves more and retries the whole call.
void *Arena::left(size_t bytes)
{
    bytes = round_up(bytes);
    if (lo_ + bytes > hi_)
        throw Error{-3, "arena: left allocation of " + std::to_string(bytes) + " bytes does not fit", true,
                    lo_ + (cap_ - hi_) + bytes};
    void *p = base_ + lo_;
    lo_ += bytes;
    if (lo_ + (cap_ - hi_) > peak_) peak_ = lo_ + (cap_ - hi_);
    return p;
}

void *Arena::right(size_t bytes)
{
    bytes = round_up(bytes);
    if (lo_ + bytes > hi_)
        throw Error{-3, "arena: right allocation of " + std::to_string(bytes) + " bytes does not fit", true,
                    lo_ + (cap_ - hi_) + bytes};
    hi_ -= bytes;
    if (lo_ + (cap_ - hi_) > peak_) peak_ = lo_ + (cap_ - hi_);
    return base_ + hi_;
}

void Arena::left_release(size_t mark) { lo_ = mark; }
void Arena::right_release(size_t mark) { hi_ = mark; }

void Arena::free_block()
{
    if (lo_ != 0 || hi_ != cap_) throw Error{-3, "arena: free_block while allocations are live"};
    if (base_) HGP_CHECK_HIP(hipFree(base_));
    base_ = nullptr;
    cap_ = lo_ = hi_ = 0;
}

void Arena::reset()
{
    lo_ = 0;
    hi_ = cap_;
    peak_ = 0;
}

// ------------------------------------------------------------------ context

std::unique_ptr<Ctx> create_context(int device)
{
    int cnt = 0;
    if (hipGetDeviceCount(&cnt) != hipSuccess || cnt == 0) {
        (void)hipGetLastError();
        throw Error{-5, "no HIP device"};
    }
    if (device < 0 || device >= cnt) throw Error{-1, "device ordinal out of range"};
    HGP_CHECK_HIP(hipSetDevice(device));
    auto c = std::make_unique<Ctx>();
    c->device = device;
    HGP_CHECK_HIP(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    HGP_CHECK_HIP(hipHostMalloc((void **)&c->pinned, kPinnedBytes, hipHostMallocDefault));
    for (auto &e : c->ev) HGP_CHECK_HIP(hipEventCreate(&e));
    return c;
}

void destroy_context(std::unique_ptr<Ctx> c)
{
    if (!c) return;
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    c->arena.reset();
    c->arena.free_block();
    for (auto &e : c->ev) (void)hipEventDestroy(e);
    (void)hipHostFree(c->pinned);
    (void)hipStreamDestroy(c->stream);
}

Ctx &context(int device)
{
    std::lock_guard<std::mutex> g(g_ctx_mu);
    auto it = g_ctx.find(device);
    if (it != g_ctx.end()) return *it->second;
    auto c = create_context(device);
    Ctx &ref = *c;
    g_ctx[device] = std::move(c);
    return ref;
}

KTimer &ktimer()
{
    static KTimer t;
    return t;
}

hipEvent_t KTimer::get()
{
    std::lock_guard<std::mutex> g(mu);
    if (used == pool.size()) {
        hipEvent_t e;
        HGP_CHECK_HIP(hipEventCreate(&e));
        pool.push_back(e);
    }
    return pool[used++];
}

int64_t KTimer::counter_slot()
{
    std::lock_guard<std::mutex> g(mu);
    if (!dcnt) {
        cnt_cap = (int64_t)kKbLanes << 14;
        if (hipMalloc((void **)&dcnt, sizeof(unsigned long long) * (size_t)cnt_cap) != hipSuccess) {
            (void)hipGetLastError();
            dcnt = nullptr;
            cnt_cap = 0;
        }
    }
    if (ncnt + kKbLanes > cnt_cap) return -1;
    int64_t s = ncnt;
    ncnt += kKbLanes;
    return s;
}

Tuning &tuning()
{
    static Tuning t;
    return t;
}

JetTrace &jet_trace()
{
    static JetTrace t;
    return t;
}

KScope::KScope(Ctx &c, const char *name, double bytes, bool counted)
    : c_(c), on_(ktimer().on), name_(name), bytes_(bytes)
{
    if (!on_) return;
    a_ = ktimer().get();
    b_ = ktimer().get();
    if (counted) {
        slot_ = ktimer().counter_slot();
        if (slot_ >= 0) {
            dev_ = ktimer().dcnt + slot_;
            HGP_CHECK_HIP(hipMemsetAsync(dev_, 0, sizeof(unsigned long long) * kKbLanes, c_.stream));
        }
    }
    HGP_CHECK_HIP(hipEventRecord(a_, c_.stream));
}

KScope::~KScope()
{
    if (!on_) return;
    if (hipEventRecord(b_, c_.stream) != hipSuccess) return;
    std::lock_guard<std::mutex> g(ktimer().mu);
    ktimer().recs.push_back({name_, a_, b_, bytes_, slot_});
}

void upload_staged(Ctx &c, void *ddst, const void *hsrc, size_t bytes)
{
    char *stage = (char *)c.pinned + kStageOffset;
    size_t cap = kPinnedBytes - kStageOffset;
    for (size_t off = 0; off < bytes; off += cap) {
        size_t b = std::min(cap, bytes - off);
        HGP_CHECK_HIP(hipStreamSynchronize(c.stream));   // staging area free again
        memcpy(stage, (const char *)hsrc + off, b);
        HGP_CHECK_HIP(hipMemcpyAsync((char *)ddst + off, stage, b, hipMemcpyHostToDevice, c.stream));
    }
    HGP_CHECK_HIP(hipStreamSynchronize(c.stream));
}

void read_scalars(Ctx &c, const void *dsrc, size_t bytes)
{
    HGP_CHECK_HIP(hipMemcpyAsync(c.pinned, dsrc, bytes, hipMemcpyDeviceToHost, c.stream));
    HGP_CHECK_HIP(hipStreamSynchronize(c.stream));
}

void read_scalars_start(Ctx &c, const void *dsrc, size_t bytes)
{
    HGP_CHECK_HIP(hipMemcpyAsync(c.pinned, dsrc, bytes, hipMemcpyDeviceToHost, c.stream));
    HGP_CHECK_HIP(hipEventRecord(c.ev[7], c.stream));
}

void read_scalars_wait(Ctx &c) { HGP_CHECK_HIP(hipEventSynchronize(c.ev[7])); }

}  // namespace hgp
