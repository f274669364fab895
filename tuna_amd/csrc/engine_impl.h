// engine_impl.h — host-side internals shared by the engine translation units
// (context, error plumbing, device buffer pool, kernel timing).
#pragma once
#include "../../include/tuna_gpu.h"
#include "tg_kernels.h"

#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <string>
#include <vector>

namespace tgi {

// thread-local message of the last failure (tg_last_error)
extern thread_local std::string g_err;

inline int fail(int code, const char* fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

#define HIPOK(expr)                                                                     \
    do {                                                                                \
        hipError_t e_ = (expr);                                                         \
        if (e_ != hipSuccess)                                                           \
            return fail(TG_EDEVICE, "%s failed: %s", #expr, hipGetErrorString(e_));   \
    } while (0)

inline uint32_t round_pow2(uint32_t n)
{
    // src/setup.hpp:53-59
    if (n <= 1) return 1;
    if ((n & (n - 1)) == 0) return n;
    while (n & (n - 1)) n &= n - 1;
    return n << 1;
}

inline uint32_t ilog2(uint64_t x)
{
    uint32_t r = 0;
    while ((1ull << r) < x) ++r;
    return r;
}

struct Buf {
    void* p = nullptr;
    size_t cap = 0;
};

struct KTime {
    double ms = 0;
    uint64_t n = 0;
};

struct BatchState;
void destroy_batch(BatchState* b);
// One job through the batched engine; fills st and the context's result.
int count_one_batched(tg_ctx* c, const tg_params* p, const uint8_t* d_bases, uint64_t n_bases,
                      const uint64_t* d_off, uint64_t n_chunks, tg_stats* st);
// The job of chunks [c0, c1) (bases [b0, b1)) of a device-resident input whose
// offsets are absolute, through the batched engine.
int count_range_batched(tg_ctx* c, const tg_params* p, const uint8_t* d_bases, uint64_t n_bases,
                        const uint64_t* d_off, uint64_t n_chunks, uint64_t c0, uint64_t c1, uint64_t b0,
                        uint64_t b1, tg_stats* st);
// A single job of more bases than this is counted in slices (tg_batch_config::slice_bases).
constexpr uint64_t DEFAULT_SLICE_BASES = (1ull << 31) - (1ull << 20);
uint64_t slice_limit(tg_ctx* c);
// Forget the batched engine's single-job result view (the context's result has moved).
void drop_single_view(tg_ctx* c);
// Resolve the device's copy of the no-allocation scratch array (format.hip).
void noalloc_setup(int device);

}  // namespace tgi

struct tg_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::map<std::string, tgi::Buf> bufs;
    bool profiling = false;
    std::vector<std::pair<std::string, std::pair<hipEvent_t, hipEvent_t>>> pending;
    std::vector<hipEvent_t> ev_pool;
    std::map<std::string, tgi::KTime> ktimes;
    hipEvent_t ev_a = nullptr, ev_b = nullptr, ev_c = nullptr;
    // second stream (default priority) of the k > 64 path, made on first use
    hipStream_t side = nullptr;
    hipEvent_t ev_side = nullptr;
    uint32_t wide_dq = 16;  // deque entries of its superkmer pass (tg_set_wide_deque)
    uint64_t input_budget = 0;  // bytes of file input kept on the device (0: half the free memory)
    // last job
    uint64_t n_out = 0;
    uint32_t nw = 1;
    uint64_t* d_keys = nullptr;
    uint32_t* d_cnt = nullptr;
    // batched jobs (count_many): buffers, captured graph and results
    tgi::BatchState* batch = nullptr;
    // single jobs: false = through the batched engine (one captured launch
    // sequence, no host round trip); true = the plain host-driven sequence
    bool plain = false;
    // debug statistics per reference partition of the last plain-sequence job
    // (work items then never span two partitions)
    bool debug = false;
    std::vector<uint64_t> dbg_inserted, dbg_unique, dbg_spilled, dbg_cap;
};

namespace tgi {

inline int ensure(tg_ctx* c, const char* name, size_t bytes, void** out)
{
    Buf& b = c->bufs[name];
    if (bytes == 0) bytes = 256;
    if (b.cap < bytes) {
        if (b.p) HIPOK(hipFree(b.p));
        b.p = nullptr;
        size_t want = bytes + bytes / 4;
        hipError_t e = hipMalloc(&b.p, want);
        if (e != hipSuccess) {
            want = bytes;
            e = hipMalloc(&b.p, want);
        }
        if (e != hipSuccess) {
            b.p = nullptr;
            b.cap = 0;
            (void)hipGetLastError();
            return fail(TG_ENOMEM, "device allocation of %zu bytes for %s failed", bytes, name);
        }
        b.cap = want;
    }
    *out = b.p;
    return TG_OK;
}

template <class T>
inline int ensure_t(tg_ctx* c, const char* name, size_t count, T** out)
{
    void* p = nullptr;
    const int rc = ensure(c, name, count * sizeof(T), &p);
    *out = static_cast<T*>(p);
    return rc;
}

inline hipEvent_t take_event(tg_ctx* c)
{
    if (!c->ev_pool.empty()) {
        hipEvent_t e = c->ev_pool.back();
        c->ev_pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    hipEventCreate(&e);
    return e;
}

// Kernel timing on the context stream: events bracket each launch.
struct Prof {
    tg_ctx* c;
    const char* name;
    hipEvent_t a = nullptr;
    Prof(tg_ctx* cc, const char* n) : c(cc), name(n)
    {
        if (c && c->profiling) {
            a = take_event(c);
            hipEventRecord(a, c->stream);
        }
    }
    ~Prof()
    {
        if (c && c->profiling) {
            hipEvent_t b = take_event(c);
            hipEventRecord(b, c->stream);
            c->pending.push_back({name, {a, b}});
        }
    }
};

inline void collect_prof(tg_ctx* c)
{
    if (c->pending.empty()) return;
    hipStreamSynchronize(c->stream);
    for (auto& pe : c->pending) {
        float ms = 0;
        hipEventElapsedTime(&ms, pe.second.first, pe.second.second);
        KTime& kt = c->ktimes[pe.first];
        kt.ms += ms;
        kt.n += 1;
        c->ev_pool.push_back(pe.second.first);
        c->ev_pool.push_back(pe.second.second);
    }
    c->pending.clear();
}

inline int check_params(const tg_params* p, uint64_t n_bases, uint64_t n_chunks)
{
    if (!p) return fail(TG_EINVAL, "null parameters");
    if (p->k < 2 || p->k > 256)
        return fail(TG_EINVAL, "k-mer length (-k %u) must be in [2, 256]", p->k);
    if (p->m < 1 || p->m > 32)
        return fail(TG_EINVAL, "minimizer length (-m %u) must be in [1, 32]", p->m);
    if (p->m >= p->k)
        return fail(TG_EINVAL,
                    "minimizer length (-m %u) must be strictly less than k-mer length (-k %u)",
                    p->m, p->k);
    (void)n_bases;
    if (n_chunks >= (1ull << 40))
        return fail(TG_EINVAL, "too many chunks (%llu)", (unsigned long long)n_chunks);
    return TG_OK;
}

// One pass of the engine (a job, or a slice of one) addresses bases, records
// and chunks with 32-bit indices.
inline int check_pass(uint64_t n_bases, uint64_t n_chunks)
{
    if (n_bases >= (1ull << 32) - 64)
        return fail(TG_EINVAL, "a pass of %llu bases exceeds the 2^32 limit of one pass",
                    (unsigned long long)n_bases);
    if (n_chunks >= (1ull << 32) - 1)
        return fail(TG_EINVAL, "too many chunks in one pass (%llu)", (unsigned long long)n_chunks);
    return TG_OK;
}

}  // namespace tgi
