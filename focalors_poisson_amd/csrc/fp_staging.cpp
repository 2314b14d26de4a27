// Pinned host staging for the descriptor tables of large batches.  A copy
// from pageable host memory is not asynchronous (the runtime has to finish
// reading the source before the call returns, which here meant waiting for
// the kernels already queued), so a batched solve's host thread would stall
// behind the device at every table upload.  Blocks of pinned memory are kept
// for the process; a block is handed out again once the event recorded after
// its copy has completed, so an upload never waits for the device unless
// every block is still in flight (then a new block is allocated).
#include <algorithm>
#include <mutex>
#include <vector>

#include "fp_internal.h"

namespace fp {

namespace {
struct Block {
    void*      p    = nullptr;
    size_t     n    = 0;
    hipEvent_t ev   = nullptr;  // recorded after the block's last copy
    bool       held = false;    // handed out, copy not yet queued
};
std::mutex         g_mu;
std::vector<Block> g_pool;
}  // namespace

void* staging_get(size_t n)
{
    std::lock_guard<std::mutex> lk(g_mu);
    for (Block& b : g_pool) {
        if (b.held || b.n < n) continue;
        if (b.ev) {
            const hipError_t e = hipEventQuery(b.ev);
            if (e == hipErrorNotReady) continue;
            FP_HIP(e);
        }
        b.held = true;
        return b.p;
    }
    Block b;
    b.n = std::max<size_t>((n + 65535) & ~(size_t)65535, 256 << 10);
    FP_HIP(hipHostMalloc(&b.p, b.n, hipHostMallocPortable));
    b.held = true;
    g_pool.push_back(b);
    return b.p;
}

TableArena*& table_arena()
{
    thread_local TableArena* a = nullptr;
    return a;
}

void staging_put(void* p, hipStream_t s)
{
    std::lock_guard<std::mutex> lk(g_mu);
    for (Block& b : g_pool)
        if (b.p == p) {
            if (!b.ev) FP_HIP(hipEventCreateWithFlags(&b.ev, hipEventDisableTiming));
            b.held = false;
            FP_HIP(hipEventRecord(b.ev, s));
            return;
        }
    throw Error(FP_E_INVALID, "staging block not from the pool");
}

}  // namespace fp
