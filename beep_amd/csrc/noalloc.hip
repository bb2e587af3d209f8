// noalloc.hip -- beep_count_noalloc: the enumeration launch sequence of
// beep_count with every scratch buffer carved from one fixed device array,
// for callers that cannot allocate device memory.
//
// The array lives for the whole process on every GPU the library runs on.
// In host code its name only stands for the symbol: its address on a device
// comes from hipGetSymbolAddress, taken once per device with that device
// current, and only that address is handed to kernels and copies.  Every call
// on one GPU shares the array and leaves the last call's contents in it; the
// launch sequence clears what must start at zero (counters, work cursors,
// hub bitmaps and hashes) with fill kernels.  Calls that overlap on one GPU
// race on it.

#include <hip/hip_runtime.h>

#include <chrono>
#include <mutex>

#include "engine.hpp"

static __device__ __align__(256) unsigned char scratch[BEEP_NOALLOC_SCRATCH_BYTES];

namespace beep {

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
char *g_scratch_ptr[kMaxDevices] = {};  // per device: the address of `scratch` there

// Address of the array on `device`, which must be current.
char *scratch_on(int device)
{
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (device < 0 || device >= kMaxDevices) throw Error(BEEP_EDEVICE, "device ordinal out of range");
    if (!g_scratch_ptr[device]) {
        void *p = nullptr;
        HIP_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) throw Error(BEEP_EDEVICE, "no device address for the allocation-free scratch array");
        g_scratch_ptr[device] = (char *)p;
    }
    return g_scratch_ptr[device];
}

// The launch sequence's scratch cut to the array: counters, hub arena and
// hash, DFS stacks and tuple buffer must fit whole; the donation ring takes
// what is left (a hub item whose subtree finds the ring full is walked where
// it is, so a smaller ring changes where work runs, not the result).
ScratchNeeds fit_to_array(ScratchNeeds n)
{
    ScratchNeeds req = n;
    req.queue_cap = req.slot_cap = 0;
    const uint64_t base = req.bytes();
    if (base > BEEP_NOALLOC_SCRATCH_BYTES)
        throw Error(BEEP_ERANGE, "the query needs " + std::to_string(base) +
                                     " bytes of scratch; beep_count_noalloc holds at most " +
                                     std::to_string((unsigned long long)BEEP_NOALLOC_SCRATCH_BYTES) +
                                     " (use beep_count)");
    if (n.queue_cap && n.slot_cap) {
        const uint64_t left = BEEP_NOALLOC_SCRATCH_BYTES - base;
        const uint64_t want = n.queue_cap * sizeof(HubChunk) + n.slot_cap * 8;
        if (left >= want + 512) return n;
        // scale both parts down alike; each keeps 256 bytes for its alignment
        const double f = left > 512 ? (double)(left - 512) / (double)want : 0.0;
        n.queue_cap = (uint64_t)((double)n.queue_cap * f);
        n.slot_cap = (uint64_t)((double)n.slot_cap * f);
        if (n.queue_cap == 0 || n.slot_cap == 0) n.queue_cap = n.slot_cap = 0;  // no ring at all
    } else {
        n.queue_cap = n.slot_cap = 0;
    }
    if (n.bytes() > BEEP_NOALLOC_SCRATCH_BYTES) throw Error(BEEP_EINVAL, "scratch layout overflow");
    return n;
}

}  // namespace

void count_noalloc(Prep &pp, const Plan &p, const beep_opts &o, void *stream, beep_result &res)
{
    hipStream_t st = (hipStream_t)stream;
    if (o.profile) throw Error(BEEP_EINVAL, "beep_count_noalloc does not take opts.profile (use beep_count)");
    if (!pp.dev || pp.dev->device != o.device || pp.dev->roots.skip_split != p.skip_split ||
        !pp.var->dev)
        throw Error(BEEP_EINVAL, "the prepared query is not resident on this device: call beep_prep_upload first");
    DevPrep &d = *pp.dev;
    HIP_CHECK(hipSetDevice(d.device));
    char *base = scratch_on(d.device);
    const ScratchNeeds need = fit_to_array(scratch_needs(pp, p, o, d.cus));
    Scratch s;
    carve(s, base, need);
    const LaunchCfg c = make_launch(pp, p, o, s, d.cus);

    unsigned long long acc[kCtrWords];
    const auto t0 = std::chrono::steady_clock::now();
    HIP_CHECK(launch_begin_sequence(s.ctr, kCtrTotal, kCtrDeadline, c.deadline_ticks, c.deadline_ticks_lds, st));
    enqueue_lds_classes(pp, c, s, &st, 1);
    enqueue_hubs(pp, c, s, st);
    HIP_CHECK(hipMemcpyAsync(acc, s.ctr, sizeof(unsigned long long) * kCtrWords, hipMemcpyDeviceToHost, st));
    HIP_CHECK(hipStreamSynchronize(st));
    res.t_enum_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    fill_result(o, acc, res);
    if (c.out.buf && o.match_buf && res.n_written) {
        HIP_CHECK(hipMemcpyAsync(o.match_buf, c.out.buf, res.n_written * (uint64_t)p.L * sizeof(uint32_t),
                                 hipMemcpyDeviceToHost, st));
        HIP_CHECK(hipStreamSynchronize(st));
    }
}

}  // namespace beep
