// Plan: device arena upload, eager replay with phase timers, graph replay
// (optionally with the phase timers captured as event-record nodes).
#include "plan.h"

#include <mutex>
#include <utility>

namespace epic {

// Device arenas of released plans, for reuse: hipFree waits for the whole
// device (every stream), which would stall the members' steps in flight
// each time a transient plan (a parcel operation's launches) is dropped.
// A block returns with an event recorded behind the last work that used it;
// the next user's stream waits for that event before its upload.  The
// runtime has refused to synchronise on such an event ("operation not
// permitted on an event last recorded in a capturing stream") although its
// stream was not capturing when it was recorded (streams that held a
// capture earlier); a refused wait falls back to a device synchronisation,
// and the refusal is cleared from the thread's last-error slot, where it
// would otherwise be read as the failure of a later, unrelated launch.
namespace {
struct PoolBlock {
    char *p;
    size_t size;
    int dev;
    hipEvent_t ev;   // may be null (never launched)
};
std::mutex pool_mu;
std::vector<PoolBlock> pool;
constexpr size_t POOL_MAX_BLOCKS = 256;

size_t pool_round(size_t n) { return (n + 4095) & ~size_t(4095); }

hipError_t pool_get(size_t bytes, hipStream_t s, char **out, size_t *got)
{
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    const size_t want = pool_round(bytes);
    {
        std::lock_guard<std::mutex> lk(pool_mu);
        size_t best = pool.size();
        for (size_t i = 0; i < pool.size(); ++i)   // best fit, at most 4x the request
            if (pool[i].dev == dev && pool[i].size >= want && pool[i].size <= 4 * want &&
                (best == pool.size() || pool[i].size < pool[best].size))
                best = i;
        if (best < pool.size()) {
            PoolBlock b = pool[best];
            pool.erase(pool.begin() + (long)best);
            if (b.ev) {
                e = hipStreamWaitEvent(s, b.ev, 0);
                if (e != hipSuccess) {   // a refused wait: every stream drained instead
                    (void)hipGetLastError();
                    e = hipDeviceSynchronize();
                }
                if (hipEventDestroy(b.ev) != hipSuccess) (void)hipGetLastError();
                if (e != hipSuccess) {
                    (void)hipFree(b.p);
                    return e;
                }
            }
            *out = b.p;
            *got = b.size;
            return hipSuccess;
        }
    }
    e = hipMalloc((void **)out, want);
    if (e != hipSuccess) return e;
    *got = want;
    return hipSuccess;
}

void pool_put(char *p, size_t size, hipStream_t last)
{
    if (!p) return;
    int dev = 0;
    hipEvent_t ev = nullptr;
    if (hipGetDevice(&dev) != hipSuccess) {
        (void)hipFree(p);
        return;
    }
    if (last) {
        if (hipEventCreateWithFlags(&ev, hipEventDisableTiming) != hipSuccess) ev = nullptr;
        if (ev && hipEventRecord(ev, last) != hipSuccess) {
            (void)hipEventDestroy(ev);
            ev = nullptr;
        }
        if (!ev) {   // no event: wait here for the stream instead
            (void)hipStreamSynchronize(last);
        }
    }
    std::lock_guard<std::mutex> lk(pool_mu);
    pool.push_back(PoolBlock{p, size, dev, ev});
    while (pool.size() > POOL_MAX_BLOCKS) {   // the oldest block leaves the pool
        PoolBlock b = pool.front();
        pool.erase(pool.begin());
        if (b.ev) {
            if (hipEventSynchronize(b.ev) != hipSuccess) {   // refused: hipFree below drains the device
                (void)hipGetLastError();
            }
            if (hipEventDestroy(b.ev) != hipSuccess) (void)hipGetLastError();
        }
        (void)hipFree(b.p);
    }
}
}  // namespace

void Plan::drop_graph()
{
    if (exec_) (void)hipGraphExecDestroy(exec_);
    if (graph_) (void)hipGraphDestroy(graph_);
    exec_ = nullptr;
    graph_ = nullptr;
    if (stamps_) (void)hipFree(stamps_);
    stamps_ = nullptr;
    stamp_phase_.clear();
}

// the device wall clock (constant frequency) into buf[i], with a vector store
__global__ void k_stamp(unsigned long long *buf, int i)
{
    if (threadIdx.x == 0) buf[i] = (unsigned long long)wall_clock64();
}

// the launches with a stamp before and after every run of one phase
hipError_t Plan::run_stamped(hipStream_t s)
{
    int cur = -1, k = 0;
    for (const Rec &r : recs_) {
        if (r.phase != cur) {
            if (cur >= 0) hipLaunchKernelGGL(k_stamp, dim3(1), dim3(64), 0, s, stamps_, 2 * k - 1);
            hipLaunchKernelGGL(k_stamp, dim3(1), dim3(64), 0, s, stamps_, 2 * k);
            ++k;
            cur = r.phase;
        }
        const void *tab = dev_ + r.tab_off;
        const void *map = dev_ + r.map_off;
        void *args[2] = {(void *)&tab, (void *)&map};
        hipError_t e = hipLaunchKernel(r.fn, dim3(r.nblocks), dim3(r.threads), args, r.lds, s);
        if (e != hipSuccess) return e;
    }
    if (cur >= 0) hipLaunchKernelGGL(k_stamp, dim3(1), dim3(64), 0, s, stamps_, 2 * k - 1);
    return hipGetLastError();
}

hipError_t Plan::graph_phase_times(std::vector<std::pair<int, float>> &out)
{
    out.clear();
    if (!timed_ || !stamps_ || stamp_phase_.empty() || !(clock_khz_ > 0.0)) return hipSuccess;
    stamps_h_.resize(2 * stamp_phase_.size());
    hipError_t e = hipMemcpy(stamps_h_.data(), stamps_, sizeof(unsigned long long) * stamps_h_.size(), hipMemcpyDeviceToHost);
    if (e != hipSuccess) return e;
    for (size_t k = 0; k < stamp_phase_.size(); ++k)
        out.emplace_back(stamp_phase_[k], (float)((double)(stamps_h_[2 * k + 1] - stamps_h_[2 * k]) / clock_khz_));
    return hipSuccess;
}

hipError_t Plan::upload(hipStream_t s)
{
    if (dev_ready_) return hipSuccess;
    drop_graph();
    if (host_.size() > dev_size_) {
        pool_put(dev_, dev_size_, last_stream_);
        dev_ = nullptr;
        dev_size_ = 0;
        hipError_t e = pool_get(host_.size(), s, &dev_, &dev_size_);
        if (e != hipSuccess) return e;
    } else if (dev_ && last_stream_ && last_stream_ != s) {   // earlier launches on another stream read the arena
        hipEvent_t ev = nullptr;
        hipError_t e = hipEventCreateWithFlags(&ev, hipEventDisableTiming);
        if (e == hipSuccess) e = hipEventRecord(ev, last_stream_);
        if (e == hipSuccess) e = hipStreamWaitEvent(s, ev, 0);
        if (ev) (void)hipEventDestroy(ev);
        if (e != hipSuccess) return e;
    }
    last_stream_ = s;
    if (!host_.empty()) {
        hipError_t e = hipMemcpyAsync(dev_, host_.data(), host_.size(), hipMemcpyHostToDevice, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        if (e != hipSuccess) return e;
    }
    dev_ready_ = true;
    return hipSuccess;
}

hipError_t Plan::run(hipStream_t s, std::vector<hipEvent_t> *ev, std::vector<int> *ev_phase)
{
    hipError_t e = upload(s);
    if (e != hipSuccess) return e;
    last_stream_ = s;
    int cur = -1;
    hipEvent_t stop = nullptr;
    for (const Rec &r : recs_) {
        if (ev && r.phase != cur) {
            if (stop) (void)hipEventRecord(stop, s);
            hipEvent_t a = nullptr, b = nullptr;
            stop = nullptr;
            if (hipEventCreate(&a) == hipSuccess && hipEventCreate(&b) == hipSuccess) {
                (void)hipEventRecord(a, s);
                ev->push_back(a);
                ev->push_back(b);
                ev_phase->push_back(r.phase);
                stop = b;
            }
            cur = r.phase;
        }
        const void *tab = dev_ + r.tab_off;
        const void *map = dev_ + r.map_off;
        void *args[2] = {(void *)&tab, (void *)&map};
        e = hipLaunchKernel(r.fn, dim3(r.nblocks), dim3(r.threads), args, r.lds, s);
        if (e != hipSuccess) return e;
    }
    if (stop) (void)hipEventRecord(stop, s);
    return hipSuccess;
}

hipError_t Plan::prepare_graph(hipStream_t s, bool timed)
{
    hipError_t e = upload(s);
    if (e != hipSuccess) return e;
    if (exec_ && timed != timed_) drop_graph();
    if (!exec_) {
        timed_ = timed;
        if (timed) {   // the stamp buffer and the phase of every run, before the capture
            int cur = -1;
            for (const Rec &r : recs_)
                if (r.phase != cur) {
                    stamp_phase_.push_back(r.phase);
                    cur = r.phase;
                }
            e = hipMalloc((void **)&stamps_, sizeof(unsigned long long) * 2 * std::max<size_t>(1, stamp_phase_.size()));
            if (e != hipSuccess) return e;
            int dev = 0, khz = 0;
            e = hipGetDevice(&dev);
            if (e == hipSuccess) e = hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev);
            if (e != hipSuccess) return e;
            clock_khz_ = (double)khz;
        }
        e = hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal);
        if (e != hipSuccess) return e;
        hipError_t er = timed ? run_stamped(s) : run(s);
        e = hipStreamEndCapture(s, &graph_);
        if (er != hipSuccess) return er;
        if (e != hipSuccess) return e;
        e = hipGraphInstantiate(&exec_, graph_, nullptr, nullptr, 0);
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

hipError_t Plan::run_graph(hipStream_t s, bool timed)
{
    hipError_t e = prepare_graph(s, timed);
    if (e != hipSuccess) return e;
    last_stream_ = s;
    return hipGraphLaunch(exec_, s);
}

hipError_t Plan::refresh_from(const Plan &o, hipStream_t s, bool *same)
{
    *same = false;
    if (!dev_ready_ || !dev_ || recs_.size() != o.recs_.size() || host_.size() != o.host_.size()) return hipSuccess;
    for (size_t i = 0; i < recs_.size(); ++i) {
        const Rec &a = recs_[i], &b = o.recs_[i];
        if (a.fn != b.fn || a.nblocks != b.nblocks || a.threads != b.threads || a.lds != b.lds || a.tab_off != b.tab_off ||
            a.map_off != b.map_off || a.phase != b.phase)
            return hipSuccess;
    }
    *same = true;
    if (host_ == o.host_) return hipSuccess;
    host_ = o.host_;
    maps_ = o.maps_;
    if (last_stream_ && last_stream_ != s) {   // the arena's readers so far ran on another stream
        hipEvent_t ev = nullptr;
        hipError_t e = hipEventCreateWithFlags(&ev, hipEventDisableTiming);
        if (e == hipSuccess) e = hipEventRecord(ev, last_stream_);
        if (e == hipSuccess) e = hipStreamWaitEvent(s, ev, 0);
        if (ev) (void)hipEventDestroy(ev);
        if (e != hipSuccess) return e;
    }
    last_stream_ = s;
    hipError_t e = hipMemcpyAsync(dev_, host_.data(), host_.size(), hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);   // host_ may change after this returns
    return e;
}

void Plan::swap(Plan &o)
{
    std::swap(host_, o.host_);
    std::swap(maps_, o.maps_);
    std::swap(recs_, o.recs_);
    std::swap(dev_, o.dev_);
    std::swap(dev_size_, o.dev_size_);
    std::swap(last_stream_, o.last_stream_);
    std::swap(dev_ready_, o.dev_ready_);
    std::swap(graph_, o.graph_);
    std::swap(exec_, o.exec_);
    std::swap(timed_, o.timed_);
    std::swap(stamps_, o.stamps_);
    std::swap(stamp_phase_, o.stamp_phase_);
    std::swap(stamps_h_, o.stamps_h_);
    std::swap(clock_khz_, o.clock_khz_);
}

void Plan::release()
{
    drop_graph();
    pool_put(dev_, dev_size_, last_stream_);
    dev_ = nullptr;
    dev_size_ = 0;
    dev_ready_ = false;
    last_stream_ = nullptr;
    host_.clear();
    maps_.clear();
    recs_.clear();
}

}  // namespace epic
