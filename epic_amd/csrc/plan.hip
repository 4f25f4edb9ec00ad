// Plan: device arena upload, eager replay with phase timers, graph replay
// (optionally with the phase timers captured as event-record nodes).
#include "plan.h"

namespace epic {

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
        if (dev_) (void)hipFree(dev_);
        dev_ = nullptr;
        dev_size_ = 0;
        hipError_t e = hipMalloc((void **)&dev_, host_.size());
        if (e != hipSuccess) return e;
        dev_size_ = host_.size();
    }
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
    return hipGraphLaunch(exec_, s);
}

void Plan::release()
{
    drop_graph();
    if (dev_) (void)hipFree(dev_);
    dev_ = nullptr;
    dev_size_ = 0;
    dev_ready_ = false;
    host_.clear();
    maps_.clear();
    recs_.clear();
}

}  // namespace epic
