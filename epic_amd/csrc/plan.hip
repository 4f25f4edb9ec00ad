// Plan: device arena upload, eager replay with phase timers, graph replay.
#include "plan.h"

namespace epic {

hipError_t Plan::upload(hipStream_t s)
{
    if (dev_ready_) return hipSuccess;
    if (exec_) {
        (void)hipGraphExecDestroy(exec_);
        exec_ = nullptr;
    }
    if (graph_) {
        (void)hipGraphDestroy(graph_);
        graph_ = nullptr;
    }
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

hipError_t Plan::prepare_graph(hipStream_t s)
{
    hipError_t e = upload(s);
    if (e != hipSuccess) return e;
    if (!exec_) {
        e = hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal);
        if (e != hipSuccess) return e;
        hipError_t er = run(s);
        e = hipStreamEndCapture(s, &graph_);
        if (er != hipSuccess) return er;
        if (e != hipSuccess) return e;
        e = hipGraphInstantiate(&exec_, graph_, nullptr, nullptr, 0);
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

hipError_t Plan::run_graph(hipStream_t s)
{
    hipError_t e = prepare_graph(s);
    if (e != hipSuccess) return e;
    return hipGraphLaunch(exec_, s);
}

void Plan::release()
{
    if (exec_) (void)hipGraphExecDestroy(exec_);
    if (graph_) (void)hipGraphDestroy(graph_);
    if (dev_) (void)hipFree(dev_);
    exec_ = nullptr;
    graph_ = nullptr;
    dev_ = nullptr;
    dev_size_ = 0;
    dev_ready_ = false;
    host_.clear();
    maps_.clear();
    recs_.clear();
}

}  // namespace epic
