// Multi-instance batch: several contexts (Hamiltonians of uneven sector
// size, each with its own uploaded angle batch) evaluated by one hipGraph
// launch.  Each context's batch sequence runs on its own stream; the graph is
// captured once per shape (batch sizes, gradient flag) from a driver stream
// that forks to every context's stream and joins them again, so the
// instances run as parallel branches of one graph.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "../../include/ansatzsynth.h"
#include "manager.hpp"

namespace ae {

struct MultiBatch {
    std::vector<Manager *> ms;
    hipStream_t s = nullptr;
    hipEvent_t fork = nullptr;
    std::vector<hipEvent_t> join;
    hipGraphExec_t g = nullptr;
    std::vector<int> shape;

    ~MultiBatch()
    {
        if (g) (void)hipGraphExecDestroy(g);
        for (hipEvent_t e : join) (void)hipEventDestroy(e);
        if (fork) (void)hipEventDestroy(fork);
        if (s) (void)hipStreamDestroy(s);
    }
};

namespace {

// every involved manager's lock, in address order
struct LockAll {
    std::vector<std::unique_lock<std::mutex>> locks;
    explicit LockAll(const std::vector<Manager *> &ms)
    {
        std::vector<Manager *> o(ms);
        std::sort(o.begin(), o.end());
        o.erase(std::unique(o.begin(), o.end()), o.end());
        for (Manager *m : o) locks.emplace_back(m->lock);
    }
};

bool run(MultiBatch &mb, bool grad, std::string &err)
{
    std::vector<Engine *> es;
    std::vector<int> shape;
    bool timing = false;
    for (Manager *m : mb.ms) {
        Engine *e = m->engine();
        if (!e || e->batch_size() < 1) {
            err = "a context has no uploaded angle batch";
            return false;
        }
        es.push_back(e);
        shape.push_back(e->batch_size());
        shape.push_back(e->generation());
        timing = timing || e->timing_on();
    }
    shape.push_back(grad ? 1 : 0);
    if (timing) {  // event timing is not captured: run eagerly
        for (Engine *e : es)
            if (!e->batch_enqueue(grad, err)) return false;
        return true;
    }
    if (mb.g && shape != mb.shape) {
        (void)hipGraphExecDestroy(mb.g);
        mb.g = nullptr;
    }
    if (!mb.g) {
        // one eager pass first: library kernels load lazily, not under capture
        for (Engine *e : es)
            if (!e->batch_enqueue(grad, err)) return false;
        for (Engine *e : es)
            if (!e->synchronize(err)) return false;
        if (hipStreamBeginCapture(mb.s, hipStreamCaptureModeThreadLocal) != hipSuccess) {
            err = "hipStreamBeginCapture failed";
            return false;
        }
        bool ok = hipEventRecord(mb.fork, mb.s) == hipSuccess;
        for (size_t i = 0; ok && i < es.size(); i++) {
            hipStream_t si = (hipStream_t)es[i]->stream();
            ok = hipStreamWaitEvent(si, mb.fork, 0) == hipSuccess && es[i]->batch_enqueue(grad, err) &&
                 hipEventRecord(mb.join[i], si) == hipSuccess && hipStreamWaitEvent(mb.s, mb.join[i], 0) == hipSuccess;
        }
        hipGraph_t graph = nullptr;
        const hipError_t ec = hipStreamEndCapture(mb.s, &graph);
        if (!ok || ec != hipSuccess) {
            if (graph) (void)hipGraphDestroy(graph);
            if (err.empty()) err = "multi-instance capture failed";
            return false;
        }
        const hipError_t ei = hipGraphInstantiate(&mb.g, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        if (ei != hipSuccess) {
            mb.g = nullptr;
            err = "hipGraphInstantiate failed";
            return false;
        }
        mb.shape = shape;
    }
    if (hipGraphLaunch(mb.g, mb.s) != hipSuccess) {
        err = "hipGraphLaunch failed";
        return false;
    }
    return true;
}

} // namespace
} // namespace ae

using ae::log_msg;
using ae::Manager;
using ae::MultiBatch;

extern "C" {

void *multiBatchCreate(int nctx, void *const *ctxs)
{
    if (nctx < 1 || ctxs == nullptr) {
        log_msg("multiBatchCreate: no contexts");
        return nullptr;
    }
    MultiBatch *mb = new (std::nothrow) MultiBatch;
    if (!mb) return nullptr;
    for (int i = 0; i < nctx; i++) {
        if (!ctxs[i]) {
            log_msg("multiBatchCreate: context %d is NULL", i);
            delete mb;
            return nullptr;
        }
        mb->ms.push_back(static_cast<Manager *>(ctxs[i]));
    }
    bool ok = hipStreamCreateWithFlags(&mb->s, hipStreamNonBlocking) == hipSuccess &&
              hipEventCreateWithFlags(&mb->fork, hipEventDisableTiming) == hipSuccess;
    mb->join.resize(nctx, nullptr);
    for (int i = 0; ok && i < nctx; i++) ok = hipEventCreateWithFlags(&mb->join[i], hipEventDisableTiming) == hipSuccess;
    if (!ok) {
        log_msg("multiBatchCreate: stream or event creation failed");
        delete mb;
        return nullptr;
    }
    return mb;
}

int multiBatchDestroy(void *mb)
{
    if (!mb) return 1;
    delete static_cast<MultiBatch *>(mb);
    return 0;
}

int multiBatchUpload(void *mbp, int index, int nvec, int nangles, const double *angles)
{
    if (!mbp) return 1;
    MultiBatch &mb = *static_cast<MultiBatch *>(mbp);
    if (index < 0 || index >= (int)mb.ms.size() || angles == nullptr) return 2;
    if (nvec < 1 || nangles < 1) return 3;
    // the previous launch may still read the angle tables
    if (hipStreamSynchronize(mb.s) != hipSuccess) {
        log_msg("multiBatchUpload: stream synchronisation failed");
        return 5;
    }
    return uploadAngleBatch(nvec, nangles, angles, mb.ms[index]);
}

int multiBatchRun(void *mbp, int withgradient)
{
    if (!mbp) return 1;
    MultiBatch &mb = *static_cast<MultiBatch *>(mbp);
    ae::LockAll lk(mb.ms);
    std::string err;
    if (!ae::run(mb, withgradient != 0, err)) {
        log_msg("multiBatchRun failed: %s", err.c_str());
        return 5;
    }
    return 0;
}

int multiBatchDownload(void *mbp, int index, double *energies, double *gradients)
{
    if (!mbp) return 1;
    MultiBatch &mb = *static_cast<MultiBatch *>(mbp);
    if (index < 0 || index >= (int)mb.ms.size() || energies == nullptr) return 2;
    if (hipStreamSynchronize(mb.s) != hipSuccess) {
        log_msg("multiBatchDownload: stream synchronisation failed");
        return 5;
    }
    return downloadAngleBatch(energies, gradients, mb.ms[index]);
}

} // extern "C"
