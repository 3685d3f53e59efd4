// graph_cache.hpp -- a device function's launch sequence as a HIP graph.
//
// A colour-phase function issues one kernel per phase (32 for a symmetric
// Gauss-Seidel step on 27-pt 128^3, 116 for k = 8 matrix powers).  The
// sequence depends only on the function, the sweep direction, the vectors
// and the power, so it is captured once per such key on a private stream,
// instantiated, and replayed with one hipGraphLaunch on the caller's
// stream.  Only plain kernel launches are captured (no memset nodes).
#pragma once
#include <hip/hip_runtime.h>

#include <deque>

#include "exec.hpp"

namespace race {

class GraphCache {
  public:
    ~GraphCache() { clear(); }
    // run f through its graph for (rev, x, b, power); on any capture error
    // the function runs directly from then on
    int run(DeviceFn* f, bool rev, hipStream_t st) {
        if (off) return f->run(rev, st);
        Entry* en = nullptr;
        for (Entry& e : cache)
            if (e.rev == rev && e.x == f->x && e.b == f->b && e.power == f->power) en = &e;
        if (!en) {
            if (!cs && hipStreamCreateWithFlags(&cs, hipStreamNonBlocking) != hipSuccess) {
                off = true;
                return f->run(rev, st);
            }
            Entry e{rev, f->x, f->b, f->power, nullptr, nullptr};
            if (hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal) != hipSuccess) {
                off = true;
                return f->run(rev, st);
            }
            const int err = f->run(rev, cs);
            const hipError_t ce = hipStreamEndCapture(cs, &e.g);
            if (err || ce != hipSuccess || !e.g ||
                hipGraphInstantiate(&e.ex, e.g, nullptr, nullptr, 0) != hipSuccess) {
                if (e.g) (void)hipGraphDestroy(e.g);
                (void)hipGetLastError();
                off = true;
                return err ? err : f->run(rev, st);
            }
            if (cache.size() >= kMaxEntries) {
                destroy(cache.front());
                cache.pop_front();
            }
            cache.push_back(e);
            en = &cache.back();
        }
        RACE_HIP(hipGraphLaunch(en->ex, st));
        return OK;
    }
    void clear() {
        for (Entry& e : cache) destroy(e);
        cache.clear();
        if (cs) (void)hipStreamDestroy(cs);
        cs = nullptr;
    }

  private:
    struct Entry {
        bool rev;
        void *x, *b;
        int power;
        hipGraph_t g;
        hipGraphExec_t ex;
    };
    static constexpr size_t kMaxEntries = 8;
    // a replay may still be queued on some stream: wait for the device first
    static void destroy(Entry& e) {
        (void)hipDeviceSynchronize();
        if (e.ex) (void)hipGraphExecDestroy(e.ex);
        if (e.g) (void)hipGraphDestroy(e.g);
    }
    std::deque<Entry> cache;
    hipStream_t cs = nullptr;
    bool off = false;
};

}  // namespace race
