// Per-kernel timing with HIP events on the launching stream.  A region
// records (name, start, stop, algorithmic bytes, algorithmic flops); the
// pipeline aggregates by name.  Disabled unless a profiler is installed for
// the calling thread.
#pragma once

#include <cstring>
#include <string>
#include <vector>

#include "sev_common.hpp"

namespace sev {

struct Prof {
    struct Rec {
        const char *name;
        hipEvent_t a, b;
        double bytes, flops;
    };
    bool on = false;
    std::string only;  // when set, only the regions of this name are recorded
    std::vector<Rec> recs;
    std::vector<hipEvent_t> pool;

    hipEvent_t take() {
        if (!pool.empty()) {
            hipEvent_t e = pool.back();
            pool.pop_back();
            return e;
        }
        hipEvent_t e;
        SEV_HIP(hipEventCreate(&e));
        return e;
    }
    void reset() {
        for (auto &r : recs) {
            pool.push_back(r.a);
            pool.push_back(r.b);
        }
        recs.clear();
    }
    ~Prof() {
        reset();
        for (auto e : pool) (void)hipEventDestroy(e);
    }
    struct Agg {
        std::string name;
        int64_t launches = 0;
        double ms = 0, bytes = 0, flops = 0;
    };
    // Caller synchronises the stream first.
    std::vector<Agg> aggregate() {
        std::vector<Agg> out;
        for (auto &r : recs) {
            float ms = 0.f;
            SEV_HIP(hipEventElapsedTime(&ms, r.a, r.b));
            Agg *g = nullptr;
            for (auto &x : out)
                if (x.name == r.name) g = &x;
            if (!g) {
                out.push_back(Agg{r.name});
                g = &out.back();
            }
            g->launches += 1;
            g->ms += ms;
            g->bytes += r.bytes;
            g->flops += r.flops;
        }
        return out;
    }
};

extern thread_local Prof *t_prof;

struct ProfScope {
    Prof *p = nullptr;
    hipStream_t st;
    Prof::Rec rec;
    ProfScope(const char *name, hipStream_t s, double bytes = 0, double flops = 0) : st(s) {
        if (t_prof && t_prof->on && (t_prof->only.empty() || t_prof->only == name)) {
            p = t_prof;
            rec.name = name;
            rec.bytes = bytes;
            rec.flops = flops;
            rec.a = p->take();
            rec.b = p->take();
            SEV_HIP(hipEventRecord(rec.a, st));
        }
    }
    ~ProfScope() {
        if (p) {
            (void)hipEventRecord(rec.b, st);
            p->recs.push_back(rec);
        }
    }
};

}  // namespace sev
