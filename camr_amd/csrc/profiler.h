// profiler.h -- named hipEvent timers around kernel launches, the analogue
// of the reference's BL_PROFILE regions (e.g. Godunov_3D.cpp:45,
// Hydro_umdrv.cpp:36) for the gfx950 path.  Disabled by default; when a
// name filter is set only launches of that name are bracketed by events.
#pragma once

#include <hip/hip_runtime.h>

#include <map>
#include <string>
#include <vector>

namespace camr {

class Prof {
  public:
    ~Prof() {
        collect();
        for (auto &e : pool_) (void)hipEventDestroy(e);
    }
    void enable(bool on, const char *filter) {
        on_ = on;
        filter_ = filter ? filter : "";
    }
    bool enabled() const { return on_; }
    bool active(const char *name) const {
        return on_ && (filter_.empty() || filter_ == name);
    }
    void begin(const char *name, hipStream_t s) {
        if (!active(name)) return;
        Rec r;
        r.name = name;
        r.a = get();
        r.b = get();
        (void)hipEventRecord(r.a, s);
        open_.push_back(r);
    }
    void end(const char *name, hipStream_t s) {
        if (!active(name) || open_.empty()) return;
        Rec r = open_.back();
        open_.pop_back();
        (void)hipEventRecord(r.b, s);
        done_.push_back(r);
    }
    // Wait for recorded events and accumulate their durations.
    void collect() {
        for (auto &r : done_) {
            float ms = 0.f;
            if (hipEventSynchronize(r.b) == hipSuccess &&
                hipEventElapsedTime(&ms, r.a, r.b) == hipSuccess) {
                auto &a = acc_[r.name];
                a.first += 1;
                a.second += ms;
            }
            pool_.push_back(r.a);
            pool_.push_back(r.b);
        }
        done_.clear();
    }
    void reset() {
        collect();
        acc_.clear();
    }
    std::string json() {
        collect();
        std::string s = "{";
        bool first = true;
        for (auto &kv : acc_) {
            char buf[256];
            snprintf(buf, sizeof(buf), "%s\"%s\": {\"count\": %ld, \"ms\": %.6f}",
                     first ? "" : ", ", kv.first.c_str(), kv.second.first,
                     kv.second.second);
            s += buf;
            first = false;
        }
        return s + "}";
    }

  private:
    struct Rec {
        std::string name;
        hipEvent_t a, b;
    };
    hipEvent_t get() {
        if (!pool_.empty()) {
            hipEvent_t e = pool_.back();
            pool_.pop_back();
            return e;
        }
        hipEvent_t e;
        (void)hipEventCreate(&e);
        return e;
    }
    bool on_ = false;
    std::string filter_;
    std::vector<Rec> open_, done_;
    std::vector<hipEvent_t> pool_;
    std::map<std::string, std::pair<long, double>> acc_;
};

// RAII region: Region r(prof, "name", stream); ... launches ...
struct Region {
    Prof *p;
    const char *n;
    hipStream_t s;
    Region(Prof *p_, const char *n_, hipStream_t s_) : p(p_), n(n_), s(s_) {
        if (p) p->begin(n, s);
    }
    ~Region() {
        if (p) p->end(n, s);
    }
};

} // namespace camr
