// profile.cpp — event-pair recorder behind prof::Scope and the cuev_profile_* ABI.
#include "profile.hpp"

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

namespace cuev {
namespace prof {
namespace {

struct Rec {
    std::string tag;
    hipEvent_t a = nullptr, b = nullptr;
    double flops = 0, bytes = 0;
};

struct State {
    std::mutex mu;
    int on = -1; // -1: not yet read from the environment
    std::vector<Rec> recs;
    std::vector<hipEvent_t> pool;
    std::vector<std::string> only; // tag prefixes to record; empty: every tag
};

State &st() {
    static State s;
    return s;
}

hipEvent_t take(State &S) {
    if (!S.pool.empty()) {
        hipEvent_t e = S.pool.back();
        S.pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}

} // namespace

bool enabled() {
    State &S = st();
    if (S.on < 0) {
        std::lock_guard<std::mutex> lk(S.mu);
        if (S.on < 0) {
            const char *e = std::getenv("CUEV_PROFILE");
            S.on = (e && e[0] && e[0] != '0') ? 1 : 0;
        }
    }
    return S.on == 1;
}

bool selected(const State &S, const char *tag) {
    if (S.only.empty()) return true;
    for (const std::string &p : S.only)
        if (std::strncmp(tag, p.c_str(), p.size()) == 0) return true;
    return false;
}

Scope::Scope(hipStream_t s, const char *tag, double flops, double bytes) {
    if (!enabled()) return;
    State &S = st();
    std::lock_guard<std::mutex> lk(S.mu);
    if (!selected(S, tag)) return;
    Rec r;
    r.tag = tag;
    r.a = take(S);
    r.b = take(S);
    r.flops = flops;
    r.bytes = bytes;
    if (!r.a || !r.b) return;
    if (hipEventRecord(r.a, s) != hipSuccess) return;
    slot_ = (int)S.recs.size();
    s_ = s;
    S.recs.push_back(r);
}

Scope::~Scope() {
    if (slot_ < 0) return;
    State &S = st();
    std::lock_guard<std::mutex> lk(S.mu);
    if (slot_ < (int)S.recs.size()) (void)hipEventRecord(S.recs[slot_].b, s_);
}

} // namespace prof
} // namespace cuev

using namespace cuev::prof;

extern "C" {

int cuev_profile_enable(int on) {
    State &S = st();
    std::lock_guard<std::mutex> lk(S.mu);
    S.on = on ? 1 : 0;
    return CUEV_STATUS_SUCCESS;
}

int cuev_profile_select(const char *tags) {
    State &S = st();
    std::lock_guard<std::mutex> lk(S.mu);
    S.only.clear();
    if (!tags) return CUEV_STATUS_SUCCESS;
    std::string cur;
    for (const char *c = tags;; ++c) {
        if (*c == ',' || *c == 0) {
            if (!cur.empty()) S.only.push_back(cur);
            cur.clear();
            if (*c == 0) break;
        } else if (*c != ' ') {
            cur += *c;
        }
    }
    return CUEV_STATUS_SUCCESS;
}

int cuev_profile_read(char *buf, int len, int reset) {
    State &S = st();
    std::lock_guard<std::mutex> lk(S.mu);
    struct Acc {
        double ms = 0, flops = 0, bytes = 0, max_ms = 0;
        long count = 0;
    };
    std::map<std::string, Acc> acc;
    for (auto &r : S.recs) {
        float ms = 0.f;
        if (hipEventSynchronize(r.b) == hipSuccess && hipEventElapsedTime(&ms, r.a, r.b) == hipSuccess) {
            Acc &a = acc[r.tag];
            a.ms += ms;
            a.flops += r.flops;
            a.bytes += r.bytes;
            a.count += 1;
            if (ms > a.max_ms) a.max_ms = ms;
        }
    }
    std::string out = "{";
    bool first = true;
    char tmp[512];
    for (auto &kv : acc) {
        std::snprintf(tmp, sizeof(tmp),
                      "%s\"%s\": {\"ms\": %.6f, \"count\": %ld, \"flops\": %.6e, \"bytes\": %.6e, \"max_ms\": %.6f}",
                      first ? "" : ", ", kv.first.c_str(), kv.second.ms, kv.second.count,
                      kv.second.flops, kv.second.bytes, kv.second.max_ms);
        out += tmp;
        first = false;
    }
    out += "}";
    if (reset) {
        for (auto &r : S.recs) {
            S.pool.push_back(r.a);
            S.pool.push_back(r.b);
        }
        S.recs.clear();
    }
    if (buf && len > 0) {
        std::strncpy(buf, out.c_str(), (size_t)len - 1);
        buf[len - 1] = 0;
    }
    return (int)out.size() + 1;
}

} // extern "C"
