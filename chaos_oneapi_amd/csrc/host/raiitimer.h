// Scope timer with the reference's behaviour (raiitimer.h:9-28): the clock
// receives whole seconds.  A second constructor also records milliseconds.
#pragma once

#include <chrono>
#include <cstdint>
#include <functional>

class RaiiTimer {
    std::chrono::high_resolution_clock::time_point t0;
    std::function<void(long long)> cb;
    uint32_t* clock = nullptr;
    double* clock_ms = nullptr;

public:
    explicit RaiiTimer(std::function<void(long long)> callback)
        : t0(std::chrono::high_resolution_clock::now()), cb(std::move(callback)) {}
    explicit RaiiTimer(uint32_t* clock_, double* clock_ms_ = nullptr)
        : t0(std::chrono::high_resolution_clock::now()), clock(clock_), clock_ms(clock_ms_) {}

    ~RaiiTimer() {
        auto t1 = std::chrono::high_resolution_clock::now();
        auto secs = static_cast<uint32_t>(std::chrono::duration_cast<std::chrono::seconds>(t1 - t0).count());
        if (clock_ms) *clock_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
        if (clock) *clock = secs;
        else if (cb) cb(secs);
    }
};
