#pragma once

// Timing harness with the reference's shape (src/bench.hpp:1-46): `warmup` untimed passes,
// then `runs` passes with device events around upload, compute and download on `stream` (the
// default stream unless the caller enqueues its work elsewhere); one-time setup happens before
// run().

#include "rfhip.h"

#include <functional>

struct BenchConfig {
    int warmup = 0;
    int runs = 1;
    rf_stream stream = nullptr;
};

struct PhaseTime {
    double mean_ms = 0.0;
    double std_ms = 0.0;  // sample standard deviation (n - 1); zero for a single run
};

struct BenchResult {
    PhaseTime h2d;
    PhaseTime kernel;
    PhaseTime d2h;
};

class Bench {
public:
    explicit Bench(BenchConfig config) : config_(config) {}

    void run(const std::function<void()>& upload, const std::function<void()>& compute,
             const std::function<void()>& download);

    const BenchConfig& config() const { return config_; }
    const BenchResult& result() const { return result_; }

private:
    BenchConfig config_;
    BenchResult result_;
};
