// tune.h -- the autotuner (tune.cpp).
#pragma once

#include <cstddef>
#include <string>

#include "engine.h"

namespace sfft {

struct TuneResult {
    PlanHints hints;
    double ms = 0;     // mean of forward and inverse, per launch
    double gbs = 0;    // algorithmic bytes (input + output) / ms
    size_t batch = 0;  // transforms per timed launch
    int candidates = 0;
};

// Times every candidate of plan_candidates() over `batch` transforms (0:
// about 1 GB per launch), forward and inverse, `reps` launches each; the
// fastest becomes the plan get_plan() builds for (n, kind, prec).
TuneResult autotune(size_t n, int kind, int prec, size_t batch, int reps);
void tune_clear();
size_t tune_count();
void tune_save(const std::string &path);
// Returns the number of entries taken; a file tuned on another GPU
// architecture or HIP runtime is ignored (0) when check_device is set.
size_t tune_load(const std::string &path, bool check_device);

}  // namespace sfft
