// The host side of the Hirschberg traceback (hirschberg.hip), as the batch code calls it.
// Internal to libnwgrad_hip: no part of the C ABI.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

namespace nwg_hb {

// One pass over the batch's split pairs: Global affine, untracked, hard.  Device pointers are
// the batch's own buffers; offsets are its host copies.
struct Job {
    hipStream_t stream = nullptr;
    int precision = 0;        // 0 float32, 1 float64
    int pmax = 0;             // 1: the prefix-max carry (traceback "hirschberg_pmax")
    int cutoff = 512;         // base-case rows
    int na = 0;
    const uint8_t* codes_a = nullptr;
    const uint8_t* codes_b = nullptr;
    const void* sub = nullptr;   // na * na, float or double as `precision`
    double go_a = 0, ge_a = 0, go_b = 0, ge_b = 0;
    double* scores = nullptr;
    double* grads = nullptr;
    int32_t* guide = nullptr;
    uint32_t* ops = nullptr;     // null: no stored paths
    int32_t* ops_n = nullptr;
    int32_t* ends = nullptr;
    const std::vector<int32_t>* pairs = nullptr;   // batch indices of the pairs that split
    const std::vector<int64_t>* off_a = nullptr;   // host offsets, N + 1
    const std::vector<int64_t>* off_b = nullptr;
    int64_t table_budget = int64_t(16) << 30;      // base-case codes of one launch, at most
};

// Measurements of the last pass.
struct Stats {
    int levels = 0;
    int64_t sweep_items = 0, nodes = 0, base_blocks = 0;
    int64_t sweep_cells = 0, base_cells = 0;
    int64_t peak_bytes = 0;   // device scratch the pass held
};

struct Work;                  // device buffers, grown and kept between passes
Work* work_new();
void work_free(Work* w);
const Stats& work_stats(const Work* w);

// Runs the pass; returns 0, or an NWG_* status with `err` set.  Synchronises `stream` once per
// recursion level (the host plans each level's panels from the previous level's joins).
int run(Work* w, const Job& job, std::string& err);

}  // namespace nwg_hb
