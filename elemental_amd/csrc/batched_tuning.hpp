// Variant table of the batched Cholesky factorization, per element type and size bin.
// Written by elemental_amd/autotune_batched.py from timings on one MI355X (see that script for
// the workloads and the candidates); edit by running the script again, not by hand.
// Measured 2026-12-03, chosen variants (median factor time of the bin's workload):
//   float64 n in [16, 128]: nb 32 occ 3 0.311 ms
//   float64 n in [129, 256]: nb 32 occ 2 1.723 ms
//   float64 n in [257, 512]: nb 16 occ 2 5.555 ms
//   float32 n in [16, 128]: nb 16 occ 3 0.298 ms
//   float32 n in [129, 256]: nb 32 occ 2 1.361 ms
//   float32 n in [257, 512]: nb 16 occ 2 4.903 ms
#pragma once

namespace elx {
namespace tuned {

struct BatchedVariant {
    int nb;   // block-column width of the left-looking factorization (16 or 32)
    int occ;  // persistent workgroups per CU (at most what the bin's LDS and registers allow)
};

// [0: double, 1: float][bin 0: n <= 128, 1: n <= 256, 2: n <= 512]
constexpr BatchedVariant kBatchedPotrf[2][3] = {
    {{32, 3}, {32, 2}, {16, 2}},
    {{16, 3}, {32, 2}, {16, 2}},
};

}  // namespace tuned
}  // namespace elx
