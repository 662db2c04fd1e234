// Variant table of the batched Cholesky factorization, per element type and size bin.
// Written by elemental_amd/autotune_batched.py from timings on one MI355X (see that script for
// the workloads and the candidates); edit by running the script again, not by hand.
// Measured 2026-12-03, chosen variants (median factor time of the bin's workload):
//   float64 n in [16, 128]: kind 0 nb 32 occ 3 0.300 ms
//   float64 n in [129, 256]: kind 2 nb 64 occ 2 1.225 ms
//   float64 n in [257, 512]: kind 2 nb 64 occ 2 4.035 ms
//   float32 n in [16, 128]: kind 0 nb 32 occ 3 0.284 ms
//   float32 n in [129, 256]: kind 2 nb 64 occ 2 0.946 ms
//   float32 n in [257, 512]: kind 2 nb 64 occ 2 2.839 ms
#pragma once

namespace elx {
namespace tuned {

struct BatchedVariant {
    int nb;    // block-column width of the left-looking factorization (16 or 32; 64 for kind 2)
    int occ;   // persistent workgroups per CU (at most what the bin's LDS and registers allow)
    int kind;  // 0: block column staged in LDS; 1: operands read straight into registers;
               // 2: 64-column panels, diagonal block factored and inverted in LDS
};

// [0: double, 1: float][bin 0: n <= 128, 1: n <= 256, 2: n <= 512]
constexpr BatchedVariant kBatchedPotrf[2][3] = {
    {{32, 3, 0}, {64, 2, 2}, {64, 2, 2}},
    {{32, 3, 0}, {64, 2, 2}, {64, 2, 2}},
};

}  // namespace tuned
}  // namespace elx
