// Internal helpers shared by the library's translation units.
#pragma once
#include <stddef.h>

namespace bsofi {

// Flop counts of include/bsofi_flops.h:19-30 (LAWN-41 formulas of
// include/bench/flops.h), as doubles.
double flops_bsof(int n, int L);
double flops_bstri(int n, int L);
double flops_bsoi(int n, int L);

}  // namespace bsofi
