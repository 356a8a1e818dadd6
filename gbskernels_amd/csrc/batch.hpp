// batch.hpp -- internal declarations shared by batch.hip and the server.
#pragma once

namespace gbsk {
// canonical flops of one evaluation: the cost model of bucket ordering and of
// the server's admission
double canonical_flops(int func, int tier, int dim);
}  // namespace gbsk
