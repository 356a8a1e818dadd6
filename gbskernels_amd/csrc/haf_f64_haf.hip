// haf_f64_haf.hip -- instantiation unit: run_haf<ArF64, LOOP=false, KAPPA=false>
#include "haf_impl.hpp"

namespace gbsk {
template int run_haf<ArF64, false, false>(int, const cplx*, int, cplx*, double*, void*, hipStream_t);
}  // namespace gbsk
