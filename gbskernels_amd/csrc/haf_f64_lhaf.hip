// haf_f64_lhaf.hip -- instantiation unit: run_haf<ArF64, LOOP=true, KAPPA=false>
#include "haf_impl.hpp"

namespace gbsk {
template int run_haf<ArF64, true, false>(int, const cplx*, int, cplx*, double*, void*, hipStream_t);
}  // namespace gbsk
