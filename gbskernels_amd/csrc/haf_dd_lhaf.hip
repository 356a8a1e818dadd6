// haf_dd_lhaf.hip -- instantiation unit: run_haf<ArDD, LOOP=true, KAPPA=false>
#include "haf_impl.hpp"

namespace gbsk {
template int run_haf<ArDD, true, false>(int, const cplx*, int, cplx*, double*, void*, hipStream_t);
}  // namespace gbsk
