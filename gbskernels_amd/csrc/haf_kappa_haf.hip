// haf_kappa_haf.hip -- instantiation unit: run_haf<ArF64, LOOP=false, KAPPA=true>
#include "haf_impl.hpp"

namespace gbsk {
template int run_haf<ArF64, false, true>(int, const cplx*, int, cplx*, double*, void*, hipStream_t);
}  // namespace gbsk
