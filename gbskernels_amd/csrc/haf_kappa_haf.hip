// haf_kappa_haf.hip -- instantiation unit: run_haf<ArF64, LOOP=false, HM_KAPPA>
#include "haf_impl.hpp"

namespace gbsk {
template int run_haf<ArF64, false, HM_KAPPA>(int, const cplx*, int, cplx*, double*, void*, hipStream_t);
}  // namespace gbsk
