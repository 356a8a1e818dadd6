// haf_kappa_lhaf.hip -- instantiation unit: run_haf<ArF64, LOOP=true, HM_KAPPA>
#include "haf_impl.hpp"

namespace gbsk {
template int run_haf<ArF64, true, HM_KAPPA>(int, const cplx*, int, cplx*, double*, void*, hipStream_t);
}  // namespace gbsk
