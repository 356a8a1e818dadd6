// haf_kappa_lhaf.hip -- instantiation unit: run_haf_multi<ArF64, LOOP=true, HM_KAPPA>
#include "haf_impl.hpp"

namespace gbsk {
template int run_haf_multi<ArF64, true, HM_KAPPA>(const HafJob*, int, hipStream_t);
}  // namespace gbsk
