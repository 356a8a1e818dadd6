// haf_kappa_haf.hip -- instantiation unit: run_haf_multi<ArF64, LOOP=false, HM_KAPPA>
#include "haf_impl.hpp"

namespace gbsk {
template int run_haf_multi<ArF64, false, HM_KAPPA>(const HafJob*, int, hipStream_t);
}  // namespace gbsk
