// haf_f64_haf.hip -- instantiation unit: run_haf_multi<ArF64, LOOP=false, HM_PLAIN>
#include "haf_impl.hpp"

namespace gbsk {
template int run_haf_multi<ArF64, false, HM_PLAIN>(const HafJob*, int, hipStream_t);
}  // namespace gbsk
