// haf_ddcert_lhaf.hip -- instantiation unit: run_haf_multi<ArDD, LOOP=true, HM_DDCERT>
#include "haf_impl.hpp"

namespace gbsk {
template int run_haf_multi<ArDD, true, HM_DDCERT>(const HafJob*, int, hipStream_t);
}  // namespace gbsk
