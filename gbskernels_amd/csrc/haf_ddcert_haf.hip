// haf_ddcert_haf.hip -- instantiation unit: run_haf_multi<ArDD, LOOP=false, HM_DDCERT>
#include "haf_impl.hpp"

namespace gbsk {
template int run_haf_multi<ArDD, false, HM_DDCERT>(const HafJob*, int, hipStream_t);
}  // namespace gbsk
