// haf_cert_haf.hip -- instantiation unit: run_haf_multi<ArF64, LOOP=false, HM_CERT>
#include "haf_impl.hpp"

namespace gbsk {
template int run_haf_multi<ArF64, false, HM_CERT>(const HafJob*, int, hipStream_t);
}  // namespace gbsk
