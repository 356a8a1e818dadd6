// haf_cert_lhaf.hip -- instantiation unit: run_haf_multi<ArF64, LOOP=true, HM_CERT>
#include "haf_impl.hpp"

namespace gbsk {
template int run_haf_multi<ArF64, true, HM_CERT>(const HafJob*, int, hipStream_t);
}  // namespace gbsk
