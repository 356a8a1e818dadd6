// haf_cert_lhaf.hip -- instantiation unit: run_haf<ArF64, LOOP=true, HM_CERT>
#include "haf_impl.hpp"

namespace gbsk {
template int run_haf<ArF64, true, HM_CERT>(int, const cplx*, int, cplx*, double*, void*, hipStream_t);
}  // namespace gbsk
