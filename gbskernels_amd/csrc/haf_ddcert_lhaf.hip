// haf_ddcert_lhaf.hip -- instantiation unit: run_haf<ArDD, LOOP=true, HM_DDCERT>
#include "haf_impl.hpp"

namespace gbsk {
template int run_haf<ArDD, true, HM_DDCERT>(int, const cplx*, int, cplx*, double*, void*, hipStream_t);
}  // namespace gbsk
