// haf_ddcert_haf.hip -- instantiation unit: run_haf<ArDD, LOOP=false, HM_DDCERT>
#include "haf_impl.hpp"

namespace gbsk {
template int run_haf<ArDD, false, HM_DDCERT>(int, const cplx*, int, cplx*, double*, void*, hipStream_t);
}  // namespace gbsk
