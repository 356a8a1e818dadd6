// haf_f64_haf.hip -- instantiation unit: run_haf<ArF64, LOOP=false, HM_PLAIN>
#include "haf_impl.hpp"

namespace gbsk {
template int run_haf<ArF64, false, HM_PLAIN>(int, const cplx*, int, cplx*, double*, void*, hipStream_t);
}  // namespace gbsk
