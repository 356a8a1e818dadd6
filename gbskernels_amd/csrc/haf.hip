#include "kernels.hpp"
namespace gbsk {
int perm_cert_max_dim(int) { return 0; }
size_t haf_workspace_bytes(bool, int, int, int) { return 0; }
int haf_launch(bool, int, int, const cplx*, int, cplx*, double*, void*, hipStream_t) { return GBSK_ERR_FUNC; }
int haf_max_dim(bool, int) { return 0; }
size_t tor_workspace_bytes(int, int, int) { return 0; }
int tor_launch(int, int, const cplx*, int, cplx*, double*, void*, hipStream_t) { return GBSK_ERR_FUNC; }
int tor_max_dim(int) { return 0; }
}
