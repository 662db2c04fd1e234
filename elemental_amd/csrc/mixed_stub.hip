// placeholder
#include "elx_common.hpp"
extern "C" {
elx_status_t elx_dgesv_mixed(void*, int64_t, int64_t, const double*, int64_t, const double*, int64_t, double*, int64_t, int, int, double, int32_t, elx_ir_stats_t*) { return ELX_ERROR_NOT_SUPPORTED; }
elx_status_t elx_dposv_mixed(void*, int64_t, int64_t, const double*, int64_t, const double*, int64_t, double*, int64_t, int, int, double, int32_t, elx_ir_stats_t*) { return ELX_ERROR_NOT_SUPPORTED; }
}
