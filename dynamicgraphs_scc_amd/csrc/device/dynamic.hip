// Dynamic edge updates (see scc_update in include/sccmi.h).
#include "common.h"
#include "engine.h"

namespace sccmi {
void destroy_dynamic(scc_ctx* c) { (void)c; }
void destroy_sched(scc_ctx* c) { (void)c; }
}  // namespace sccmi

extern "C" int scc_update(scc_ctx* c, const int32_t*, const int32_t*, int64_t, const int32_t*,
                          const int32_t*, int64_t, scc_update_stats*) {
  return sccmi::set_error(c, SCC_ERR_STATE, "not implemented");
}
