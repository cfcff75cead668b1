// bgk_kind_bf16.hip -- the step loop (every hot kernel) of one storage kind,
// instantiated apart so the library compiles in parallel
#include "bgk_host.hpp"

namespace bgk_host {
template int run_t<bgk::bf16_t, float>(bgk_ctx *, int);
template int enqueue_t<bgk::bf16_t, float>(bgk_ctx *, int);
}  // namespace bgk_host
