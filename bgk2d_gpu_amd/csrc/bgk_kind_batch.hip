// bgk_kind_batch.hip -- the batch kernels of bgk_serve.hpp, instantiated apart
// so the library compiles in parallel
#include "bgk_host.hpp"

namespace bgk_host {
template void launch_batch_s<float>(bgk_batch_t *, int, int, int, hipStream_t);
template void launch_batch_tb<float>(bgk_batch_t *, int, int, int, hipStream_t);
template int occupancy_tb<float>(int, int, int *);
template void launch_batch_s<_Float16>(bgk_batch_t *, int, int, int, hipStream_t);
template void launch_batch_tb<_Float16>(bgk_batch_t *, int, int, int, hipStream_t);
template int occupancy_tb<_Float16>(int, int, int *);
template void launch_batch_s<bgk::bf16_t>(bgk_batch_t *, int, int, int, hipStream_t);
template void launch_batch_tb<bgk::bf16_t>(bgk_batch_t *, int, int, int, hipStream_t);
template int occupancy_tb<bgk::bf16_t>(int, int, int *);
}  // namespace bgk_host
