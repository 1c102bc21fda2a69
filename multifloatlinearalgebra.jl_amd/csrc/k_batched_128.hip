// Batched factor + solve: the 128-thread workgroup shape (its own
// translation unit; see k_batched_body.hpp).
#include "k_batched_body.hpp"

namespace mfk {

hipError_t batched_launch_128(int N, hipStream_t s, const BatchArgs &a) { return launch_bt<128>(N, s, a); }

}  // namespace mfk
