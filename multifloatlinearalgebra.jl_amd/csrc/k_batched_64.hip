// Batched factor + solve: the 64-thread workgroup shape (its own
// translation unit; see k_batched_body.hpp).
#include "k_batched_body.hpp"

namespace mfk {

hipError_t batched_launch_64(int N, hipStream_t s, const BatchArgs &a) { return launch_bt<64>(N, s, a); }

}  // namespace mfk
