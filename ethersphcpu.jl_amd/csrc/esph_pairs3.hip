// esph_pairs3.hip — 3D instances of the pair kernels and launchers.
#include "esph_pairs.hpp"

namespace esph {

template hipError_t launch_continuity_d<3>(const Bucket& b, hipStream_t st);
template hipError_t launch_forces_d<3>(const Bucket& b, hipStream_t st);
template hipError_t launch_shepard_d<3>(const Bucket& b, hipStream_t st);

}  // namespace esph
