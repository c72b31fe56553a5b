// Node-level path cuts: the record-writing walks for chains of 3 specular vertices.
#include "pathcuts_kernels.hpp"

namespace palz {
namespace cuts {

template void launch_walk<3, kWrite>(const CutsArgs&, const CutEntry*, uint32_t, uint32_t*, hipStream_t);
template void launch_walk<3, kEmit>(const CutsArgs&, const CutEntry*, uint32_t, uint32_t*, hipStream_t);

}  // namespace cuts
}  // namespace palz
