// Per-tuple interval culling in front of the Newton solve.
#pragma once

#include "common.hpp"

namespace palz {

// admissible[p] = 0 when path p (chain length 1..3) provably cannot reach the
// solver's tolerance for any of its scene's wavelengths, else 1; paths of
// other lengths get 1.  `scenes` is indexed by the path's scene (a path naming
// a scene >= n_scenes gets 1); `n_dev` (optional, device) caps n.
// mode 1: the interval test; mode 2: the interval test and then, for paths
// it keeps, the affine test (a path is culled when either proves it has no
// solution).
void launch_cull_paths(const palz_scene_desc* scenes, uint32_t n_scenes, const palz_path* paths, uint32_t n,
                       const uint32_t* n_dev, uint8_t* admissible, hipStream_t stream, int mode = 1);

}  // namespace palz
