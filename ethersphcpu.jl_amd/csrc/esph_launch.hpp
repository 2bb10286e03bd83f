// esph_launch.hpp — host-side launchers of the kernels in esph_kernels.hip.
#pragma once

#include <hip/hip_runtime.h>

#include "esph_device.hpp"

namespace esph {

// Scenes of one (dim, kernel, precision) class; the pair kernels are
// compiled per class, so a batch launches one kernel per bucket and phase.
struct Bucket {
    int dim;
    int kernel;
    int prec;
    const SceneDev* scenes;  // device table
    const Tile* tiles;       // device: particle tiles of the bucket's scenes
    int ntiles;
    const Tile* ovf_tiles;   // device: kOvfTiles tiles per scene over SceneDev::ovf
    int nscenes;
    bool has_thermal;        // some scene of the bucket is thermal
    bool has_plain;          // some scene of the bucket is neither thermal nor FLAG_EOS_WALK
    bool has_eos = false;    // some scene of the bucket has FLAG_EOS_WALK but not FLAG_BRICK
    bool has_brick = false;  // some scene of the bucket has FLAG_BRICK (brick walks)
    bool has_tile = true;    // some scene of the bucket has not (tile walks)
    const Tile* btiles = nullptr;  // device: one entry (scene, 0) per FLAG_BRICK scene
    int nbtiles = 0;
    int depth = 2;           // gather depth of the force pass (1, 2 or 4; see the tuner)
    int bs = 256;            // workgroup size of the force pass (64, 128 or 256; see the tuner)
    int cdepth = 2, cbs = 256;  // the same for the continuity pass
    int sdepth = 2, sbs = 256;  // and for the Shepard pass
    // tuner runs: the main instances only, the continuity pass dry (its
    // result to SceneDev::rhon, the state untouched), the Shepard pass
    // whether or not a re-initialisation is due (its result is rhon anyway)
    bool tuning = false;
};

// Tables for the dimension-agnostic sort kernels of a whole batch.
struct SortPlan {
    const SceneDev* scenes;
    int nscenes;
    const Tile* ptiles;  // particle tiles (capacity-sized)
    int nptiles;
    const Tile* ctiles;  // cell tiles of kCellTile cells
    int nctiles;
};

template <int D>
hipError_t launch_continuity_d(const Bucket& b, hipStream_t st);
template <int D>
hipError_t launch_forces_d(const Bucket& b, hipStream_t st);
template <int D>
hipError_t launch_shepard_d(const Bucket& b, hipStream_t st);
hipError_t launch_continuity(const Bucket& b, hipStream_t st);
hipError_t launch_forces(const Bucket& b, hipStream_t st);
hipError_t launch_shepard(const Bucket& b, hipStream_t st);
hipError_t launch_sort(const SortPlan& p, int mode, hipStream_t st);
hipError_t launch_pairs(const Bucket& b, int pairset, int32_t* counts, const int64_t* offs, int32_t* out,
                        hipStream_t st);
hipError_t launch_walk_pairs(const Bucket& b, int pairset, int32_t* counts, const int64_t* offs, int32_t* out,
                             hipStream_t st);
hipError_t launch_kernel_eval(int kernel, int64_t n, const double* r, double* w, double* dw, double h,
                              double sigma, hipStream_t st);
hipError_t launch_probe(const Bucket& b, int64_t np, const double* pts, int field, double none_value, double* out,
                        hipStream_t st);

}  // namespace esph
