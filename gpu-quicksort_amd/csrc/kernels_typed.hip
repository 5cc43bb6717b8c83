// kernels_typed.hip -- the key-type kernels of sortdev.h for one key type
// (GQS_DT: 0 uint32, 1 float, 2 double) and their launchers.  The build
// compiles this file once per key type.
#include "sortdev.h"

#ifndef GQS_DT
#error "compile once per key type with -DGQS_DT=0, 1 or 2"
#endif

namespace gqs {

template <int DT>
hipError_t typed_level_step(int step, const LevelArgs& a, const Grid& g, hipStream_t st) {
    switch (step) {
        case 0: hipLaunchKernelGGL(k_splitters<DT>, dim3(g.splitters), dim3(kSplWG), 0, st, a); break;
        case 2: hipLaunchKernelGGL(k_count<DT>, dim3(g.part), dim3(kPartWG), 0, st, a); break;
        case 6:
            hipLaunchKernelGGL(k_scatter<DT>, dim3(std::max(1, g.part * kPartWG / kScatterWG)), dim3(kScatterWG), 0, st,
                               a);
            break;
        case 7: hipLaunchKernelGGL((k_tilepart<DT, false>), dim3(g.tilepart), dim3(kScatterWG), 0, st, a); break;
        case 10: hipLaunchKernelGGL((k_tilepart<DT, true>), dim3(g.tilepart), dim3(kScatterWG), 0, st, a); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

template <int DT>
hipError_t typed_leaf(const LeafArgs& a, const Grid& g, hipStream_t st) {
    hipLaunchKernelGGL((k_leaf_small<DT>), dim3(g.leaf), dim3(kLeafWG), 0, st, a);
    // gathered leaves (a tile-local level) are never sorted by one wave
    // the leaves of a tile-local level are claimed from a counter by as many
    // work-groups as are resident; others go by a static stride
    const bool claimed = a.list == nullptr && a.lv.tile_local;
    if (a.leaf_max > a.bitonic_cutoff || a.is_last || a.lv.tile_local) {
        LeafArgs b = a;
        b.next = claimed ? &a.lv.ctl->leaf_next : nullptr;
        hipLaunchKernelGGL((k_leaf_big<DT, 512>), dim3(claimed ? g.leaf_big : g.leaf), dim3(512), 0, st, b);
    }
    if (a.leaf_max > a.leaf_cap) {  // leaves above the 512-thread capacity: the large-leaf list
        LeafArgs h = a;
        h.list = a.huge;
        h.list_count = a.huge_count;
        h.list_cap = a.lv.huge_cap;
        h.next = claimed ? &a.lv.ctl->huge_next : nullptr;
        hipLaunchKernelGGL((k_leaf_big<DT, 1024>), dim3(g.huge), dim3(1024), 0, st, h);
    }
    return hipGetLastError();
}

template hipError_t typed_level_step<GQS_DT>(int, const LevelArgs&, const Grid&, hipStream_t);
template hipError_t typed_leaf<GQS_DT>(const LeafArgs&, const Grid&, hipStream_t);

}  // namespace gqs
