// Instances and launch of the split-BF16 matrix-core stage kernel (mixed precision), kept in a
// translation unit of their own.
#include <hip/hip_runtime.h>

#include "dg_kernels_bf16.hpp"

namespace dgk {
namespace {

template <int DIM, int P, int NF, int NSPLIT>
void launchShape(int mode, const StageArgs<float>& a, const void* afrag, int64_t nGroups, unsigned grid,
                 hipStream_t st) {
    const auto* af = static_cast<const u32x4_t*>(afrag);
    const dim3 g(grid), b(Bf16Shape<DIM, P>::kBlockSize);
    switch (mode) {
        case kModeFirst:
            hipLaunchKernelGGL((stageKernelBf16<DIM, P, NF, NSPLIT, kModeFirst>), g, b, 0, st, a, af, nGroups);
            break;
        case kModeMid:
            hipLaunchKernelGGL((stageKernelBf16<DIM, P, NF, NSPLIT, kModeMid>), g, b, 0, st, a, af, nGroups);
            break;
        case kModeFinal:
            hipLaunchKernelGGL((stageKernelBf16<DIM, P, NF, NSPLIT, kModeFinal>), g, b, 0, st, a, af, nGroups);
            break;
        case kModeEuler:
            hipLaunchKernelGGL((stageKernelBf16<DIM, P, NF, NSPLIT, kModeEuler>), g, b, 0, st, a, af, nGroups);
            break;
        default:
            hipLaunchKernelGGL((stageKernelBf16<DIM, P, NF, NSPLIT, kModeRHS>), g, b, 0, st, a, af, nGroups);
            break;
    }
}

template <int DIM, int P, int NF>
bool launchSplit(int nsplit, int mode, const StageArgs<float>& a, const void* afrag, int64_t nGroups, unsigned grid,
                 hipStream_t st) {
    if (nsplit == 3) {
        if constexpr (bf16Supported<DIM, P, 3>()) {
            launchShape<DIM, P, NF, 3>(mode, a, afrag, nGroups, grid, st);
            return true;
        }
    } else if (nsplit == 2) {
        if constexpr (bf16Supported<DIM, P, 2>()) {
            launchShape<DIM, P, NF, 2>(mode, a, afrag, nGroups, grid, st);
            return true;
        }
    }
    return false;
}

// Tetrahedra of order 2..5 (the orders the matrix-core form pays for: the FP32 contractions
// of lower orders are a small part of a stage).
template <int NF>
bool launchP(int p, int nsplit, int mode, const StageArgs<float>& a, const void* afrag, int64_t nGroups,
             unsigned grid, hipStream_t st) {
    switch (p) {
        case 2: return launchSplit<3, 2, NF>(nsplit, mode, a, afrag, nGroups, grid, st);
        case 3: return launchSplit<3, 3, NF>(nsplit, mode, a, afrag, nGroups, grid, st);
        case 4: return launchSplit<3, 4, NF>(nsplit, mode, a, afrag, nGroups, grid, st);
        case 5: return launchSplit<3, 5, NF>(nsplit, mode, a, afrag, nGroups, grid, st);
        default: return false;
    }
}

}  // namespace

bool stageBf16Built(int dim, int p, int nf, int nsplit) {
    return dim == 3 && p >= 2 && p <= 5 && (nf == 4 || nf == 5) && (nsplit == 2 || nsplit == 3);
}

bool launchStageBf16(int dim, int p, int nf, int nsplit, int mode, const StageArgs<float>& a, const void* afrag,
                     int64_t nGroups, unsigned grid, hipStream_t stream) {
    if (!stageBf16Built(dim, p, nf, nsplit)) return false;
    return nf == 4 ? launchP<4>(p, nsplit, mode, a, afrag, nGroups, grid, stream)
                   : launchP<5>(p, nsplit, mode, a, afrag, nGroups, grid, stream);
}

}  // namespace dgk
