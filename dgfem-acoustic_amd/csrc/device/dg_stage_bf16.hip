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
    const dim3 g(grid), b(Bf16Shape<DIM, P, NF>::kBlockSize);
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

// Tetrahedra of order 2..5 (the orders the matrix-core form pays for: the FP32 contractions
// of lower orders are a small part of a stage), 4 or 5 fields, 2 or 3 BF16 terms.
template <typename Fn>
bool dispatch(int dim, int p, int nf, int nsplit, Fn&& fn) {
    if (dim != 3 || (nf != 4 && nf != 5) || (nsplit != 2 && nsplit != 3)) return false;
    auto bySplit = [&](auto pc, auto nfc) {
        constexpr int P = decltype(pc)::value, NF = decltype(nfc)::value;
        if (nsplit == 3) {
            if constexpr (bf16Supported<3, P, NF, 3>()) return fn(std::integral_constant<int, P>{}, nfc,
                                                                  std::integral_constant<int, 3>{});
        } else {
            if constexpr (bf16Supported<3, P, NF, 2>()) return fn(std::integral_constant<int, P>{}, nfc,
                                                                  std::integral_constant<int, 2>{});
        }
        return false;
    };
    auto byNf = [&](auto pc) {
        return nf == 4 ? bySplit(pc, std::integral_constant<int, 4>{}) : bySplit(pc, std::integral_constant<int, 5>{});
    };
    switch (p) {
        case 2: return byNf(std::integral_constant<int, 2>{});
        case 3: return byNf(std::integral_constant<int, 3>{});
        case 4: return byNf(std::integral_constant<int, 4>{});
        case 5: return byNf(std::integral_constant<int, 5>{});
        default: return false;
    }
}

}  // namespace

bool stageBf16Built(int dim, int p, int nf, int nsplit) {
    return dispatch(dim, p, nf, nsplit, [](auto, auto, auto) { return true; });
}

int stageBf16BlocksPerCu(int dim, int p, int nf, int nsplit) {
    int b = 0;
    dispatch(dim, p, nf, nsplit, [&](auto pc, auto nfc, auto sc) {
        b = bf16BlocksPerCu<3, decltype(pc)::value, decltype(nfc)::value, decltype(sc)::value>();
        return true;
    });
    return b;
}

bool launchStageBf16(int dim, int p, int nf, int nsplit, int mode, const StageArgs<float>& a, const void* afrag,
                     int64_t nGroups, unsigned grid, hipStream_t stream) {
    return dispatch(dim, p, nf, nsplit, [&](auto pc, auto nfc, auto sc) {
        launchShape<3, decltype(pc)::value, decltype(nfc)::value, decltype(sc)::value>(mode, a, afrag, nGroups, grid,
                                                                                      stream);
        return true;
    });
}

}  // namespace dgk
