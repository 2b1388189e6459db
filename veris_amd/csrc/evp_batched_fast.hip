// FP32-storage instantiations of the batched EVP substep kernel
// (evp_kernel.h), built with -ffp-contract=fast so multiply-adds fuse:
//  - evp_batched_launch_f32: FP32 arithmetic, FP64 Delta, zeta and
//    replacement pressure;
//  - evp_batched_launch_f32_f64: FP64 arithmetic on FP32-stored fields (the
//    HBM traffic of FP32, the rounding of FP64 inside each substep);
//  - evp_batched_launch_f32_s64: FP32-stored step constants around an FP64
//    state (u, v, sigma and the previous-step velocities) with FP64
//    arithmetic: the only rounding is the one of the constant inputs.
#include "evp_kernel.h"

namespace vh {

void evp_batched_launch_f32(int cfg, const EvpArgBlock<float>* ab, int which, int launch,
                            int n_tiles, bool so, bool plain, hipStream_t st) {
  evp_launch_cfg<float, double, true>(cfg, ab, which, launch, n_tiles, 0, st, so, plain);
}

void evp_batched_launch_f32_f64(int cfg, const EvpArgBlock<float>* ab, int which, int launch,
                                int n_tiles, bool so, bool plain, hipStream_t st) {
  evp_launch_cfg<float, double, true, double>(cfg, ab, which, launch, n_tiles, 0, st, so, plain);
}

void evp_batched_launch_f32_s64(int cfg, const EvpArgBlock<float, double>* ab, int which,
                                int launch, int n_tiles, bool so, bool plain, hipStream_t st) {
  evp_launch_cfg<float, double, true, double, double>(cfg, ab, which, launch, n_tiles, 0, st, so, plain);
}

}  // namespace vh
