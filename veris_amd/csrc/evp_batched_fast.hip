// FP32 instantiation of the batched EVP substep kernel (evp_kernel.h): FP32
// bulk with FP64 Delta, zeta and replacement pressure.  Built with
// -ffp-contract=fast so multiply-adds fuse.
#include "evp_kernel.h"

namespace vh {

void evp_batched_launch_f32(int cfg, const EvpArgBlock<float>* ab, int which, int launch,
                            int n_tiles, hipStream_t st) {
  evp_launch_cfg<float, double, true>(cfg, ab, which, launch, n_tiles, 0, st);
}

}  // namespace vh
