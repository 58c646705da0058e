// ocfd_viscous_mixed.hip -- FP32 instantiations of the viscous kernels for
// the mixed-precision form (Ev, the derivatives, stresses and fluxes in FP32;
// du accumulated in FP32; the stage combination and primitives FP64).
#include "ocfd_visc.h"
#include "ocfd_visc_fused.h"

namespace ocfd {

hipError_t launch_visc_flux_mixed(const Block &B, hipStream_t st) { return visc_flux_as<float>(B, st); }

hipError_t launch_visc_div_mixed(const Block &B, hipStream_t st) { return visc_div_as<float>(B, st); }

hipError_t launch_visc_div_update_mixed(const Block &B, int krk, int refresh_planes, int *neg, hipStream_t st) {
  return visc_div_update_as<float>(B, krk, refresh_planes, neg, st);
}

hipError_t launch_visc_fused_mixed(const Block &B, int krk, int refresh, int *neg, hipStream_t st) {
  return visc_fused_as<float>(B, krk, refresh, neg, st);
}

hipError_t launch_visc_fused_batch(const BoxSet &S, int krk, int *neg, int ntiles, hipStream_t st) {
  return visc_fused_batch_as<float>(S, krk, neg, ntiles, st);
}

int64_t fused_batch_tiles(const Geom &g, int kseg) { return fused_blocks(g, fused_batch_kseg(g.nz, kseg)); }

hipError_t launch_visc_flux_batch_mixed(const BoxSet &S, bool cd8, int ntiles, hipStream_t st) {
  return visc_flux_batch_as<float>(S, cd8, ntiles, st);
}

hipError_t launch_visc_div_update_batch_mixed(const BoxSet &S, bool visc, bool cd8, int krk, int *neg, int ntiles,
                                              hipStream_t st) {
  return visc_div_update_batch_as<float>(S, visc, cd8, krk, neg, ntiles, st);
}

}  // namespace ocfd
