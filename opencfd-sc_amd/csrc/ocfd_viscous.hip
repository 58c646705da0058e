// ocfd_viscous.hip -- launchers of the viscous right-hand side (du_viscous,
// OCFD_viscous.f90:29-149) and of the fused stage update; the FP64
// instantiations of the kernels in ocfd_visc.h.  The mixed-precision form's
// FP32 instantiations live in ocfd_viscous_mixed.hip.
#include "ocfd_visc.h"
#include "ocfd_visc_fused.h"

namespace ocfd {

hipError_t launch_visc_flux_mixed(const Block &B, hipStream_t st);
hipError_t launch_visc_div_mixed(const Block &B, hipStream_t st);
hipError_t launch_visc_div_update_mixed(const Block &B, int krk, int refresh_planes, int *neg, hipStream_t st);
hipError_t launch_visc_flux_batch_mixed(const BoxSet &S, bool cd8, int ntiles, hipStream_t st);
hipError_t launch_visc_div_update_batch_mixed(const BoxSet &S, bool visc, bool cd8, int krk, int *neg, int ntiles,
                                              hipStream_t st);

hipError_t launch_visc_flux(const Block &B, hipStream_t st) {
  if (!prims_match(B)) return hipErrorInvalidValue;
  if (B.prec == OCFD_PREC_MIXED_CODE) return launch_visc_flux_mixed(B, st);
  return visc_flux_as<double>(B, st);
}

hipError_t launch_visc_div(const Block &B, hipStream_t st) {
  if (B.prec == OCFD_PREC_MIXED_CODE) return launch_visc_div_mixed(B, st);
  return visc_div_as<double>(B, st);
}

hipError_t launch_visc_div_update(const Block &B, int krk, int refresh_planes, int *neg, hipStream_t st) {
  if (!prims_match(B)) return hipErrorInvalidValue;
  if (B.prec == OCFD_PREC_MIXED_CODE) return launch_visc_div_update_mixed(B, krk, refresh_planes, neg, st);
  return visc_div_update_as<double>(B, krk, refresh_planes, neg, st);
}

hipError_t launch_visc_flux_batch(const BoxSet &S, int prec, bool cd8, int ntiles, hipStream_t st) {
  if (prec == OCFD_PREC_MIXED_CODE) return launch_visc_flux_batch_mixed(S, cd8, ntiles, st);
  return visc_flux_batch_as<double>(S, cd8, ntiles, st);
}

hipError_t launch_visc_div_update_batch(const BoxSet &S, int prec, bool visc, bool cd8, int krk, int *neg,
                                        int ntiles, hipStream_t st) {
  if (prec == OCFD_PREC_MIXED_CODE) return launch_visc_div_update_batch_mixed(S, visc, cd8, krk, neg, ntiles, st);
  return visc_div_update_batch_as<double>(S, visc, cd8, krk, neg, ntiles, st);
}

hipError_t launch_visc_fused_mixed(const Block &B, int krk, int refresh, int *neg, hipStream_t st);

// periodic blocks with uniform metrics and CD6, whole 32 x 8 tiles, at least
// 16 planes; the FP32 instance only (the mixed-precision form)
bool visc_fused_supported(const Block &B) {
  return B.prec == OCFD_PREC_MIXED_CODE && all_periodic(B) && B.S.scheme_vis != SCH_CD8 && B.g.nx % FTX == 0 &&
         B.g.ny % FTY == 0 && B.g.nz >= 16;
}

hipError_t launch_visc_fused(const Block &B, int krk, int refresh, int *neg, hipStream_t st) {
  if (!visc_fused_supported(B) || !prims_match(B)) return hipErrorInvalidValue;
  return launch_visc_fused_mixed(B, krk, refresh, neg, st);
}

int64_t march_tiles(const Geom &g, int kseg) { return march_blocks(g, kseg < g.nz ? kseg : g.nz); }

}  // namespace ocfd
