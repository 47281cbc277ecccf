// Instantiations of the local labelling kernel for the cont mode.
#include "local.hpp"

namespace cc3d_amd {
hipError_t launch_local_cont(int dt, int nb, int qk, bool tall, const LocalParams& p, hipStream_t s) {
  return launch_local_mode<M_CONT>(dt, nb, qk, tall, p, s);
}
hipError_t launch_local_batch_cont(int dt, int nb, bool tall, const LocalParams& p, const BatchVol* vols,
                                   const uint32_t* vix, int64_t gtile0, int64_t gtiles, hipStream_t s) {
  return launch_local_batch_mode<M_CONT>(dt, nb, tall, p, vols, vix, gtile0, gtiles, s);
}
hipError_t launch_local_batch_plane_cont(int dt, int nb, const LocalParams& p, const BatchVol* vols,
                                         const uint32_t* vix, int64_t gtile0, int64_t gtiles, hipStream_t s) {
  return launch_local_batch_plane_mode<M_CONT>(dt, nb, p, vols, vix, gtile0, gtiles, s);
}
}  // namespace cc3d_amd
