// Instantiations of the row labelling kernel for the multilabel mode.
#include "rows.hpp"

namespace cc3d_amd {
hipError_t launch_rows_multi(int dt, int nb, const LocalParams& p, hipStream_t s) {
  return launch_rows_mode<M_MULTI>(dt, nb, p, s);
}
hipError_t launch_rows_batch_multi(int dt, int nb, const LocalParams& p, const BatchVol* vols, const uint32_t* vix,
                                   int64_t gtile0, int64_t gtiles, hipStream_t s) {
  return launch_rows_batch_mode<M_MULTI>(dt, nb, p, vols, vix, gtile0, gtiles, s);
}
}  // namespace cc3d_amd
