// Instantiations of the row labelling kernel for the binary mode.
#include "rows.hpp"

namespace cc3d_amd {
hipError_t launch_rows_binary(int dt, int nb, const LocalParams& p, hipStream_t s) {
  return launch_rows_mode<M_BINARY>(dt, nb, p, s);
}
hipError_t launch_rows_batch_binary(int dt, int nb, const LocalParams& p, const BatchVol* vols, const uint32_t* vix,
                                    int64_t gtile0, int64_t gtiles, hipStream_t s) {
  return launch_rows_batch_mode<M_BINARY>(dt, nb, p, vols, vix, gtile0, gtiles, s);
}
}  // namespace cc3d_amd
