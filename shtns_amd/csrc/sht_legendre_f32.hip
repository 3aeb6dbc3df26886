// sht_legendre_f32.hip -- Legendre kernels with the contraction on the FP32 matrix cores
// (recurrence, seeds and final sums in FP64; see Contract<1> in sht_legendre.h).
#include "sht_legendre.h"

namespace shtns_amd {

void launch_legendre_synth_tiles_f32(const LegendreArgs &a, int spec_ltop, double2 *F, long four_dist,
                                     const double *qmix, int nsteps_total, int tile0, int ntl, hipStream_t s) {
    synth_tiles_launch<kPrecF32>(a, spec_ltop, F, four_dist, qmix, nsteps_total, tile0, ntl, s);
}

void launch_legendre_anal_f32(const LegendreArgs &a, const double2 *F, long four_dist, double2 *Qlm, long spec_dist,
                              int spec_ltop, hipStream_t s) {
    anal_launch<kPrecF32>(a, F, four_dist, Qlm, spec_dist, spec_ltop, s);
}

}  // namespace shtns_amd
