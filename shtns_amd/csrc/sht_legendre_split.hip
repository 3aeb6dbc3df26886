// sht_legendre_split.hip -- Legendre kernels with the contraction as three FP32 matrix-core
// products of hi/lo operand pairs (recurrence, seeds and final sums in FP64; see Contract<2>
// in sht_legendre.h).
#include "sht_legendre.h"

namespace shtns_amd {

void launch_legendre_synth_tiles_split(const LegendreArgs &a, int spec_ltop, double2 *F, long four_dist,
                                       const double *qmix, int nsteps_total, int tile0, int ntl, hipStream_t s) {
    synth_tiles_launch<kPrecSplit>(a, spec_ltop, F, four_dist, qmix, nsteps_total, tile0, ntl, s);
}

void launch_legendre_anal_split(const LegendreArgs &a, const double2 *F, long four_dist, double2 *Qlm,
                                long spec_dist, int spec_ltop, hipStream_t s) {
    anal_launch<kPrecSplit>(a, F, four_dist, Qlm, spec_dist, spec_ltop, s);
}

}  // namespace shtns_amd
