// FP64 mode: built with -ffp-contract=off so every product and sum rounds
// exactly where the reference FP64 solver rounds.
#include "swe_launch_impl.h"

namespace braids {
template void run_stage<double>(const MeshView<double> &, StateView<double>, StateView<double>, StateView<double>,
                                const double *, Work<double>, Control *, double *, int, double, double,
                                const LaunchCfg &, hipStream_t);
template void run_rates<double>(const MeshView<double> &, StateView<double>, const double *, Work<double>, Control *,
                                double *, double *, double, const LaunchCfg &, hipStream_t);
}  // namespace braids
