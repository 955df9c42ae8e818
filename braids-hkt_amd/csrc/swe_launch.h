// Host-side launch sequences for one precision.  Each precision is compiled
// in its own translation unit (swe_f64.hip: no FMA contraction, so FP64 mode
// follows the reference operation order; swe_f32.hip: contraction allowed).
#pragma once
#include <hip/hip_runtime.h>
#include "swe_types.h"

namespace braids {

// Kernel kinds timed by the optional event profiler.
enum : int { KK_PREP = 0, KK_TRACES, KK_FACE, KK_DRAIN, KK_GATHER, KK_LEDGER,
             KK_RDRAG, KK_RTHETA, KK_TURBINE, KK_RUPDATE, KK_COUNT };

struct LaunchCfg {
    int block = 256;    // threads per workgroup (k_prep/k_face/k_drain/k_gather)
    int remap = 1;      // XCD-aware block remap
    // When non-null: ev[(slot*KK_COUNT + kind)*2 + {0,1}] bracket each launch
    // whose kind bit is set in ev_mask.
    hipEvent_t *ev = nullptr;
    int ev_mask = 0;
};

// mode: 0,1,2 = SSPRK(3,3) stage index (combination + prepare [+ Manning]);
//       3 = plain Euler stage (no combination, no preparation);
// Launches prep, traces, face, drain, gather and ledger on `stream`.
template <class R>
void run_stage(const MeshView<R> &m, StateView<R> in, StateView<R> base, StateView<R> out,
               const double *z, Work<R> w, Control *ctl, double *ledger_stage, int mode,
               double dt, double t, const LaunchCfg &cfg, hipStream_t stream);

// Reduce the partial maxima of the stages into the Control block.
template <class R>
void run_finalize(Work<R> w, Control *ctl, hipStream_t stream);

// Spatial rate without draining or CFL rejection (bend_native_rates):
// rate = -sum (L/A)(F - 1/2 g h^2 n); water = fl(1,:).
template <class R>
void run_rates(const MeshView<R> &m, StateView<R> in, const double *z, Work<R> w, Control *ctl,
               double *rate, double *water, double t, const LaunchCfg &cfg, hipStream_t stream);

// River SSPRK(2,2) substep buffers.
struct RiverBuffers {
    double *theta = nullptr;            // [nc] inventory factor of the current stage
    double *wlim[2] = {nullptr, nullptr};   // [nf] limited water of open boundary faces, per stage
    double *ledger = nullptr;           // [nb] inward water of the substep
    DragView drag{0, nullptr};          // mode 0: no drag
};

// One river substep (RiverExplicit._single): base -> work (drag half-step)
// -> q1 -> out.  tv may be null (no rotors).  Stage slots 0 and 1 of ctl.
template <class R>
void run_river_substep(const MeshView<R> &m, StateView<R> base, StateView<R> work, StateView<R> q1,
                       StateView<R> out, const double *z, Work<R> w, Control *ctl, const RiverBuffers &rb,
                       const TurbineView *tv, double dt, double t, const LaunchCfg &cfg, hipStream_t stream);

// Footprint weights of (rotor, cell) pairs (device arrays).
void run_footprint(int npairs, const double *tri, const double *D, const double *nodes, const double *weights,
                   double *value, double *clipped, int *order, hipStream_t stream);

}  // namespace braids
