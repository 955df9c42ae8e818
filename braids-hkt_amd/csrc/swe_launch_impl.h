// Definitions of the per-precision launch sequences (included once per
// precision translation unit, which then instantiates them).
#pragma once
#include "swe_launch.h"
#include "swe_kernels.h"
#include <type_traits>

namespace braids {

static inline int grid_for(int n, int block) { return n > 0 ? (n + block - 1) / block : 1; }

struct EvScope {
    const LaunchCfg &cfg; hipStream_t s; int idx;
    bool on;
    EvScope(const LaunchCfg &c, hipStream_t st, int slot, int kind)
        : cfg(c), s(st), idx((slot * KK_COUNT + kind) * 2), on(c.ev && ((c.ev_mask >> kind) & 1))
    {
        if (on) (void)hipEventRecord(cfg.ev[idx], s);
    }
    ~EvScope()
    {
        if (on) (void)hipEventRecord(cfg.ev[idx + 1], s);
    }
};

template <class R, int SCHEME>
static void launch_prep(const MeshView<R> &m, StateView<R> in, const double *z, Work<R> w, Control *ctl,
                        int slot, const LaunchCfg &cfg, hipStream_t s)
{
    const int B = cfg.block;
    if constexpr (std::is_same<R, double>::value && SCHEME == 5)
        hipLaunchKernelGGL(k_prep_ls64, dim3(grid_for(m.nc, B)), dim3(B), 0, s, m, in, z, w, ctl, slot, cfg.remap);
    else
        hipLaunchKernelGGL((k_prep<R, SCHEME>), dim3(grid_for(m.nc, B)), dim3(B), 0, s, m, in, z, w, ctl, slot, cfg.remap);
}

template <class R, int SCHEME>
static void launch_face(const MeshView<R> &m, StateView<R> in, const double *z, Work<R> w,
                        const LaunchCfg &cfg, hipStream_t s)
{
    const int B = cfg.block;
    if constexpr (std::is_same<R, float>::value)
        hipLaunchKernelGGL((k_face_f32<SCHEME>), dim3(grid_for(m.nf, B)), dim3(B), 0, s, m, in, z, w, cfg.remap);
    else
        hipLaunchKernelGGL((k_face<R, SCHEME>), dim3(grid_for(m.nf, B)), dim3(B), 0, s, m, in, z, w, cfg.remap);
}

template <class R, int SCHEME>
static void stage_kernels(const MeshView<R> &m, StateView<R> in, StateView<R> base, StateView<R> out,
                          const double *z, Work<R> w, Control *ctl, double *ledger_stage, int mode,
                          double dt, double t, const LaunchCfg &cfg, hipStream_t s)
{
    const int B = cfg.block;
    { EvScope e(cfg, s, mode, KK_PREP); launch_prep<R, SCHEME>(m, in, z, w, ctl, mode, cfg, s); }
    if (m.nb > 0) {
        EvScope e(cfg, s, mode, KK_TRACES);
        hipLaunchKernelGGL((k_traces<R, SCHEME>), dim3(m.nb), dim3(TRACE_BLOCK), 0, s, m, in, z, w, ctl, mode, t);
    }
    { EvScope e(cfg, s, mode, KK_FACE); launch_face<R, SCHEME>(m, in, z, w, cfg, s); }
    { EvScope e(cfg, s, mode, KK_DRAIN);
      hipLaunchKernelGGL((k_drain<R>), dim3(grid_for(m.nc, B)), dim3(B), 0, s, m, in, w, ctl, mode, dt, cfg.remap); }
    {
    EvScope eg(cfg, s, mode, KK_GATHER);
    switch (mode) {
    case 0: hipLaunchKernelGGL((k_gather<R, 0>), dim3(grid_for(m.nc, B)), dim3(B), 0, s, m, in, base, out, w, ctl, dt, cfg.remap); break;
    case 1: hipLaunchKernelGGL((k_gather<R, 1>), dim3(grid_for(m.nc, B)), dim3(B), 0, s, m, in, base, out, w, ctl, dt, cfg.remap); break;
    case 2: hipLaunchKernelGGL((k_gather<R, 2>), dim3(grid_for(m.nc, B)), dim3(B), 0, s, m, in, base, out, w, ctl, dt, cfg.remap); break;
    default: hipLaunchKernelGGL((k_gather<R, 3>), dim3(grid_for(m.nc, B)), dim3(B), 0, s, m, in, base, out, w, ctl, dt, cfg.remap); break;
    }
    }
    if (m.nb > 0 && ledger_stage) {
        EvScope e(cfg, s, mode, KK_LEDGER);
        hipLaunchKernelGGL((k_ledger<R>), dim3(m.nb), dim3(TRACE_BLOCK), 0, s, m, w, ledger_stage, dt);
    }
}

template <class R>
void run_stage(const MeshView<R> &m, StateView<R> in, StateView<R> base, StateView<R> out,
               const double *z, Work<R> w, Control *ctl, double *ledger_stage, int mode,
               double dt, double t, const LaunchCfg &cfg, hipStream_t stream)
{
    if (m.scheme == 5) stage_kernels<R, 5>(m, in, base, out, z, w, ctl, ledger_stage, mode, dt, t, cfg, stream);
    else stage_kernels<R, 1>(m, in, base, out, z, w, ctl, ledger_stage, mode, dt, t, cfg, stream);
}

template <class R, int SCHEME>
static void rate_kernels(const MeshView<R> &m, StateView<R> in, const double *z, Work<R> w, Control *ctl,
                         double *rate, double *water, double t, const LaunchCfg &cfg, hipStream_t s)
{
    const int B = cfg.block;
    launch_prep<R, SCHEME>(m, in, z, w, ctl, 3, cfg, s);
    if (m.nb > 0)
        hipLaunchKernelGGL((k_traces<R, SCHEME>), dim3(m.nb), dim3(TRACE_BLOCK), 0, s, m, in, z, w, ctl, 3, t);
    launch_face<R, SCHEME>(m, in, z, w, cfg, s);
    const int n = m.nc > m.nf ? m.nc : m.nf;
    hipLaunchKernelGGL((k_rates<R>), dim3(grid_for(n, B)), dim3(B), 0, s, m, in, w, rate, water, 0);
}

template <class R>
void run_rates(const MeshView<R> &m, StateView<R> in, const double *z, Work<R> w, Control *ctl,
               double *rate, double *water, double t, const LaunchCfg &cfg, hipStream_t stream)
{
    if (m.scheme == 5) rate_kernels<R, 5>(m, in, z, w, ctl, rate, water, t, cfg, stream);
    else rate_kernels<R, 1>(m, in, z, w, ctl, rate, water, t, cfg, stream);
}

}  // namespace braids
