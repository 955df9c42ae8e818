// Reference-named operator kernels of the river path: the entries that the
// reference loads through WaveEngine and calls with Engine.launch(name, n,
// *args) (src/braids_support/sediment/river_morphodynamics/sediment_wave.py:22-58).
// Same names, argument order and meaning as the single-source headers
//   explicit_frequency                       river_explicit.h:13-25
//   midpoint_inventory/faces/gather          river_midpoint.h:13-41
//   river_cell_drag (+ river_chezy)          river_hydraulic_drag.h:3-16,
//                                            river_roughness.py:42-47
// with device pointers, launched on the stream set by braids_engine_set_stream.
// FP64; built without FMA contraction.
#include <hip/hip_runtime.h>
#include "braids_hip.h"
#include "braids_engine.h"

namespace {

hipStream_t g_stream = nullptr;

__global__ void k_explicit_frequency(int n, int nf, double grav, double dry, const double *q,
                                     const double *speed, const double *trace, const int *mask,
                                     const int *owner, const double *normal, const double *area,
                                     const double *length, const int *off, const int *faces, double *out)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double sum = 0.;
    for (int k = off[i]; k < off[i + 1]; ++k) {
        int f = faces[k];
        double s = speed[f];
        if (mask[f]) {
            int o = owner[f];
            double h = q[o], hb = trace[f], nx = normal[f], ny = normal[nf + f];
            double a = h > dry ? fabs((q[n + o] * nx + q[2 * n + o] * ny) / h) + sqrt(grav * h) : 0.;
            double b = hb > dry ? fabs((trace[nf + f] * nx + trace[2 * nf + f] * ny) / hb) + sqrt(grav * hb) : 0.;
            s = fmax(a, b);
        }
        sum += length[f] * s;
    }
    out[i] = sum / area[i];
}

__global__ void k_midpoint_inventory(int n, double dt, const double *base, const double *area,
                                     const double *length, const int *off, const int *faces,
                                     const int *signs, const double *water, double *theta)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double out = 0.;
    for (int k = off[i]; k < off[i + 1]; ++k) {
        int f = faces[k];
        out += fmax(signs[k] * water[f], 0.) * length[f];
    }
    theta[i] = 1.;
    if (out > 0. && dt * out >= base[i] * area[i])
        theta[i] = fmin(1., base[i] * area[i] / (dt * out)) * (1. - 32. * 2.2204460492503131e-16);
}

__global__ void k_midpoint_faces(int nf, int n, double dry, const double *q, const double *theta,
                                 const int *o, const int *j, const int *bid, double *water,
                                 double *correction, int *status)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nf) return;
    status[i] = 0; correction[i] = 0.; correction[nf + i] = 0.;
    if (water[i] == 0.) return;
    int donor = water[i] < 0. ? j[i] : o[i];
    if (donor < 0 || theta[donor] == 1.) return;
    if (j[i] < 0 && bid[i] >= 0) { status[i] = 1; return; }
    double delta = (theta[donor] - 1.) * water[i];
    water[i] *= theta[donor];
    if (q[donor] > dry) {
        correction[i] = delta * q[n + donor] / q[donor];
        correction[nf + i] = delta * q[2 * n + donor] / q[donor];
    }
}

__global__ void k_midpoint_gather(int n, int nf, const double *area, const double *length, const int *off,
                                  const int *faces, const int *signs, const double *water,
                                  const double *correction, double *rate)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double mass = 0., mx = 0., my = 0.;
    for (int k = off[i]; k < off[i + 1]; ++k) {
        int f = faces[k];
        double w = signs[k] * length[f] / area[i];
        mass -= w * water[f]; mx -= w * correction[f]; my -= w * correction[nf + f];
    }
    rate[i] = mass; rate[n + i] += mx; rate[2 * n + i] += my;
}

// river_chezy: 1 Chezy C, 2 Manning h^(1/6)/n, 3 log law with ks, 4 Nikuradse-type
__device__ __forceinline__ double river_chezy(int method, double depth, double value)
{
    if (method == 1) return value;
    if (method == 2) return pow(depth, 1. / 6.) / value;
    if (method == 3) return sqrt(9.81) / .41 * log1p(30. * depth / (exp(1.) * value));
    return 18. * log10(12. * fmax(.5, depth / value));
}

__global__ void k_river_cell_drag(int n, double dt, double gravity, double dry, int mode, double value,
                                  const double *values, int feedback, const double *components,
                                  double *q, double *chezy)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (values) value = values[i];
    double h = q[i];
    if (h <= dry) { chezy[i] = river_chezy(mode, dry, value); q[n + i] = q[2 * n + i] = 0.; return; }
    double c = river_chezy(mode, h, value);
    if (feedback) {
        double ks2 = 0.;
        for (int k = 2; k < 5; ++k) { double r = components[k * n + i]; ks2 += r * r; }
        double ks = fmax(.001, fmax(1e-6, fmin(sqrt(ks2), .5 * h)));
        c = 18. * log10(fmax(12. * h / ks, 1.0129));
    }
    chezy[i] = c;
    double speed = sqrt(q[n + i] * q[n + i] + q[2 * n + i] * q[2 * n + i]) / h;
    double factor = 1. / (1. + dt * gravity * speed / (c * c * h));
    q[n + i] *= factor; q[2 * n + i] *= factor;
}

inline dim3 grid(int n) { return dim3(n > 0 ? (n + 255) / 256 : 1); }

int finish()
{
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? BRAIDS_OK : BRAIDS_EHIP;
}

}  // namespace

extern "C" int braids_engine_set_stream(void *stream)
{
    g_stream = (hipStream_t)stream;
    return BRAIDS_OK;
}

extern "C" int explicit_frequency(int n, int nf, double grav, double dry, const double *q, const double *speed,
                                  const double *trace, const int *mask, const int *owner, const double *normal,
                                  const double *area, const double *length, const int *off, const int *faces,
                                  double *out)
{
    if (n <= 0) return BRAIDS_OK;
    hipLaunchKernelGGL(k_explicit_frequency, grid(n), dim3(256), 0, g_stream, n, nf, grav, dry, q, speed, trace,
                       mask, owner, normal, area, length, off, faces, out);
    return finish();
}

extern "C" int midpoint_inventory(int n, double dt, const double *base, const double *area, const double *length,
                                  const int *off, const int *faces, const int *signs, const double *water,
                                  double *theta)
{
    if (n <= 0) return BRAIDS_OK;
    hipLaunchKernelGGL(k_midpoint_inventory, grid(n), dim3(256), 0, g_stream, n, dt, base, area, length, off, faces,
                       signs, water, theta);
    return finish();
}

extern "C" int midpoint_faces(int nf, int n, double dry, const double *q, const double *theta, const int *o,
                              const int *j, const int *bid, double *water, double *correction, int *status)
{
    if (nf <= 0) return BRAIDS_OK;
    hipLaunchKernelGGL(k_midpoint_faces, grid(nf), dim3(256), 0, g_stream, nf, n, dry, q, theta, o, j, bid, water,
                       correction, status);
    return finish();
}

extern "C" int midpoint_gather(int n, int nf, const double *area, const double *length, const int *off,
                               const int *faces, const int *signs, const double *water, const double *correction,
                               double *rate)
{
    if (n <= 0) return BRAIDS_OK;
    hipLaunchKernelGGL(k_midpoint_gather, grid(n), dim3(256), 0, g_stream, n, nf, area, length, off, faces, signs,
                       water, correction, rate);
    return finish();
}

extern "C" int river_cell_drag(int n, double dt, double gravity, double dry, int mode, double value, int feedback,
                               const double *components, double *q, double *chezy)
{
    if (n <= 0) return BRAIDS_OK;
    hipLaunchKernelGGL(k_river_cell_drag, grid(n), dim3(256), 0, g_stream, n, dt, gravity, dry, mode, value,
                       (const double *)nullptr, feedback, components, q, chezy);
    return finish();
}

extern "C" int river_cell_drag_spatial(int n, double dt, double gravity, double dry, int mode, const double *values,
                                       int feedback, const double *components, double *q, double *chezy)
{
    if (n <= 0) return BRAIDS_OK;
    hipLaunchKernelGGL(k_river_cell_drag, grid(n), dim3(256), 0, g_stream, n, dt, gravity, dry, mode, 0.0, values,
                       feedback, components, q, chezy);
    return finish();
}
