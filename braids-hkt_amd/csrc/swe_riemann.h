// Augmented f-wave Riemann solver for gfx950, templated on the working
// precision R.  Algorithm: GeoClaw 5.14 riemanntype / riemann_aug_JCP / rpn2
// (solver/clawpack-v5.14.0/riemann/src/geoclaw_riemann_utils.f:2-292,
// :547-658; rpn2_geoclaw.f:52-295) as called once per face by wave_flux
// (solver/cut_cells/cutcell_solver.f90:9135-9159).
//
// R = double reproduces the reference operation order term by term (the
// translation unit is built with -ffp-contract=off).  R = float keeps the
// bed elevations in FP64: bed differences and the dry-step wall tests are
// formed in double and only then rounded, so a lake at rest over a large
// absolute elevation stays balanced.
#pragma once
#include <hip/hip_runtime.h>

namespace braids {

template <class R> struct Num;
template <> struct Num<double> {
    static __device__ __forceinline__ double sqrt_(double x) { return __builtin_sqrt(x); }
    static constexpr double eps() { return 2.2204460492503131e-16; }
};
template <> struct Num<float> {
    static __device__ __forceinline__ float sqrt_(float x) { return __builtin_sqrtf(x); }
    static constexpr float eps() { return 1.1920928955078125e-07f; }
};

template <class T> __device__ __forceinline__ T vmax(T a, T b) { return a > b ? a : (b > a ? b : a); }
template <class T> __device__ __forceinline__ T vmin(T a, T b) { return a < b ? a : (b < a ? b : a); }

template <class R> struct RType {
    R hm, s1m, s2m;
};

// riemanntype with maxiter = 1 (only hm, s1m, s2m are consumed downstream)
template <class R>
__device__ __forceinline__ RType<R> riemanntype(R hL, R hR, R uL, R uR, R g, R drytol)
{
    using N = Num<R>;
    RType<R> o;
    R h_min = vmin(hR, hL), h_max = vmax(hR, hL), delu = uR - uL;
    if (h_min <= drytol) {
        o.hm = R(0);
        o.s1m = uR + uL - R(2) * N::sqrt_(g * hR) + R(2) * N::sqrt_(g * hL);
        o.s2m = o.s1m;
    } else {
        R F_min = delu + R(2) * (N::sqrt_(g * h_min) - N::sqrt_(g * h_max));
        R F_max = delu + (h_max - h_min) * (N::sqrt_(R(.5) * g * (h_max + h_min) / (h_max * h_min)));
        if (F_min > R(0)) {
            R t = vmax(R(0), -delu + R(2) * (N::sqrt_(g * hL) + N::sqrt_(g * hR)));
            o.hm = (R(1) / (R(16) * g)) * (t * t);
            o.s1m = uL + R(2) * N::sqrt_(g * hL) - R(3) * N::sqrt_(g * o.hm);
            o.s2m = uR - R(2) * N::sqrt_(g * hR) + R(3) * N::sqrt_(g * o.hm);
        } else if (F_max <= R(0)) {
            R h0 = h_max;
            R gL = N::sqrt_(R(.5) * g * (R(1) / h0 + R(1) / hL));
            R gR = N::sqrt_(R(.5) * g * (R(1) / h0 + R(1) / hR));
            R F0 = delu + (h0 - hL) * gL + (h0 - hR) * gR;
            R dfdh = gL - g * (h0 - hL) / (R(4) * (h0 * h0) * gL) + gR - g * (h0 - hR) / (R(4) * (h0 * h0) * gR);
            R slope = R(2) * N::sqrt_(h0) * dfdh;
            R t = N::sqrt_(h0) - F0 / slope;
            o.hm = t * t;
            R u1m = uL - (o.hm - hL) * N::sqrt_((R(.5) * g) * (R(1) / o.hm + R(1) / hL));
            R u2m = uR + (o.hm - hR) * N::sqrt_((R(.5) * g) * (R(1) / o.hm + R(1) / hR));
            o.s1m = u1m - N::sqrt_(g * o.hm);
            o.s2m = u2m + N::sqrt_(g * o.hm);
        } else {
            R h0 = h_min;
            R F0 = delu + R(2) * (N::sqrt_(g * h0) - N::sqrt_(g * h_max))
                 + (h0 - h_min) * N::sqrt_(R(.5) * g * (R(1) / h0 + R(1) / h_min));
            R slope = (F_max - F0) / (h_max - h_min);
            o.hm = h0 - F0 / slope;
            if (hL > hR) {
                o.s1m = uL + R(2) * N::sqrt_(g * hL) - R(3) * N::sqrt_(g * o.hm);
                o.s2m = uL + R(2) * N::sqrt_(g * hL) - N::sqrt_(g * o.hm);
            } else {
                o.s2m = uR - R(2) * N::sqrt_(g * hR) + R(3) * N::sqrt_(g * o.hm);
                o.s1m = uR - R(2) * N::sqrt_(g * hR) + N::sqrt_(g * o.hm);
            }
        }
    }
    return o;
}

// Face flux of wave_flux: states in (h, hu, hv) Cartesian components, bed
// elevations in FP64, unit normal (nx, ny).  Outputs fl[3], fr[3] (shared
// mass flux in fl[0] == fr[0]) and the face wave speed.
template <class R>
__device__ __forceinline__ void wave_flux(const R l_in[3], const R r_in[3], double bl, double br,
                                          R nx, R ny, R g, R drytol, R fl[3], R fr[3], R &speed)
{
    using N = Num<R>;
    const R tx = -ny, ty = nx;
    R lq[3] = {l_in[0], l_in[1] * nx + l_in[2] * ny, l_in[1] * tx + l_in[2] * ty};
    R rq[3] = {r_in[0], r_in[1] * nx + r_in[2] * ny, r_in[1] * tx + r_in[2] * ty};
    // ---- rpn2 (ixy = 1): left state qr(:,0) = l, right state ql(:,1) = r
    R s[3] = {R(0), R(0), R(0)};
    R fwave[3][3] = {{R(0), R(0), R(0)}, {R(0), R(0), R(0)}, {R(0), R(0), R(0)}};
    if (lq[0] < R(0)) { lq[0] = R(0); lq[1] = R(0); lq[2] = R(0); }
    if (rq[0] < R(0)) { rq[0] = R(0); rq[1] = R(0); rq[2] = R(0); }
    if (!(lq[0] <= drytol && rq[0] <= drytol)) {
        R hL = lq[0], hR = rq[0], huL = lq[1], huR = rq[1], hvL = lq[2], hvR = rq[2];
        double bL = bl, bR = br;
        R uL, uR, vL, vR, phiL, phiR;
        R wall0 = R(1), wall1 = R(1), wall2 = R(1);
        if (hR > drytol) { uR = huR / hR; vR = hvR / hR; phiR = R(0.5) * g * (hR * hR) + (huR * huR) / hR; }
        else { hR = R(0); huR = R(0); hvR = R(0); uR = R(0); vR = R(0); phiR = R(0); }
        if (hL > drytol) { uL = huL / hL; vL = hvL / hL; phiL = R(0.5) * g * (hL * hL) + (huL * huL) / hL; }
        else { hL = R(0); huL = R(0); hvL = R(0); uL = R(0); vL = R(0); phiL = R(0); }
        if (hR <= drytol) {
            RType<R> t = riemanntype<R>(hL, hL, uL, -uL, g, drytol);
            R hstartest = vmax(hL, t.hm);
            if ((double)hstartest + bL < bR) {
                wall1 = R(0); wall2 = R(0);
                hR = hL; huR = -huL; bR = bL; phiR = phiL; uR = -uL; vR = vL;
            } else if ((double)hL + bL < bR) {
                bR = (double)hL + bL;
            }
        } else if (hL <= drytol) {
            RType<R> t = riemanntype<R>(hR, hR, -uR, uR, g, drytol);
            R hstartest = vmax(hR, t.hm);
            if ((double)hstartest + bR < bL) {
                wall0 = R(0); wall1 = R(0);
                hL = hR; huL = -huR; bL = bR; phiL = phiR; uL = -uR; vL = vR;
            } else if ((double)hR + bR < bL) {
                bL = (double)hR + bR;
            }
        }
        R sL = uL - N::sqrt_(g * hL);
        R sR = uR + N::sqrt_(g * hR);
        R uhat = (N::sqrt_(g * hL) * uL + N::sqrt_(g * hR) * uR) / (N::sqrt_(g * hR) + N::sqrt_(g * hL));
        R chat = N::sqrt_(g * R(0.5) * (hR + hL));
        R sE1 = vmin(sL, uhat - chat), sE2 = vmax(sR, uhat + chat);
        // ---- riemann_aug_JCP, maxiter = 1
        R delh = hR - hL, delhu = huR - huL, delphi = phiR - phiL;
        R delb = (R)((bR - bL) + 0.0);   // + pR/(rho g) - pL/(rho g) with pL = pR = 0
        RType<R> rt = riemanntype<R>(hL, hR, uL, uR, g, drytol);
        R lam0 = vmin(sE1, rt.s2m), lam2 = vmax(sE2, rt.s1m);
        sE1 = lam0; sE2 = lam2;
        R hstarHLL = vmax((huL - huR + sE2 * hR - sE1 * hL) / (sE2 - sE1), R(0));
        R lam1 = R(0.5) * (lam0 + lam2);
        // eigenvectors: r(:,0) = (1, lam0, lam0^2), r(:,1) = (0, 0, 1), r(:,2) = (1, lam2, lam2^2)
        R r00 = R(1), r10 = lam0, r20 = lam0 * lam0;
        R r01 = R(0), r11 = R(0), r21 = R(1);
        R r02 = R(1), r12 = lam2, r22 = lam2 * lam2;
        R criticaltol = vmax(drytol * g, R(1e-6));
        R criticaltol_2 = N::sqrt_(criticaltol);
        R hbar = vmax(R(0.5) * (hL + hR), R(0));
        R su = uL + uR;
        R s1s2bar = R(0.25) * (su * su) - g * hbar;
        R s1s2tilde = vmax(R(0), uL * uR) - g * hbar;
        bool sonic = false;
        if (fabs(s1s2bar) <= criticaltol) sonic = true;
        else if (s1s2bar * s1s2tilde <= criticaltol * criticaltol) sonic = true;
        else if (s1s2bar * sE1 * sE2 <= criticaltol * criticaltol) sonic = true;
        else if (vmin(fabs(sE1), fabs(sE2)) < criticaltol_2) sonic = true;
        else if (sE1 < criticaltol_2 && rt.s1m > -criticaltol_2) sonic = true;
        else if (sE2 > -criticaltol_2 && rt.s2m < criticaltol_2) sonic = true;
        else if ((uL + N::sqrt_(g * hL)) * (uR + N::sqrt_(g * hR)) < R(0)) sonic = true;
        else if ((uL - N::sqrt_(g * hL)) * (uR - N::sqrt_(g * hR)) < R(0)) sonic = true;
        R deldelh = sonic ? -delb : delb * g * hbar / s1s2bar;
        if (sE1 < -criticaltol && sE2 > criticaltol) {
            deldelh = vmin(deldelh, hstarHLL * (sE2 - sE1) / sE2);
            deldelh = vmax(deldelh, hstarHLL * (sE2 - sE1) / sE1);
        } else if (sE1 >= criticaltol) {
            deldelh = vmin(deldelh, hstarHLL * (sE2 - sE1) / sE1);
            deldelh = vmax(deldelh, -hL);
        } else if (sE2 <= -criticaltol) {
            deldelh = vmin(deldelh, hR);
            deldelh = vmax(deldelh, hstarHLL * (sE2 - sE1) / sE2);
        }
        R deldelphi = sonic ? -g * hbar * delb : -delb * g * hbar * s1s2tilde / s1s2bar;
        deldelphi = vmin(deldelphi, g * vmax(-hL * delb, -hR * delb));
        deldelphi = vmax(deldelphi, g * vmin(-hL * delb, -hR * delb));
        R d0 = delh - deldelh, d1 = delhu, d2 = delphi - deldelphi;
        // Cramer's rule exactly as written in the reference (3x3 determinants)
        R rm[3][3] = {{r00, r01, r02}, {r10, r11, r12}, {r20, r21, r22}};
        R del[3] = {d0, d1, d2};
        R det1 = rm[0][0] * (rm[1][1] * rm[2][2] - rm[1][2] * rm[2][1]);
        R det2 = rm[0][1] * (rm[1][0] * rm[2][2] - rm[1][2] * rm[2][0]);
        R det3 = rm[0][2] * (rm[1][0] * rm[2][1] - rm[1][1] * rm[2][0]);
        R determinant = det1 - det2 + det3;
        R beta[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            R A[3][3];
#pragma unroll
            for (int i = 0; i < 3; ++i)
#pragma unroll
                for (int j = 0; j < 3; ++j) A[i][j] = (j == k) ? del[i] : rm[i][j];
            det1 = A[0][0] * (A[1][1] * A[2][2] - A[1][2] * A[2][1]);
            det2 = A[0][1] * (A[1][0] * A[2][2] - A[1][2] * A[2][0]);
            det3 = A[0][2] * (A[1][0] * A[2][1] - A[1][1] * A[2][0]);
            beta[k] = (det1 - det2 + det3) / determinant;
        }
        R lam[3] = {lam0, lam1, lam2};
        R rr1[3] = {r10, r11, r12}, rr2[3] = {r20, r21, r22};
        R fw[3][3];
        for (int mw = 0; mw < 3; ++mw) {
            fw[0][mw] = beta[mw] * rr1[mw];
            fw[1][mw] = beta[mw] * rr2[mw];
            fw[2][mw] = beta[mw] * rr1[mw];
        }
        fw[2][0] = fw[2][0] * vL;
        fw[2][2] = fw[2][2] * vR;
        fw[2][1] = R(0);
        R hustar_interface = huL + fw[0][0];
        if (hustar_interface <= R(0)) fw[2][0] = fw[2][0] + (hR * uR * vR - hL * uL * vL - fw[2][0] - fw[2][2]);
        else fw[2][2] = fw[2][2] + (hR * uR * vR - hL * uL * vL - fw[2][0] - fw[2][2]);
        R wall[3] = {wall0, wall1, wall2};
        for (int mw = 0; mw < 3; ++mw) {
            s[mw] = lam[mw] * wall[mw];
            fwave[0][mw] = fw[0][mw] * wall[mw];
            fwave[1][mw] = fw[1][mw] * wall[mw];
            fwave[2][mw] = fw[2][mw] * wall[mw];
        }
    }
    R am[3] = {R(0), R(0), R(0)}, ap[3] = {R(0), R(0), R(0)};
    for (int mw = 0; mw < 3; ++mw) {
        if (s[mw] < R(-1e-14)) {
            for (int j = 0; j < 3; ++j) am[j] = am[j] + fwave[j][mw];
        } else if (s[mw] > R(1e-14)) {
            for (int j = 0; j < 3; ++j) ap[j] = ap[j] + fwave[j][mw];
        } else {
            for (int j = 0; j < 3; ++j) { am[j] = am[j] + R(0.5) * fwave[j][mw]; ap[j] = ap[j] + R(0.5) * fwave[j][mw]; }
        }
    }
    R pl[3] = {R(0), R(0), R(0)}, pr[3] = {R(0), R(0), R(0)};
    if (lq[0] > drytol) { pl[0] = lq[1]; pl[1] = (lq[1] * lq[1]) / lq[0] + R(.5) * g * (lq[0] * lq[0]); pl[2] = lq[1] * lq[2] / lq[0]; }
    if (rq[0] > drytol) { pr[0] = rq[1]; pr[1] = (rq[1] * rq[1]) / rq[0] + R(.5) * g * (rq[0] * rq[0]); pr[2] = rq[1] * rq[2] / rq[0]; }
    R fluxl[3], fluxr[3];
    for (int j = 0; j < 3; ++j) { fluxl[j] = pl[j] + am[j]; fluxr[j] = pr[j] - ap[j]; }
    fl[0] = fluxl[0]; fl[1] = fluxl[1] * nx + fluxl[2] * tx; fl[2] = fluxl[1] * ny + fluxl[2] * ty;
    fr[0] = fluxr[0]; fr[1] = fluxr[1] * nx + fluxr[2] * tx; fr[2] = fluxr[1] * ny + fluxr[2] * ty;
    fl[0] = R(.5) * (fl[0] + fr[0]); fr[0] = fl[0];
    speed = vmax(vmax(fabs(s[0]), fabs(s[1])), fabs(s[2]));
}

// dynamic_runup_opens_face: cutcell_solver.f90:9161-9180
template <class R>
__device__ __forceinline__ bool runup_opens_face(const R l[3], const R r[3], double bl, double br,
                                                 R nx, R ny, R g, R dry)
{
    if (l[0] > dry && r[0] <= dry) {
        R h = l[0], approach = (l[1] * nx + l[2] * ny) / h;
        if (approach <= R(0)) return false;
        RType<R> t = riemanntype<R>(h, h, approach, -approach, g, dry);
        return t.hm > h + dry && (double)t.hm + bl > br + (double)dry;
    } else if (r[0] > dry && l[0] <= dry) {
        R h = r[0], approach = -(r[1] * nx + r[2] * ny) / h;
        if (approach <= R(0)) return false;
        RType<R> t = riemanntype<R>(h, h, approach, -approach, g, dry);
        return t.hm > h + dry && (double)t.hm + br > bl + (double)dry;
    }
    return false;
}

}  // namespace braids

namespace braids {

// FP32 form of wave_flux that keeps a lake at rest exactly balanced.
// Differences from the FP64 form (same waves otherwise):
//  * face depths dl/dr and bed elevations arrive in FP64; after the dry-step
//    wall tests the free-surface jump deta = (hR + bR) - (hL + bL) is formed
//    in FP64, and the first and third components of the decomposed jump are
//    written as deta-based terms plus a remainder that is exactly zero when
//    the velocities vanish:
//      del1 = deta - (deldelh + delb),  del3 = dyn + g hbar deta - (g hbar delb + deldelphi);
//  * the physical fluxes returned exclude the face hydrostatic pressure
//    1/2 g h^2; the caller adds every pressure term in FP64.
__device__ __forceinline__ void wave_flux_f32(const float l_in[3], const float r_in[3], double dl, double dr,
                                              double bl, double br, float nx, float ny, float g, float drytol,
                                              float fl[3], float fr[3], float &speed)
{
    typedef float R;
    const R tx = -ny, ty = nx;
    R lq[3] = {l_in[0], l_in[1] * nx + l_in[2] * ny, l_in[1] * tx + l_in[2] * ty};
    R rq[3] = {r_in[0], r_in[1] * nx + r_in[2] * ny, r_in[1] * tx + r_in[2] * ty};
    R s[3] = {0.f, 0.f, 0.f};
    R fwave[3][3] = {{0.f, 0.f, 0.f}, {0.f, 0.f, 0.f}, {0.f, 0.f, 0.f}};
    if (!(lq[0] <= drytol && rq[0] <= drytol)) {
        R hL = lq[0], hR = rq[0], huL = lq[1], huR = rq[1], hvL = lq[2], hvR = rq[2];
        double hLd = dl, hRd = dr, bL = bl, bR = br;
        R uL, uR, vL, vR;
        R wall0 = 1.f, wall1 = 1.f, wall2 = 1.f;
        if (hR > drytol) { uR = huR / hR; vR = hvR / hR; }
        else { hR = 0.f; hRd = 0.0; huR = 0.f; hvR = 0.f; uR = 0.f; vR = 0.f; }
        if (hL > drytol) { uL = huL / hL; vL = hvL / hL; }
        else { hL = 0.f; hLd = 0.0; huL = 0.f; hvL = 0.f; uL = 0.f; vL = 0.f; }
        if (hR <= drytol) {
            RType<R> t = riemanntype<R>(hL, hL, uL, -uL, g, drytol);
            R hstartest = vmax(hL, t.hm);
            if ((double)hstartest + bL < bR) {
                wall1 = 0.f; wall2 = 0.f;
                hR = hL; hRd = hLd; huR = -huL; bR = bL; uR = -uL; vR = vL;
            } else if (hLd + bL < bR) {
                bR = hLd + bL;
            }
        } else if (hL <= drytol) {
            RType<R> t = riemanntype<R>(hR, hR, -uR, uR, g, drytol);
            R hstartest = vmax(hR, t.hm);
            if ((double)hstartest + bR < bL) {
                wall0 = 0.f; wall1 = 0.f;
                hL = hR; hLd = hRd; huL = -huR; bL = bR; uL = -uR; vL = vR;
            } else if (hRd + bR < bL) {
                bL = hRd + bR;
            }
        }
        const R deta = (R)((hRd + bR) - (hLd + bL));
        const R delb = (R)(bR - bL);
        R sL = uL - __builtin_sqrtf(g * hL);
        R sR = uR + __builtin_sqrtf(g * hR);
        R uhat = (__builtin_sqrtf(g * hL) * uL + __builtin_sqrtf(g * hR) * uR) / (__builtin_sqrtf(g * hR) + __builtin_sqrtf(g * hL));
        R chat = __builtin_sqrtf(g * 0.5f * (hR + hL));
        R sE1 = vmin(sL, uhat - chat), sE2 = vmax(sR, uhat + chat);
        R delh = hR - hL, delhu = huR - huL;
        R dynphi = huR * uR - huL * uL;                 // (hu^2/h)_R - (hu^2/h)_L
        RType<R> rt = riemanntype<R>(hL, hR, uL, uR, g, drytol);
        R lam0 = vmin(sE1, rt.s2m), lam2 = vmax(sE2, rt.s1m);
        sE1 = lam0; sE2 = lam2;
        R hstarHLL = vmax((huL - huR + sE2 * hR - sE1 * hL) / (sE2 - sE1), 0.f);
        R lam1 = 0.5f * (lam0 + lam2);
        R criticaltol = vmax(drytol * g, 1e-6f);
        R criticaltol_2 = __builtin_sqrtf(criticaltol);
        R hbar = vmax(0.5f * (hL + hR), 0.f);
        R su = uL + uR;
        R s1s2bar = 0.25f * (su * su) - g * hbar;
        R s1s2tilde = vmax(0.f, uL * uR) - g * hbar;
        bool sonic = false;
        if (fabsf(s1s2bar) <= criticaltol) sonic = true;
        else if (s1s2bar * s1s2tilde <= criticaltol * criticaltol) sonic = true;
        else if (s1s2bar * sE1 * sE2 <= criticaltol * criticaltol) sonic = true;
        else if (vmin(fabsf(sE1), fabsf(sE2)) < criticaltol_2) sonic = true;
        else if (sE1 < criticaltol_2 && rt.s1m > -criticaltol_2) sonic = true;
        else if (sE2 > -criticaltol_2 && rt.s2m < criticaltol_2) sonic = true;
        else if ((uL + __builtin_sqrtf(g * hL)) * (uR + __builtin_sqrtf(g * hR)) < 0.f) sonic = true;
        else if ((uL - __builtin_sqrtf(g * hL)) * (uR - __builtin_sqrtf(g * hR)) < 0.f) sonic = true;
        // deldelh + delb, exactly zero at rest
        R ddh_bal = sonic ? 0.f : delb * (0.25f * (su * su)) / s1s2bar;
        R deldelh = ddh_bal - delb;
        bool clipped_h = false;
        {
            R d0 = deldelh;
            if (sE1 < -criticaltol && sE2 > criticaltol) {
                deldelh = vmin(deldelh, hstarHLL * (sE2 - sE1) / sE2);
                deldelh = vmax(deldelh, hstarHLL * (sE2 - sE1) / sE1);
            } else if (sE1 >= criticaltol) {
                deldelh = vmin(deldelh, hstarHLL * (sE2 - sE1) / sE1);
                deldelh = vmax(deldelh, -hL);
            } else if (sE2 <= -criticaltol) {
                deldelh = vmin(deldelh, hR);
                deldelh = vmax(deldelh, hstarHLL * (sE2 - sE1) / sE2);
            }
            clipped_h = deldelh != d0;
        }
        // g hbar delb + deldelphi, exactly zero at rest
        R ddp_bal = sonic ? 0.f : g * hbar * delb * (s1s2bar - s1s2tilde) / s1s2bar;
        R deldelphi = ddp_bal - g * hbar * delb;
        bool clipped_p = false;
        {
            R d0 = deldelphi;
            deldelphi = vmin(deldelphi, g * vmax(-hL * delb, -hR * delb));
            deldelphi = vmax(deldelphi, g * vmin(-hL * delb, -hR * delb));
            clipped_p = deldelphi != d0;
        }
        R d0 = clipped_h ? delh - deldelh : deta - ddh_bal;
        R d1 = delhu;
        R d2 = clipped_p ? (0.5f * g * (hR * hR) - 0.5f * g * (hL * hL) + dynphi) - deldelphi
                         : dynphi + g * hbar * deta - ddp_bal;
        // eigenvectors (1, lam0, lam0^2), (0, 0, 1), (1, lam2, lam2^2): solve directly
        R det = lam2 - lam0;
        R b0 = (lam2 * d0 - d1) / det;
        R b2 = (d1 - lam0 * d0) / det;
        R b1 = d2 - b0 * lam0 * lam0 - b2 * lam2 * lam2;
        R fw[3][3];
        fw[0][0] = b0 * lam0; fw[1][0] = b0 * lam0 * lam0; fw[2][0] = b0 * lam0 * vL;
        fw[0][1] = 0.f;       fw[1][1] = b1;               fw[2][1] = 0.f;
        fw[0][2] = b2 * lam2; fw[1][2] = b2 * lam2 * lam2; fw[2][2] = b2 * lam2 * vR;
        R hustar_interface = huL + fw[0][0];
        R trans = huR * vR - huL * vL - fw[2][0] - fw[2][2];
        if (hustar_interface <= 0.f) fw[2][0] += trans; else fw[2][2] += trans;
        R lam[3] = {lam0, lam1, lam2}, wall[3] = {wall0, wall1, wall2};
        for (int mw = 0; mw < 3; ++mw) {
            s[mw] = lam[mw] * wall[mw];
            fwave[0][mw] = fw[0][mw] * wall[mw];
            fwave[1][mw] = fw[1][mw] * wall[mw];
            fwave[2][mw] = fw[2][mw] * wall[mw];
        }
    }
    R am[3] = {0.f, 0.f, 0.f}, ap[3] = {0.f, 0.f, 0.f};
    for (int mw = 0; mw < 3; ++mw) {
        if (s[mw] < -1e-14f) { for (int j = 0; j < 3; ++j) am[j] += fwave[j][mw]; }
        else if (s[mw] > 1e-14f) { for (int j = 0; j < 3; ++j) ap[j] += fwave[j][mw]; }
        else { for (int j = 0; j < 3; ++j) { am[j] += 0.5f * fwave[j][mw]; ap[j] += 0.5f * fwave[j][mw]; } }
    }
    R pl[3] = {0.f, 0.f, 0.f}, pr[3] = {0.f, 0.f, 0.f};
    if (lq[0] > drytol) { pl[0] = lq[1]; pl[1] = (lq[1] * lq[1]) / lq[0]; pl[2] = lq[1] * lq[2] / lq[0]; }
    if (rq[0] > drytol) { pr[0] = rq[1]; pr[1] = (rq[1] * rq[1]) / rq[0]; pr[2] = rq[1] * rq[2] / rq[0]; }
    R fluxl[3], fluxr[3];
    for (int j = 0; j < 3; ++j) { fluxl[j] = pl[j] + am[j]; fluxr[j] = pr[j] - ap[j]; }
    fl[0] = fluxl[0]; fl[1] = fluxl[1] * nx + fluxl[2] * tx; fl[2] = fluxl[1] * ny + fluxl[2] * ty;
    fr[0] = fluxr[0]; fr[1] = fluxr[1] * nx + fluxr[2] * tx; fr[2] = fluxr[1] * ny + fluxr[2] * ty;
    fl[0] = 0.5f * (fl[0] + fr[0]); fr[0] = fl[0];
    speed = vmax(vmax(fabsf(s[0]), fabsf(s[1])), fabsf(s[2]));
}

}  // namespace braids
