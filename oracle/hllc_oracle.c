/*
 * Oracle (test infrastructure only): plain-C FP64 restatement of the HLLC
 * MUSCL-Hancock polygon operators of the reference's hydraulic owner
 * (src/braids_core/hydraulics/hllc.h, hllc_turbines.h, hllc_advanced.h).
 * Same argument order and meaning as the reference entries, so the tests
 * can drive the restatement and the reference build with one argument list.
 *
 *   ohl_reconstruct         hllc.h:15-46   LS gradient of (eta, u, v, z) with
 *                                          the d/|d|^2 weights and the
 *                                          precomputed inverse metric, the
 *                                          per-variable face-midpoint limiter,
 *                                          the Hancock half-step predictor
 *   ohl_flux                hllc.h:48-72   HLLC, HLLE at dry faces and at
 *                                          inadmissible contact states
 *   ohl_predict_source      hllc.h:74-77
 *   ohl_faces               hllc.h:79-114  hydrostatic reconstruction at
 *                                          max(z_L, z_R) and the cell-referenced
 *                                          pressure split
 *   ohl_budget              hllc.h:116-132 donor inventory factor with the
 *                                          prescribed-debit reservation and
 *                                          the face-sum CFL frequency
 *   ohl_limit               hllc.h:134-143
 *   ohl_update              hllc.h:145-155
 *   ohl_turbine_pressure    hllc_turbines.h:13-30
 *   ohl_advanced_frequency  hllc_advanced.h:13-22
 *   ohl_exterior            this build's exterior provider (the reference's
 *                           exterior owner is outside the restatement): the
 *                           HLLC flux between the owner's predictor state and
 *                           the open-boundary trace, rotated back to x/y.
 *
 * Layout: component-major SoA, q[m*n + i]; CSR (off, faces, signs) lists the
 * incident faces of each cell in ascending face order, sign +1 for the owner.
 * Built with -ffp-contract=off so every product and sum rounds on its own.
 */
#include <math.h>

/* (eta, u, v, z) of one cell: velocities vanish on dry cells. */
static void cell_values(int n, int i, double dry, const double *q, const double *z, double v[4])
{
    const double h = q[i];
    v[0] = h + z[i];
    v[1] = h > dry ? q[n + i] / h : 0.;
    v[2] = h > dry ? q[2 * n + i] / h : 0.;
    v[3] = z[i];
}

void ohl_reconstruct(int n, int nf, double gravity, double dry, double dt, const double *q, const double *z,
                     const double *xy, const double *mid, const double *normal, const int *owner,
                     const int *neighbour, const int *selected, const int *off, const int *faces,
                     const double *inverse, double *grad, double *half, double *qhalf, int *status)
{
    for (int i = 0; i < n; ++i) {
        double c[4], sx[4] = {0., 0., 0., 0.}, sy[4] = {0., 0., 0., 0.}, lo[4], hi[4], G[8];
        cell_values(n, i, dry, q, z, c);
        for (int m = 0; m < 4; ++m) { lo[m] = c[m]; hi[m] = c[m]; }
        /* least-squares right-hand side over the incident faces */
        for (int k = off[i]; k < off[i + 1]; ++k) {
            const int f = faces[k];
            const int j = owner[f] == i ? neighbour[f] : owner[f];
            double e[4], dx, dy;
            if (j >= 0) {
                dx = xy[j] - xy[i];
                dy = xy[n + j] - xy[n + i];
                cell_values(n, j, dry, q, z, e);
                /* a dry bank above the cell's surface: use the cell's own level */
                if (e[0] - e[3] <= dry && e[3] >= c[0]) e[0] = c[0];
            } else {
                if (selected[f]) continue;        /* open faces take no part */
                /* wall: mirror image of the cell across the face line */
                const double nx = normal[f], ny = normal[nf + f];
                const double dist = (mid[f] - xy[i]) * nx + (mid[nf + f] - xy[n + i]) * ny;
                dx = 2. * dist * nx;
                dy = 2. * dist * ny;
                for (int m = 0; m < 4; ++m) e[m] = c[m];
                const double un = c[1] * nx + c[2] * ny;
                e[1] -= 2. * un * nx;
                e[2] -= 2. * un * ny;
            }
            const double d2 = dx * dx + dy * dy;
            if (d2 <= 0.) continue;
            for (int m = 0; m < 4; ++m) {
                const double dv = e[m] - c[m];
                sx[m] += dv * dx / d2;
                sy[m] += dv * dy / d2;
                lo[m] = fmin(lo[m], e[m]);
                hi[m] = fmax(hi[m], e[m]);
            }
        }
        /* gradient = inverse metric x right-hand side, limited so that no face
           midpoint leaves the neighbour range */
        const double a = inverse[i], b = inverse[n + i], d = inverse[2 * n + i];
        for (int m = 0; m < 4; ++m) {
            const double gx = a * sx[m] + b * sy[m];
            const double gy = b * sx[m] + d * sy[m];
            double lim = 1.;
            for (int k = off[i]; k < off[i + 1]; ++k) {
                const int f = faces[k];
                const double rise = gx * (mid[f] - xy[i]) + gy * (mid[nf + f] - xy[n + i]);
                if (rise > 0.) lim = fmin(lim, (hi[m] - c[m]) / rise);
                else if (rise < 0.) lim = fmin(lim, (lo[m] - c[m]) / rise);
            }
            G[2 * m] = lim * gx;
            G[2 * m + 1] = lim * gy;
        }
        double h = c[0] - c[3];
        if (h <= dry)
            for (int k = 0; k < 6; ++k) G[k] = 0.;      /* the bed gradient stays */
        for (int k = 0; k < 8; ++k) grad[k * n + i] = G[k];
        /* Hancock predictor over dt/2 of the primitive system */
        const double u = c[1], w = c[2];
        double eta = c[0] - .5 * dt * (u * (G[0] - G[6]) + w * (G[1] - G[7]) + h * (G[2] + G[5]));
        double uh = u - .5 * dt * (u * G[2] + w * G[3] + gravity * G[0]);
        double wh = w - .5 * dt * (u * G[4] + w * G[5] + gravity * G[1]);
        if (eta < c[3]) { eta = c[0]; uh = u; wh = w; }
        if (h <= dry) { eta = c[0]; uh = 0.; wh = 0.; }
        half[i] = eta; half[n + i] = uh; half[2 * n + i] = wh; half[3 * n + i] = c[3];
        h = fmax(0., eta - c[3]);
        qhalf[i] = h; qhalf[n + i] = h * uh; qhalf[2 * n + i] = h * wh;
        status[i] = !isfinite(eta) || !isfinite(uh) || !isfinite(wh);
    }
}

/* Normal-frame HLLC flux of (h, u_n, u_t) left/right; out = (mass, normal
   momentum, tangential momentum), *speed the largest signal speed. */
void ohl_flux(double hl, double ul, double vl, double hr, double ur, double vr, double g, double dry, double *out,
              double *speed)
{
    const double cl = sqrt(g * hl), cr = sqrt(g * hr);
    double sl = fmin(ul - cl, ur - cr), sr = fmax(ul + cl, ur + cr);
    if (hl <= dry) sl = ur - 2. * cr;                 /* dry-bed fronts */
    if (hr <= dry) sr = ul + 2. * cl;
    *speed = fmax(fabs(sl), fabs(sr));
    const double fL[3] = {hl * ul, hl * ul * ul + .5 * g * hl * hl, hl * ul * vl};
    const double fR[3] = {hr * ur, hr * ur * ur + .5 * g * hr * hr, hr * ur * vr};
    if (hl <= dry && hr <= dry) { out[0] = out[1] = out[2] = 0.; *speed = 0.; return; }
    if (sl >= 0.) { out[0] = fL[0]; out[1] = fL[1]; out[2] = fL[2]; return; }
    if (sr <= 0.) { out[0] = fR[0]; out[1] = fR[1]; out[2] = fR[2]; return; }
    const double ml = hl * (sl - ul), mr = hr * (sr - ur), den = ml - mr;
    const double sm = fabs(den) > 1e-30 ? (.5 * g * (hr * hr - hl * hl) + ml * ul - mr * ur) / den : 0.;
    if (hl > dry && hr > dry && sl < sm && sm < sr) {
        const double hs = sm >= 0. ? ml / (sl - sm) : mr / (sr - sm);
        if (hs >= 0. && isfinite(hs)) {
            if (sm >= 0.) {
                out[0] = fL[0] + sl * (hs - hl);
                out[1] = fL[1] + sl * (hs * sm - hl * ul);
                out[2] = fL[2] + sl * (hs * vl - hl * vl);
            } else {
                out[0] = fR[0] + sr * (hs - hr);
                out[1] = fR[1] + sr * (hs * sm - hr * ur);
                out[2] = fR[2] + sr * (hs * vr - hr * vr);
            }
            return;
        }
    }
    /* HLLE */
    const double uL[3] = {hl, hl * ul, hl * vl}, uR[3] = {hr, hr * ur, hr * vr};
    for (int k = 0; k < 3; ++k) out[k] = (sr * fL[k] - sl * fR[k] + sl * sr * (uR[k] - uL[k])) / (sr - sl);
}

void ohl_predict_source(int n, double dry, double dt, const double *rate, double *half, double *qhalf)
{
    for (int i = 0; i < n; ++i) {
        const double h = qhalf[i];
        qhalf[n + i] += .5 * dt * rate[n + i];
        qhalf[2 * n + i] += .5 * dt * rate[2 * n + i];
        half[n + i] = h > dry ? qhalf[n + i] / h : 0.;
        half[2 * n + i] = h > dry ? qhalf[2 * n + i] / h : 0.;
    }
}

/* Linear extrapolation of the four half-step values of cell c to face f. */
static void extrapolate(int n, int nf, int c, int f, const double *half, const double *grad, const double *xy,
                        const double *mid, double v[4])
{
    const double rx = mid[f] - xy[c], ry = mid[nf + f] - xy[n + c];
    for (int m = 0; m < 4; ++m) v[m] = half[m * n + c] + grad[2 * m * n + c] * rx + grad[(2 * m + 1) * n + c] * ry;
}

void ohl_faces(int n, int nf, double gravity, double dry, const double *half, const double *grad, const double *xy,
               const double *mid, const double *normal, const int *owner, const int *neighbour,
               const int *selected, const double *external, const double *external_trace, double *fl, double *fr,
               double *speed)
{
    for (int f = 0; f < nf; ++f) {
        const int l = owner[f], r = neighbour[f];
        const double nx = normal[f], ny = normal[nf + f];
        if (selected[f]) {
            /* exterior flux as provided, net of the cell-centre pressure */
            const double h = fmax(0., half[l] - half[3 * n + l]);
            fl[f] = fr[f] = external[f];
            fl[nf + f] = fr[nf + f] = external[nf + f] - .5 * gravity * h * h * nx;
            fl[2 * nf + f] = fr[2 * nf + f] = external[2 * nf + f] - .5 * gravity * h * h * ny;
            const double hb = external_trace[f];
            const double ub = hb > dry ? (external_trace[nf + f] * nx + external_trace[2 * nf + f] * ny) / hb : 0.;
            speed[f] = fmax(fabs(half[n + l] * nx + half[2 * n + l] * ny) + sqrt(gravity * h),
                            fabs(ub) + sqrt(gravity * fmax(hb, 0.)));
            continue;
        }
        double L[4], R[4];
        extrapolate(n, nf, l, f, half, grad, xy, mid, L);
        if (r < 0) for (int m = 0; m < 4; ++m) R[m] = L[m];
        else extrapolate(n, nf, r, f, half, grad, xy, mid, R);
        /* a dry side keeps its actual cell bed and carries no water */
        const int ldry = half[l] - half[3 * n + l] <= dry;
        const int rdry = r >= 0 && half[r] - half[3 * n + r] <= dry;
        if (ldry || rdry) {
            L[3] = half[3 * n + l];
            R[3] = r < 0 ? L[3] : half[3 * n + r];
        }
        if (ldry) { L[0] = L[3]; L[1] = 0.; L[2] = 0.; }
        if (rdry) { R[0] = R[3]; R[1] = 0.; R[2] = 0.; }
        if (r < 0) {                                   /* wall: reflected normal velocity */
            const double un = R[1] * nx + R[2] * ny;
            R[1] -= 2. * un * nx;
            R[2] -= 2. * un * ny;
        }
        const double zs = fmax(L[3], R[3]);
        const double hl = fmax(0., L[0] - L[3]), hr = fmax(0., R[0] - R[3]);
        const double hls = fmax(0., L[0] - zs), hrs = fmax(0., R[0] - zs);
        double out[3], s;
        ohl_flux(hls, L[1] * nx + L[2] * ny, -L[1] * ny + L[2] * nx, hrs, R[1] * nx + R[2] * ny,
                 -R[1] * ny + R[2] * nx, gravity, dry, out, &s);
        const double el = fmax(0., half[l] - L[3]);
        const double er = r < 0 ? el : fmax(0., half[r] - R[3]);
        const double pl = .5 * gravity * (hl * hl - hls * hls - el * el);
        const double pr = .5 * gravity * (hr * hr - hrs * hrs - er * er);
        fl[f] = fr[f] = r < 0 ? 0. : out[0];
        fl[nf + f] = (out[1] + pl) * nx - out[2] * ny;
        fl[2 * nf + f] = (out[1] + pl) * ny + out[2] * nx;
        fr[nf + f] = (out[1] + pr) * nx - out[2] * ny;
        fr[2 * nf + f] = (out[1] + pr) * ny + out[2] * nx;
        speed[f] = s;
    }
}

void ohl_budget(int n, int nf, double dt, const double *q, const double *area, const double *length, const int *off,
                const int *faces, const int *signs, const double *fl, const double *speed, const int *selected,
                const int *status, double *theta, double *frequency)
{
    (void)nf;
    for (int i = 0; i < n; ++i) {
        double internal = 0., prescribed = 0., sum = 0.;
        for (int k = off[i]; k < off[i + 1]; ++k) {
            const int f = faces[k];
            const double debit = fmax(0., signs[k] * fl[f]) * length[f];
            if (selected[f]) prescribed += debit;
            else internal += debit;
            sum += speed[f] * length[f];
        }
        frequency[i] = status[i] ? INFINITY : sum / area[i];
        theta[i] = 1.;
        const double stock = q[i] * area[i] / dt;
        if (prescribed > stock) theta[i] = -1.;        /* unfunded exterior withdrawal */
        else if (internal > 0. && internal >= stock - prescribed)
            theta[i] = fmin(1., (stock - prescribed) / internal) * (1. - 32. * 2.2204460492503131e-16);
    }
}

void ohl_limit(int n, int nf, double dry, const double *q, const double *theta, const int *owner,
               const int *neighbour, const int *selected, double *fl, double *fr, int *status)
{
    for (int f = 0; f < nf; ++f) {
        status[f] = 0;
        const double water = fl[f];
        const int donor = water < 0. ? neighbour[f] : owner[f];
        if (donor < 0 || water == 0.) continue;
        if (selected[f]) { status[f] = theta[donor] < 0. ? 1 : 0; continue; }
        if (theta[donor] >= 1.) continue;
        const double factor = fmax(0., theta[donor]);
        const double delta = (factor - 1.) * water;
        fl[f] = fr[f] = water * factor;
        if (q[donor] > dry)
            for (int m = 1; m < 3; ++m) {
                const double dm = delta * q[m * n + donor] / q[donor];
                fl[m * nf + f] += dm;
                fr[m * nf + f] += dm;
            }
    }
}

void ohl_update(int n, int nf, double dt, double dry, const double *q, const double *area, const double *length,
                const int *off, const int *faces, const int *signs, const double *fl, const double *fr,
                const int *face_status, double *next, double *status)
{
    for (int i = 0; i < n; ++i) {
        double v[3] = {q[i], q[n + i], q[2 * n + i]};
        int bad = 0;
        for (int k = off[i]; k < off[i + 1]; ++k) {
            const int f = faces[k], s = signs[k];
            const double a = dt * s * length[f] / area[i];
            const double *F = s > 0 ? fl : fr;
            for (int m = 0; m < 3; ++m) v[m] -= a * F[m * nf + f];
            bad = bad || face_status[f];
        }
        if (v[0] <= dry) { v[1] = 0.; v[2] = 0.; }
        for (int m = 0; m < 3; ++m) {
            next[m * n + i] = v[m];
            bad = bad || !isfinite(v[m]);
        }
        status[i] = bad || v[0] < 0. ? 1. : 0.;
    }
}

void ohl_turbine_pressure(int n, int nf, double gravity, double dry, const double *half, const double *grad,
                          const double *xy, const double *mid, const double *normal, const int *owner,
                          const int *nb, double *fl, double *fr)
{
    for (int f = 0; f < nf; ++f) {
        const int a = owner[f], b = nb[f];
        if (b < 0) continue;
        if (half[a] - half[3 * n + a] <= dry || half[b] - half[3 * n + b] <= dry) continue;
        double L[4], R[4];
        extrapolate(n, nf, a, f, half, grad, xy, mid, L);
        extrapolate(n, nf, b, f, half, grad, xy, mid, R);
        const double bed = fmax(L[3], R[3]), hl = L[0] - bed, hr = R[0] - bed, h = .5 * (hl + hr);
        if (fmin(hl, hr) <= dry || fabs(hr - hl) > .005 * h) continue;   /* fully wet, small jump only */
        const double c = sqrt(gravity * h);
        const double froude = fmin(1., fmax(hypot(L[1], L[2]), hypot(R[1], R[2])) / c);
        const double nx = normal[f], ny = normal[nf + f];
        const double du = (R[1] - L[1]) * nx + (R[2] - L[2]) * ny;
        const double delta = .5 * (1. - froude) * c * (2. * hl * hr / (hl + hr)) * du;
        fl[nf + f] += delta * nx; fr[nf + f] += delta * nx;
        fl[2 * nf + f] += delta * ny; fr[2 * nf + f] += delta * ny;
    }
}

void ohl_advanced_frequency(int n, int nf, double dry, double nu, const double *q, const double *metric,
                            const int *owner, const int *nb, const int *off, const int *faces,
                            const double *length, const double *area, double *out)
{
    (void)nf;
    for (int i = 0; i < n; ++i) {
        double sum = 0.;
        if (q[i] > dry)
            for (int k = off[i]; k < off[i + 1]; ++k) {
                const int f = faces[k], j = nb[f], o = owner[f];
                if (j >= 0 && q[o] > dry && q[j] > dry) sum += nu * length[f] * metric[f];
            }
        out[i] = sum / area[i];
    }
}

/* Exterior flux of the open faces: the HLLC flux between the owner's
   predictor state and the boundary trace (both over the owner's bed),
   rotated back to x/y.  Faces with selected[f] == 0 are left untouched. */
void ohl_exterior(int n, int nf, double gravity, double dry, const double *qhalf, const double *trace,
                  const int *owner, const double *normal, const int *selected, double *external)
{
    for (int f = 0; f < nf; ++f) {
        if (!selected[f]) continue;
        const int o = owner[f];
        const double nx = normal[f], ny = normal[nf + f];
        const double hl = qhalf[o], hr = trace[f];
        const double ul = hl > dry ? (qhalf[n + o] * nx + qhalf[2 * n + o] * ny) / hl : 0.;
        const double vl = hl > dry ? (-qhalf[n + o] * ny + qhalf[2 * n + o] * nx) / hl : 0.;
        const double ur = hr > dry ? (trace[nf + f] * nx + trace[2 * nf + f] * ny) / hr : 0.;
        const double vr = hr > dry ? (-trace[nf + f] * ny + trace[2 * nf + f] * nx) / hr : 0.;
        double out[3], s;
        ohl_flux(hl, ul, vl, hr, ur, vr, gravity, dry, out, &s);
        external[f] = out[0];
        external[nf + f] = out[1] * nx - out[2] * ny;
        external[2 * nf + f] = out[1] * ny + out[2] * nx;
    }
}

/* Inverse least-squares metric of each cell (hllc.py:33-43): M = sum d d^T /
   |d|^2 over the neighbours and the wall mirror images (open faces skipped),
   inverted in closed form; a metric whose smaller eigenvalue is at most
   1e-12 of the larger gets the rank-one pseudo-inverse v v^T / lambda_max
   (numpy.linalg.pinv with rcond 1e-12).  out[3n] = (inv00, inv01, inv11). */
void ohl_metric_inverse(int n, int nf, const double *xy, const double *mid, const double *normal, const int *owner,
                        const int *neighbour, const int *selected, const int *off, const int *faces,
                        const int *signs, double *out)
{
    for (int i = 0; i < n; ++i) {
        double a = 0., b = 0., d = 0.;
        for (int k = off[i]; k < off[i + 1]; ++k) {
            const int f = faces[k];
            const int j = signs[k] > 0 ? neighbour[f] : owner[f];
            double dx, dy;
            if (j >= 0) {
                dx = xy[j] - xy[i];
                dy = xy[n + j] - xy[n + i];
            } else {
                if (selected[f]) continue;
                const double nx = normal[f], ny = normal[nf + f];
                const double dist = (mid[f] - xy[i]) * nx + (mid[nf + f] - xy[n + i]) * ny;
                dx = 2. * dist * nx;
                dy = 2. * dist * ny;
            }
            const double nn = dx * dx + dy * dy;
            if (nn > 0.) { a += dx * dx / nn; b += dx * dy / nn; d += dy * dy / nn; }
        }
        const double tr = a + d, disc = sqrt(.25 * (a - d) * (a - d) + b * b);
        const double l1 = .5 * tr + disc, l2 = .5 * tr - disc;
        if (l1 <= 0.) { out[i] = out[n + i] = out[2 * n + i] = 0.; continue; }
        if (l2 > 1e-12 * l1) {
            const double det = a * d - b * b;
            out[i] = d / det; out[n + i] = -b / det; out[2 * n + i] = a / det;
        } else {
            /* eigenvector of l1 */
            double vx = b, vy = l1 - a;
            if (fabs(vx) + fabs(vy) == 0.) { vx = l1 - d; vy = b; }
            if (fabs(vx) + fabs(vy) == 0.) { vx = 1.; vy = 0.; }
            const double s = 1. / (vx * vx + vy * vy);
            out[i] = vx * vx * s / l1; out[n + i] = vx * vy * s / l1; out[2 * n + i] = vy * vy * s / l1;
        }
    }
}
