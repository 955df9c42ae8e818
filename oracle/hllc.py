"""Oracle for the HLLC MUSCL-Hancock hydraulic owner (test infrastructure only).

Binds the plain-C restatement of the HLLC operators (hllc_oracle.c in
liboracle.so) and, where they have been compiled, the reference's own
single-source headers (hllc.h, hllc_turbines.h, hllc_advanced.h built with
g++ into _ref/).  HLLCOracle restates the owner around them:

  * geometry and the inverse least-squares metric   hllc.py:23-45
  * _spatial_step: Strang drag, reconstruction and Hancock predictor, the
    momentum source at the predictor, the exterior flux, faces, the
    low-Froude face correction of an installed rotor array
    (river_turbines.py:38-42), budget, CFL check, limiter, update, momentum
    source at the half time, closing drag         hllc.py:76-118
  * advance: step hint, truncation at the interval end, rejections
                                                   hllc.py:131-160

The exterior provider is this build's: first-order open-boundary traces
(modes 2 and 6, the Fortran trace rules) of the predictor state at
t + dt/2, and the HLLC flux between the owner's predictor state and the
trace.  The momentum source is the Thetis rotor array of the river oracle.
"""
import ctypes

import numpy as np

from . import river, swe

HERE = swe.HERE
REF_HLLC = HERE / '_ref' / 'libhllc_ref.so'
REF_TURBINES = HERE / '_ref' / 'libhllc_turbines_ref.so'
REF_ADVANCED = HERE / '_ref' / 'libhllc_advanced_ref.so'

_D = ctypes.c_double
_P = ctypes.c_void_p
_I = ctypes.c_int

OPS = {
    'hllc_reconstruct': [_I, _I, _D, _D, _D] + [_P] * 15,
    'hllc_predict_source': [_I, _D, _D] + [_P] * 3,
    'hllc_faces': [_I, _I, _D, _D] + [_P] * 13,
    'hllc_budget': [_I, _I, _D] + [_P] * 12,
    'hllc_limit': [_I, _I, _D] + [_P] * 8,
    'hllc_update': [_I, _I, _D, _D] + [_P] * 11,
    'hllc_turbine_pressure': [_I, _I, _D, _D] + [_P] * 9,
    'hllc_advanced_frequency': [_I, _I, _D, _D] + [_P] * 9,
}


def _conv(a):
    if isinstance(a, np.ndarray):
        assert a.flags.c_contiguous
        return ctypes.c_void_p(a.ctypes.data)
    return a


class _Op:
    def __init__(self, fn):
        self.fn = fn

    def __call__(self, *args):
        self.fn(*[_conv(a) for a in args])


def lib():
    L = river.lib()
    if not getattr(L, '_hllc_bound', False):
        for name, args in OPS.items():
            fn = getattr(L, 'o' + name.replace('hllc_', 'hl_'))
            fn.argtypes = args; fn.restype = None
        L.ohl_flux.argtypes = [_D] * 8 + [_P, _P]; L.ohl_flux.restype = None
        L.ohl_exterior.argtypes = [_I, _I, _D, _D] + [_P] * 6; L.ohl_exterior.restype = None
        L.ohl_metric_inverse.argtypes = [_I, _I] + [_P] * 10; L.ohl_metric_inverse.restype = None
        L.om_boundary_traces.argtypes = [_P, _P, _P, _D, _P, _P]; L.om_boundary_traces.restype = _I
        L._hllc_bound = True
    return L


def oracle_ops():
    L = lib()
    return {n: _Op(getattr(L, 'o' + n.replace('hllc_', 'hl_'))) for n in OPS}


def reference_ops():
    """The reference's compiled HLLC headers: {name: callable}, or None."""
    if not (REF_HLLC.exists() and REF_TURBINES.exists() and REF_ADVANCED.exists()):
        return None
    out = {}
    for path in (REF_HLLC, REF_TURBINES, REF_ADVANCED):
        L = ctypes.CDLL(str(path))
        L.mobile_threads.argtypes = [_I]
        L.mobile_threads(1)
        for n, args in OPS.items():
            fn = getattr(L, n, None)
            if fn is None:
                continue
            fn.argtypes = args; fn.restype = None
            out[n] = _Op(fn)
    return out


def flux(hl, ul, vl, hr, ur, vr, g=9.81, dry=1e-8):
    """Normal-frame HLLC flux: (out[3], speed)."""
    out = np.zeros(3); s = np.zeros(1)
    lib().ohl_flux(hl, ul, vl, hr, ur, vr, g, dry, out.ctypes.data, s.ctypes.data)
    return out, float(s[0])


def open_faces(mesh):
    return ((mesh.neighbour < 0) & (mesh.bid >= 0)).astype(np.int32)


def metric_inverse(mesh):
    """[3, nc] (inv00, inv01, inv11) of the inverse LS metric of every cell
    (hllc.py:33-43), by the closed-form restatement in hllc_oracle.c."""
    c = np.ascontiguousarray
    n, nf = mesh.nc, mesh.nf
    out = np.zeros((3, n))
    args = [c(mesh.xy, np.float64), c(mesh.mid, np.float64), c(mesh.normal, np.float64), c(mesh.owner, np.int32),
            c(mesh.neighbour, np.int32), open_faces(mesh), c(mesh.offset, np.int32), c(mesh.faces_csr, np.int32),
            c(mesh.signs, np.int32), out]
    lib().ohl_metric_inverse(n, nf, *[a.ctypes.data for a in args])
    return out


def geometry(mesh):
    c = np.ascontiguousarray
    return dict(xy=c(mesh.xy, np.float64), mid=c(mesh.mid, np.float64), normal=c(mesh.normal, np.float64),
                owner=c(mesh.owner, np.int32), neighbour=c(mesh.neighbour, np.int32),
                length=c(mesh.length, np.float64), area=c(mesh.area, np.float64),
                off=c(mesh.offset, np.int32), faces=c(mesh.faces_csr, np.int32), signs=c(mesh.signs, np.int32),
                inverse=c(metric_inverse(mesh)), selected=open_faces(mesh))


class Rejected(Exception):
    def __init__(self, kind):
        super().__init__(kind)
        self.kind = kind


class HLLCOracle:
    """FP64 restatement of HLLCHydraulics (Hancock integration) on one mesh."""

    def __init__(self, mesh, cfl=.9, grav=9.81, dry=1e-8, drag_mode=2, roughness=None, rows=None,
                 pressure=None, ops=None):
        self.mesh = mesh
        self.n, self.nf, self.nb = mesh.nc, mesh.nf, mesh.nb
        self.g = geometry(mesh)
        self.ops = ops or oracle_ops()
        self.grav, self.dry, self.cfl = float(grav), float(dry), float(cfl)
        self.drag_mode = int(drag_mode)
        self.rough = np.ascontiguousarray(mesh.manning if roughness is None else roughness, dtype=np.float64)
        self.L = lib()
        self.exterior_on = bool(self.g['selected'].any())
        self.bnd = swe.Oracle(mesh, scheme=1, grav=grav, dry=dry) if self.exterior_on else None
        self.turb = river.RiverOracle(mesh, rows=rows, drag_mode=0, grav=grav, dry=dry) if rows else None
        self.pressure = bool(rows) if pressure is None else bool(pressure)
        self.hint = None
        self.peak = 0.
        self.last_diag = None

    # -- pieces -----------------------------------------------------------
    def friction(self, q, dt):
        if self.drag_mode <= 0:
            return q
        return river.cell_drag(q, dt, self.drag_mode, self.rough, self.grav, self.dry)[0]

    def momentum(self, q, z):
        rate, diag, rc = self.turb.turbines(q, z)
        if rc < 0:
            raise RuntimeError('turbine induction outside its real domain')
        self.last_diag = diag
        return rate

    def exterior(self, qhalf, z, t):
        n, nf = self.n, self.nf
        trace = np.zeros((3, nf)); fail = np.zeros(1, np.int32)
        ok = self.L.om_boundary_traces(self.bnd.ctx, qhalf.ctypes.data, np.ascontiguousarray(z).ctypes.data,
                                       float(t), trace.ctypes.data, fail.ctypes.data)
        if not ok:
            raise Rejected('exterior')
        ext = np.zeros((3, nf))
        g = self.g
        self.L.ohl_exterior(n, nf, self.grav, self.dry, qhalf.ctypes.data, trace.ctypes.data,
                            g['owner'].ctypes.data, g['normal'].ctypes.data, g['selected'].ctypes.data,
                            ext.ctypes.data)
        return ext, trace

    # -- HLLCHydraulics._spatial_step (Hancock, predictor 1/2) ---------------
    def spatial_step(self, q, z, dt, t):
        n, nf, g, o = self.n, self.nf, self.g, self.ops
        z = np.ascontiguousarray(z, dtype=np.float64)
        self.peak = 0.
        work = self.friction(np.array(q, dtype=np.float64, order='C'), .5 * dt)
        grad = np.zeros((8, n)); half = np.zeros((4, n)); qhalf = np.zeros((3, n))
        pstat = np.zeros(n, np.int32)
        o['hllc_reconstruct'](n, nf, self.grav, self.dry, float(dt), work, z, g['xy'], g['mid'], g['normal'],
                              g['owner'], g['neighbour'], g['selected'], g['off'], g['faces'], g['inverse'],
                              grad, half, qhalf, pstat)
        if self.turb is not None:
            rate = self.momentum(work, z)
            o['hllc_predict_source'](n, self.dry, float(dt), rate, half, qhalf)
        ext = np.zeros((3, nf)); trace = np.zeros((3, nf))
        if self.exterior_on:
            ext, trace = self.exterior(qhalf, z, t + .5 * dt)
        fl = np.zeros((3, nf)); fr = np.zeros((3, nf)); speed = np.zeros(nf)
        o['hllc_faces'](n, nf, self.grav, self.dry, half, grad, g['xy'], g['mid'], g['normal'], g['owner'],
                        g['neighbour'], g['selected'], ext, trace, fl, fr, speed)
        if self.pressure:
            o['hllc_turbine_pressure'](n, nf, self.grav, self.dry, half, grad, g['xy'], g['mid'], g['normal'],
                                       g['owner'], g['neighbour'], fl, fr)
        theta = np.ones(n); freq = np.zeros(n)
        o['hllc_budget'](n, nf, float(dt), work, g['area'], g['length'], g['off'], g['faces'], g['signs'], fl,
                         speed, g['selected'], pstat, theta, freq)
        peak = float(np.max(freq))
        self.peak = peak
        if not np.isfinite(peak) or dt * peak > self.cfl * (1 + 1e-12):
            raise Rejected('cfl')
        fstat = np.zeros(nf, np.int32)
        o['hllc_limit'](n, nf, self.dry, work, theta, g['owner'], g['neighbour'], g['selected'], fl, fr, fstat)
        nxt = np.zeros((3, n)); stat = np.zeros(n)
        o['hllc_update'](n, nf, float(dt), self.dry, work, g['area'], g['length'], g['off'], g['faces'],
                         g['signs'], fl, fr, fstat, nxt, stat)
        if stat.any():
            raise Rejected('positivity')
        result = nxt.copy()
        if self.turb is not None:
            result += dt * self.momentum(qhalf, z)
            if not np.isfinite(result).all():
                raise Rejected('momentum')
        result = self.friction(result, .5 * dt)
        self.detail = dict(work=work, grad=grad, half=half, qhalf=qhalf, fl=fl, fr=fr, speed=speed, theta=theta,
                           frequency=freq, ext=ext, trace=trace)
        return result, fl[0].copy(), peak

    single = spatial_step

    # -- HLLCHydraulics.advance -------------------------------------------------
    def advance(self, q, z, dt, time=0.):
        elapsed = 0.; hint = self.hint or dt
        water = np.zeros(self.nf); steps = []; retries = 0; kinds = []; cfls = []
        q = np.array(q, dtype=np.float64, order='C')
        while not steps or elapsed < dt - 1e-13 * max(1., dt):
            trial = min(hint, dt - elapsed)
            truncated = trial < hint * (1 - 1e-12)
            try:
                result, flux_, frequency = self.single(q, z, trial, time + elapsed)
            except Rejected as exc:
                kinds.append(exc.kind)
                retries += 1
                hint = min(.5 * trial, .98 * self.cfl / self.peak) if self.peak > 0 else .5 * trial
                if retries > 100 or hint < dt * 2 ** -20:
                    raise
                continue
            q = result; water += trial * flux_; elapsed += trial; steps.append(trial); cfls.append(trial * frequency)
            growth = hint if truncated else 2 * trial
            hint = min(growth, .98 * self.cfl / frequency) if frequency > 0 else growth
        self.hint = float(hint)
        return dict(q=q, water=water / dt, substeps=len(steps), retries=retries, rejections=kinds, steps=steps,
                    hint=hint, cfl_max=max(cfls))

    def run(self, q, z, nsteps, time=0., ledger=None):
        """nsteps accepted substeps of the hint alone (no interval end): the
        advance control with an unbounded interval.  Returns q, time, the
        boundary ledger (inward water per boundary) and the rejections."""
        q = np.array(q, dtype=np.float64, order='C')
        hint = self.hint
        if not hint:
            raise ValueError('run needs a step hint')
        led = np.zeros(self.nb) if ledger is None else ledger
        acc = 0; kinds = []
        bo, bf, L = self.mesh.boundary_offset, self.mesh.boundary_faces, self.mesh.length
        while acc < nsteps:
            trial = hint
            try:
                result, flux_, frequency = self.single(q, z, trial, time)
            except Rejected as exc:
                kinds.append(exc.kind)
                hint = min(.5 * trial, .98 * self.cfl / self.peak) if self.peak > 0 else .5 * trial
                if len(kinds) > 100 * (acc + 1):
                    raise
                continue
            for b in range(self.nb):
                s = 0.
                for f in bf[bo[b]:bo[b + 1]]:
                    s = s - trial * L[f] * flux_[f]
                led[b] += s
            q = result; time += trial; acc += 1
            growth = 2 * trial
            hint = min(growth, .98 * self.cfl / frequency) if frequency > 0 else growth
        self.hint = float(hint)
        return dict(q=q, time=time, ledger=led, rejections=kinds)
