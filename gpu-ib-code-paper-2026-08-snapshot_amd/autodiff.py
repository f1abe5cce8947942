"""Reverse-mode derivatives of the oscillating-sphere time step.

The forward step is the same sequence of kernels as ``HipDriver.step``
(``ib_driver_np.py:412-532``, accumulated explicit IB forcing) and the load of
``HipLoadCorrection.advance`` (``ib_interior.py:268-308``), expressed as
``torch.autograd.Function`` nodes whose backward passes are the hand-written
adjoint kernels of ``csrc/adjoint.hip``:

* ``CalcVel``: U = curl(s) with the inlet values set, vc = interp_f2c(U) plus
  the ghost rows; backward = ghost fold + interp^T + curl^T (= rot).
* ``MomentumResidual``: base - a vc - kc conv(U, vc) + kd diff(vc) + kf F;
  backward is bilinear in (U, vc).
* ``CurlRhs``: rot(integr_c2f(res_c)); backward = integr_c2f^T curl.
* ``Helmholtz``: x = s + G P res_e (the PCG correction of ``helm_solve``).
  The elliptic operator is symmetric, so G is too and the backward pass is
  the same PCG, applied to the gradient's component in the operator's range.
* ``Spread`` / ``Interp``: Roma-Peskin transfer; interpolation is spreading's
  transpose with the weight hh^3 in place of ds^3, and the marker-position
  derivative uses d(delta)/dr of the same stencil.
* ``FluidForce`` / ``InteriorMomentum``: the two reductions of the load.

``PeriodForceObjective`` runs K steps from rest, J = mean over the steps of
``body_force_x``, and computes dJ/dA and dJ/dX0 (oscillation amplitude and
reference marker positions) with checkpoint recomputation: the forward pass
keeps the step state (s, s_old, held Lagrangian force) every
``checkpoint_every`` steps, and the backward pass recomputes one segment at a
time with the graph recorded, back-propagating the adjoint state across
segments.  The fluid viscosity is held fixed when A changes.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import _native as N
from .core import dev_stream, ptr

TAPS = 27


def _f64(shape, dev):
    return torch.zeros(shape, dtype=torch.float64, device=dev)


class StepContext:
    """Device tables shared by the differentiable step (one runtime, float64)."""

    def __init__(self, rt, tol1=1e-10, tol2=1e-9, cg_max=None):
        self.rt = rt
        core, drv = rt.core, rt.driver
        if drv.precision != "float64":
            raise ValueError("the adjoint path runs in float64")
        if drv.ib is None or drv.motion is None or drv.ib_force_mode != "accumulated_explicit":
            raise ValueError("the adjoint path covers a moving body with accumulated explicit forcing")
        self.core, self.drv, self.ib = core, drv, drv.ib
        self.dev = dev = core.device
        self.tol1, self.tol2 = float(tol1), float(tol2)
        self.cg_max = int(cg_max or drv.cg_max_iterations)
        self.dt = float(rt.dt)
        self.inner = int(rt.inner_iterations)
        fg = np.full(core.nface, -1, np.int32)
        fg[core.tables.bf] = np.arange(core.nghost, dtype=np.int32)
        self.face_ghost = torch.from_numpy(fg).to(dev)
        keep = np.ones(core.nedge)
        keep[np.asarray(drv.inlet_edges, bool)] = 0.0
        self.inlet_keep = torch.from_numpy(keep).to(dev)
        self.not_dir = torch.from_numpy((~np.asarray(drv.dir_mask, bool)).astype(np.float64)).to(dev)
        self.vol_ci = torch.from_numpy(np.asarray(core.vol_ci, np.float64)[:, None].copy()).to(dev)
        ib = self.ib
        self.nx, self.ny, self.nz = (int(v) for v in ib.n)
        self.pos_ref = [float(v) for v in ib.pos_ref]
        self.dx, self.hh = float(ib.dx), float(ib.hh)
        self.ds3 = drv._ds3
        self.ones_m = torch.ones_like(self.ds3)
        mo = drv.motion
        self.axis = torch.from_numpy(np.asarray(mo.axis, np.float64)).to(dev)
        self.omega, self.phase = float(mo._w), float(mo.phase)
        self.k_ib = float(drv.rou * drv.ib_alpha)
        lc = rt.load_correction
        self.lc = lc
        c0 = np.asarray(rt.config["ib"]["center"], np.float64)
        self.center0 = torch.from_numpy(c0).to(dev)
        self.cg_iterations = 0
        self.adjoint_cg_iterations = 0
        self.projection_iterations = 0
        self.log = None                      # optional callable(str): one line per adjoint solve
        nb = C.c_int64(0)
        N.call("pyib_ib_sort_temp_bytes", C.c_int64(mo.ref.shape[0] * TAPS), C.byref(nb))
        self._sort_bytes = max(int(nb.value), 1)

    # ---------------------------------------------------------- motion
    def motion(self, A, X0, t):
        """Oscillation (ib_kinematics.py:31-40) in torch: markers, velocity, sphere centre."""
        sv = float(np.sin(self.omega * t + self.phase))
        cv = float(np.cos(self.omega * t + self.phase))
        disp = (A * sv) * self.axis
        X = X0 + disp
        vel = ((A * self.omega) * cv * self.axis).expand(X0.shape[0], 3)
        return X, vel, self.center0 + disp

    def stencil(self, X):
        """Cells, weights and the cell-sorted (key, value) lists of the 27-tap stencil of X."""
        X = X.detach().contiguous()
        m = int(X.shape[0])
        n = m * TAPS
        dev = self.dev
        st = {"m": m, "X": X,
              "cell": torch.empty(n, dtype=torch.int32, device=dev), "w": torch.empty(n, dtype=torch.float64,
                                                                                       device=dev),
              "keys": torch.empty(n, dtype=torch.int32, device=dev), "vals": torch.empty(n, dtype=torch.int32,
                                                                                          device=dev)}
        kt = torch.empty(n, dtype=torch.int32, device=dev)
        vt = torch.empty(n, dtype=torch.int32, device=dev)
        tmp = torch.empty(self._sort_bytes, dtype=torch.uint8, device=dev)
        pr = self.pos_ref
        N.call("pyib_ib_stencil", C.c_int64(m), ptr(X), pr[0], pr[1], pr[2], self.dx, self.hh, self.nx, self.ny,
               self.nz, ptr(st["cell"]), ptr(st["w"]), dev_stream())
        N.call("pyib_ib_sort", C.c_int64(m), ptr(st["cell"]), ptr(kt), ptr(vt), ptr(st["keys"]), ptr(st["vals"]),
               ptr(tmp), C.c_int64(tmp.numel()), dev_stream())
        return st

    def dweights(self, st):
        if "dw" not in st:
            dw = torch.empty((st["m"] * TAPS, 3), dtype=torch.float64, device=self.dev)
            pr = self.pos_ref
            N.call("pyib_ib_dweights", C.c_int64(st["m"]), ptr(st["X"]), pr[0], pr[1], pr[2], self.dx, self.hh,
                   self.nx, self.ny, self.nz, ptr(dw), dev_stream())
            st["dw"] = dw
        return st["dw"]


# ---------------------------------------------------------------- Functions
class CalcVel(torch.autograd.Function):
    @staticmethod
    def forward(ctx, K, s):
        core, drv = K.core, K.drv
        s = s.contiguous()
        s_out = torch.empty_like(s)
        U = torch.empty(core.nface, dtype=torch.float64, device=K.dev)
        vc = torch.empty((core.ncell + core.nghost, 3), dtype=torch.float64, device=K.dev)
        lid = (C.c_double * 3)(*drv._lid)
        N.call("pyib_calc_vel", core.struct(), ptr(s), ptr(s_out), ptr(U), ptr(vc), C.cast(lid, C.c_void_p),
               dev_stream())
        ctx.K = K
        return s_out, U, vc

    @staticmethod
    def backward(ctx, g_sout, g_U, g_vc):
        K = ctx.K
        core = K.core
        gvc = g_vc.contiguous().clone() if g_vc is not None else _f64((core.ncell + core.nghost, 3), K.dev)
        gU_in = g_U.contiguous() if g_U is not None else None
        gU = torch.empty(core.nface, dtype=torch.float64, device=K.dev)
        gs = torch.empty(core.nedge, dtype=torch.float64, device=K.dev)
        N.call("pyib_calc_vel_adjoint", core.struct(), ptr(K.face_ghost), ptr(gvc), ptr(gU_in), ptr(gU), ptr(gs),
               dev_stream())
        if g_sout is not None:
            gs = gs + g_sout * K.inlet_keep
        return None, gs


class MomentumResidual(torch.autograd.Function):
    """out = [base - a vc | a vc] - kc conv(U, vc) + kd diff(vc) + kf F  (ib_driver_np.py:458-505)."""

    @staticmethod
    def forward(ctx, K, U, vc, base, force, a_vc, k_conv, k_diff, k_force):
        core = K.core
        out = torch.empty((core.ncell, 3), dtype=torch.float64, device=K.dev)
        N.call("pyib_momentum_residual", core.struct(), ptr(U), ptr(vc), float(K.drv.visc), ptr(base), float(a_vc),
               float(k_conv), float(k_diff), ptr(force), float(k_force), ptr(out), dev_stream())
        ctx.K = K
        ctx.coef = (float(a_vc), float(k_conv), float(k_diff), float(k_force))
        ctx.has_base, ctx.has_force = base is not None, force is not None
        ctx.save_for_backward(U, vc)
        return out

    @staticmethod
    def backward(ctx, g):
        K = ctx.K
        core = K.core
        U, vc = ctx.saved_tensors
        a_vc, kc, kd, kf = ctx.coef
        g = g.contiguous()
        gvc = _f64((core.ncell + core.nghost, 3), K.dev)
        gU = _f64(core.nface, K.dev)
        N.call("pyib_momentum_residual_adjoint", core.struct(), ptr(K.face_ghost), ptr(U), ptr(vc),
               float(K.drv.visc), int(ctx.has_base), a_vc, kc, kd, ptr(g), ptr(gvc), ptr(gU), dev_stream())
        g_base = g if ctx.has_base else None
        g_force = kf * g if ctx.has_force else None
        return None, gU, gvc, g_base, g_force, None, None, None, None


class CurlRhs(torch.autograd.Function):
    @staticmethod
    def forward(ctx, K, res_c):
        core = K.core
        out = torch.empty(core.nedge, dtype=torch.float64, device=K.dev)
        N.call("pyib_curl_rhs", core.struct(), core.work_struct(), ptr(res_c.contiguous()), ptr(out), dev_stream())
        ctx.K = K
        return out

    @staticmethod
    def backward(ctx, g):
        K = ctx.K
        core = K.core
        out = torch.empty((core.ncell, 3), dtype=torch.float64, device=K.dev)
        N.call("pyib_curl_rhs_adjoint", core.struct(), core.work_struct(), ptr(g.contiguous()), ptr(out),
               dev_stream())
        return None, out


def _bt(K, w):
    """B^T w = P rot(integr_c2f(vol_ci w)) with B = interp_f2c curl P, so that A = B^T (W V) B."""
    core = K.core
    return core.rot(core.integr_c2f((w * K.vol_ci).contiguous())) * K.not_dir


def _b_dinv(K, e):
    """B D^-1 e = interp_f2c(curl(prec e)) (prec is zero on the Dirichlet edges)."""
    core = K.core
    return core.interp_f2c(core.curl((e * K.drv.prec).contiguous()))


def project_on_range(K, g, tol=1e-13, maxit=20000, check=25):
    """Component of g in range(A) = range(B^T), in the metric of the Jacobi-PCG:
    g_r = B^T w with (B D^-1 B^T) w = B D^-1 g (a consistent cell-space system, plain CG).
    The CG scalars stay on the device; the host reads |r|^2 every ``check`` iterations."""
    b = _b_dinv(K, g)
    w = torch.zeros_like(b)
    r = b.clone()
    p = r.clone()
    rr = (r * r).sum()
    stop = (tol ** 2) * float(rr)
    it = 0
    while it < maxit:
        if it % check == 0 and float(rr) <= stop:
            break
        q = _b_dinv(K, _bt(K, p))
        alpha = rr / (p * q).sum()
        w.addcmul_(p, alpha)
        r.addcmul_(q, alpha, value=-1.0)
        rr_new = (r * r).sum()
        p.mul_(rr_new / rr).add_(r)
        rr = rr_new
        it += 1
    K.projection_iterations += it
    return _bt(K, w)


class Helmholtz(torch.autograd.Function):
    """x = helm_solve(res_e, s) = s + G P res_e (ib_driver_np.py:311-410).

    The Jacobi-PCG from a zero correction is CG on D^-1/2 A D^-1/2 (D = diag, Dirichlet
    rows removed), so its map is G = D^-1/2 Ahat^+ D^-1/2: symmetric, and zero on the null
    space of A (gauge gradients and the face modes the cell average does not see).  The
    forward right-hand side rot(integr_c2f(.)) always lies in range(A); the incoming
    gradient does not (the convective flux depends on U directly), and CG on an
    inconsistent right-hand side diverges.  The backward pass therefore first takes the
    component of the gradient in range(A) (``project_on_range``, which G does not change)
    and then applies the same PCG to it.
    """

    @staticmethod
    def forward(ctx, K, res_e, s):
        x, m = K.drv.helm_solve(res_e.contiguous(), s.contiguous(), K.dt, tol1=K.tol1, tol2=K.tol2, cg_max=K.cg_max)
        K.cg_iterations += int(m)
        ctx.K = K
        return x

    @staticmethod
    def backward(ctx, g):
        K = ctx.K
        g = g.contiguous()
        p0 = K.projection_iterations
        g_r = project_on_range(K, g, tol=K.tol2)
        g_res, m = K.drv.helm_solve(g_r, torch.zeros_like(g), K.dt, tol1=K.tol1, tol2=K.tol2, cg_max=K.cg_max)
        K.adjoint_cg_iterations += int(m)
        if K.log:
            K.log(f"adjoint Helmholtz solve: projection CG {K.projection_iterations - p0}, PCG {int(m)}")
        return None, g_res, g


class Spread(torch.autograd.Function):
    """E[c] = scale * sum_(m,t) f[m] w[m,t] ds3[m]  (ib_immersed.py:95-111)."""

    @staticmethod
    def forward(ctx, K, st, f, X, scale):
        f = f.contiguous()
        out = _f64((K.core.ncell, 3), K.dev)
        N.call("pyib_ib_spread", N.F64, C.c_int64(st["m"]), ptr(st["keys"]), ptr(st["vals"]), ptr(st["w"]), ptr(f),
               ptr(K.ds3), float(scale), 0, ptr(out), dev_stream())
        ctx.K, ctx.st, ctx.scale = K, st, float(scale)
        ctx.save_for_backward(f)
        return out

    @staticmethod
    def backward(ctx, g):
        K, st, scale = ctx.K, ctx.st, ctx.scale
        (f,) = ctx.saved_tensors
        g = g.contiguous()
        gf = torch.empty((st["m"], 3), dtype=torch.float64, device=K.dev)
        N.call("pyib_ib_gather", C.c_int64(st["m"]), ptr(st["cell"]), ptr(st["w"]), ptr(g), ptr(K.ds3), ptr(gf),
               dev_stream())
        gX = _f64((st["m"], 3), K.dev)
        N.call("pyib_ib_marker_grad", C.c_int64(st["m"]), ptr(st["cell"]), ptr(K.dweights(st)), ptr(g), ptr(f),
               ptr(K.ds3), scale, ptr(gX), dev_stream())
        return None, None, scale * gf, gX, None


class Interp(torch.autograd.Function):
    """u[m] = hh^3 sum_t vc[cell] w[m,t]  (ib_immersed.py:83-93)."""

    @staticmethod
    def forward(ctx, K, st, vc, X):
        vc = vc.contiguous()
        out = torch.empty((st["m"], 3), dtype=torch.float64, device=K.dev)
        N.call("pyib_ib_interp", N.F64, C.c_int64(st["m"]), ptr(st["cell"]), ptr(st["w"]), K.hh ** 3, ptr(vc),
               ptr(out), dev_stream())
        ctx.K, ctx.st, ctx.nrows = K, st, int(vc.shape[0])
        ctx.save_for_backward(vc)
        return out

    @staticmethod
    def backward(ctx, g):
        K, st = ctx.K, ctx.st
        (vc,) = ctx.saved_tensors
        g = g.contiguous()
        gvc = _f64((ctx.nrows, 3), K.dev)
        N.call("pyib_ib_spread", N.F64, C.c_int64(st["m"]), ptr(st["keys"]), ptr(st["vals"]), ptr(st["w"]), ptr(g),
               ptr(K.ones_m), K.hh ** 3, 0, ptr(gvc), dev_stream())
        gX = _f64((st["m"], 3), K.dev)
        N.call("pyib_ib_marker_grad", C.c_int64(st["m"]), ptr(st["cell"]), ptr(K.dweights(st)), ptr(vc), ptr(g),
               ptr(None), K.hh ** 3, ptr(gX), dev_stream())
        return None, None, gvc, gX


class FluidForce(torch.autograd.Function):
    """F_fluid = sum over the force-support cells of F_E vol (ib_interior.py:290-296)."""

    @staticmethod
    def forward(ctx, K, E):
        lc = K.lc
        out = _f64(8, K.dev)
        part = _f64(8 * N.MAX_PARTIALS, K.dev)
        N.call("pyib_fluid_load", N.F64, C.c_int64(lc._fidx.numel()), ptr(lc._fidx), ptr(lc._fpos), ptr(lc._fvol),
               ptr(E.contiguous()), ptr(part), ptr(out), dev_stream())
        ctx.K, ctx.shape = K, tuple(E.shape)
        return out[:3].clone()

    @staticmethod
    def backward(ctx, g):
        K = ctx.K
        lc = K.lc
        gv = g.detach().cpu().numpy()
        gE = _f64(ctx.shape, K.dev)
        N.call("pyib_fluid_load_adjoint", C.c_int64(lc._fidx.numel()), ptr(lc._fidx), ptr(lc._fvol), float(gv[0]),
               float(gv[1]), float(gv[2]), ptr(gE), dev_stream())
        return None, gE


class InteriorMomentum(torch.autograd.Function):
    """P = rho sum vc wv(center) over the smoothed sphere (ib_interior.py:57-70)."""

    @staticmethod
    def forward(ctx, K, vc, center):
        lc = K.lc
        c = center.detach().cpu().numpy()
        out = _f64(16, K.dev)
        part = _f64(8 * N.MAX_PARTIALS, K.dev)
        vc = vc.contiguous()
        N.call("pyib_interior_momenta", N.F64, C.c_int64(lc._cand.numel()), ptr(lc._cand), ptr(lc._cpos),
               ptr(lc._cvol), ptr(vc), float(c[0]), float(c[1]), float(c[2]), lc.radius, lc.smoothing_width,
               lc.density, int(lc._fixed), ptr(lc._fwv), ptr(part), ptr(out), dev_stream())
        ctx.K, ctx.c = K, c
        ctx.save_for_backward(vc)
        return lc.density * out[:3]

    @staticmethod
    def backward(ctx, g):
        K, c = ctx.K, ctx.c
        lc = K.lc
        (vc,) = ctx.saved_tensors
        gv = g.detach().cpu().numpy()
        gvc = torch.zeros_like(vc)
        part = _f64(8 * N.MAX_PARTIALS, K.dev)
        gc = _f64(3, K.dev)
        if lc._fixed:
            raise ValueError("a fixed body has no centre dependence to differentiate")
        N.call("pyib_interior_momentum_adjoint", C.c_int64(lc._cand.numel()), ptr(lc._cand), ptr(lc._cpos),
               ptr(lc._cvol), ptr(vc), float(c[0]), float(c[1]), float(c[2]), lc.radius, lc.smoothing_width,
               lc.density, float(gv[0]), float(gv[1]), float(gv[2]), ptr(gvc), ptr(part), ptr(gc), dev_stream())
        return None, gvc, gc


# ------------------------------------------------------------------- step
def step(K, A, X0, s, s_old, fL, n):
    """Step n (time (n-1) dt -> n dt) of ``HipDriver.step`` with accumulated explicit forcing.
    Returns (s, s_old, fL, vc, E, center)."""
    core, drv = K.core, K.drv
    dt, rou, ac, ad = K.dt, drv.rou, drv.ac, drv.ad
    t = n * dt
    X, vel, center = K.motion(A, X0, t)
    st = K.stencil(X)
    E = Spread.apply(K, st, fL, X, 1.0)
    _, U, vc = CalcVel.apply(K, s)
    res0 = MomentumResidual.apply(K, U, vc, None, None, rou, (1 - ac) * dt * rou, (1 - ad) * dt, 0.0)
    guess = s + ((s - s_old) * dt) / dt
    s_prev = s
    s_c, U, vc = CalcVel.apply(K, guess)
    for i in range(1, K.inner + 1):
        res = MomentumResidual.apply(K, U, vc, res0, E, rou, ac * dt * rou, ad * dt, dt)
        rhs = CurlRhs.apply(K, res)
        s_c = Helmholtz.apply(K, rhs, s_c)
        s_c, U, vc = CalcVel.apply(K, s_c)
        if i == 1:
            ui = Interp.apply(K, st, vc, X)
            f_lag = (vel - ui) / dt
            E = E + Spread.apply(K, st, f_lag, X, K.k_ib)
            fL = fL + K.k_ib * f_lag
    return s_c, s_prev, fL, vc, E, center


class PeriodForceObjective:
    """J(A, X0) = (1/K) sum_{n=1..K} body_force_x(n) from rest, and its gradient."""

    def __init__(self, rt, steps, checkpoint_every=20, tol1=1e-10, tol2=1e-9, cg_max=None, log=None):
        self.log = log                       # optional callable(str): one line per segment
        self.K = StepContext(rt, tol1=tol1, tol2=tol2, cg_max=cg_max)
        self.K.log = log
        self.steps = int(steps)
        self.every = max(1, int(checkpoint_every))
        mo = rt.driver.motion
        dev = self.K.dev
        self.A0 = float(mo.A)
        self.X00 = torch.from_numpy(np.ascontiguousarray(mo.ref, np.float64)).to(dev)

    def _initial(self, A, X0):
        K = self.K
        core = K.core
        s0 = torch.zeros(core.nedge, dtype=torch.float64, device=K.dev)
        fL = torch.zeros((X0.shape[0], 3), dtype=torch.float64, device=K.dev)
        return s0, s0.clone(), fL

    def _segment(self, A, X0, state, n0, n1):
        """Steps n0+1..n1 from ``state``; returns (sum of body_force_x, end state)."""
        K = self.K
        s, s_old, fL = state
        _, _, vc0 = CalcVel.apply(K, s)
        _, _, c_prev = K.motion(A, X0, n0 * K.dt)
        P_prev = InteriorMomentum.apply(K, vc0, c_prev)
        acc = torch.zeros((), dtype=torch.float64, device=K.dev)
        for n in range(n0 + 1, n1 + 1):
            s, s_old, fL, vc, E, center = step(K, A, X0, s, s_old, fL, n)
            ff = FluidForce.apply(K, E)
            P = InteriorMomentum.apply(K, vc, center)
            acc = acc + (-ff[0] + (P[0] - P_prev[0]) / K.dt)
            P_prev = P
        return acc, (s, s_old, fL)

    def value(self, A=None, X0=None):
        """J only (no graph)."""
        A = torch.tensor(self.A0 if A is None else float(A), dtype=torch.float64, device=self.K.dev)
        X0 = self.X00 if X0 is None else X0
        with torch.no_grad():
            acc, _ = self._segment(A, X0, self._initial(A, X0), 0, self.steps)
        if self.log:
            self.log(f"J evaluated: CG {self.K.cg_iterations}")
        return float(acc.item()) / self.steps

    def value_and_grad(self, A=None, X0=None):
        """(J, dJ/dA, dJ/dX0) with checkpoint recomputation."""
        K = self.K
        dev = K.dev
        A_val = self.A0 if A is None else float(A)
        X0_val = self.X00 if X0 is None else X0
        A_c = torch.tensor(A_val, dtype=torch.float64, device=dev)
        bounds = list(range(0, self.steps, self.every)) + [self.steps]
        states = []
        total = 0.0
        with torch.no_grad():
            state = self._initial(A_c, X0_val)
            for a, b in zip(bounds[:-1], bounds[1:]):
                states.append(tuple(x.clone() for x in state))
                acc, state = self._segment(A_c, X0_val, state, a, b)
                total += float(acc.item())
                if self.log:
                    self.log(f"forward steps {a + 1}-{b}: CG {K.cg_iterations}")
        gA = torch.zeros((), dtype=torch.float64, device=dev)
        gX = torch.zeros_like(X0_val)
        g_state = None
        for seg in range(len(states) - 1, -1, -1):
            a, b = bounds[seg], bounds[seg + 1]
            A_l = torch.tensor(A_val, dtype=torch.float64, device=dev, requires_grad=True)
            X_l = X0_val.detach().clone().requires_grad_(True)
            st0 = tuple(x.detach().clone().requires_grad_(True) for x in states[seg])
            with torch.enable_grad():
                acc, st1 = self._segment(A_l, X_l, st0, a, b)
                outs, grads = [acc / self.steps], [torch.ones((), dtype=torch.float64, device=dev)]
                if g_state is not None:
                    for o, g in zip(st1, g_state):
                        if g is not None and o.requires_grad:
                            outs.append(o)
                            grads.append(g)
                ins = [A_l, X_l] + list(st0)
                res = torch.autograd.grad(outs, ins, grads, allow_unused=True)
            gA = gA + (res[0] if res[0] is not None else 0.0)
            gX = gX + (res[1] if res[1] is not None else 0.0)
            g_state = res[2:]
            if self.log:
                self.log(f"backward steps {b}-{a + 1}: adjoint CG {K.adjoint_cg_iterations}, "
                         f"projection CG {K.projection_iterations}")
        return total / self.steps, float(gA.item()), gX


def central_difference(obj, direction_A=0.0, direction_X=None, eps=1.0):
    """(J(p + eps d) - J(p - eps d)) / (2 eps) along (direction_A, direction_X)."""
    X0 = obj.X00
    dX = torch.zeros_like(X0) if direction_X is None else direction_X
    jp = obj.value(obj.A0 + eps * direction_A, X0 + eps * dX)
    jm = obj.value(obj.A0 - eps * direction_A, X0 - eps * dX)
    return (jp - jm) / (2.0 * eps)


def direction(m, norm, seed=20261203, device="cuda"):
    """Seeded random marker-position direction of the given Euclidean norm."""
    d = np.random.default_rng(seed).standard_normal((m, 3))
    d *= norm / np.linalg.norm(d)
    return torch.from_numpy(d).to(device)


__all__ = ["StepContext", "CalcVel", "MomentumResidual", "CurlRhs", "Helmholtz", "Spread", "Interp", "FluidForce",
           "InteriorMomentum", "step", "PeriodForceObjective", "central_difference", "direction"]
