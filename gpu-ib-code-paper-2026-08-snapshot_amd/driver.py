"""Crank-Nicolson stream-function stepper on the GPU (``--device hip``).

``HipDriver`` mirrors ``src/pyib/solver/ib_driver_np.py`` (class ``Driver``):
the same boundary rules, the same step structure (``step`` 412-532) and the
same Jacobi-PCG contract of ``helm_solve`` (311-410): warm start from ``s``,
the mixed stopping test ``z.r < |x|^2 tol1^2 + tol2^2 eta0`` evaluated only
every ``cg_check_interval`` iterations, guarded alpha/beta, and the
true-residual epilogue that decides convergence.  All field arithmetic runs in
gfx950 kernels; the host only reads the scalar stopping test at the check
cadence and the four closing diagnostics of each solve.

``precision="mixed"`` replaces the float64 CG by defect correction: float32
PCG sweeps on the float32 operator solve for the correction, every outer
residual is recomputed with the float64 operator, and the outer loop stops on
the same float64 mixed criterion as the plain path.  Dot products are always
accumulated in float64.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import _native as N
from .core import dev_stream, ptr
from .tables import ghost_kinds


class HipDriver:
    cg_max_iterations = 1000
    fail_on_cg_nonconvergence = False
    REFERENCE_TOL1, REFERENCE_TOL2 = 1e-10, 1e-9    # helm_solve's tolerances (ib_driver_np.py:311)
    max_residual_replacements = 8
    residual_replacements_last_solve = 0
    cg_timer = None        # optional: object with begin()/end(event, (n, precision)) around PCG blocks

    def __init__(self, core, mesh, visc, rou=1.0, ac=0.5, ad=0.5, *, cg_max_iterations=1000,
                 fail_on_cg_nonconvergence=False, cg_check_interval=20, precision="float64",
                 mixed_inner_tol=0.1, mixed_max_outer=1000, structured=True, layout=None):
        self.core = core
        self.grid = None
        if structured:
            from .grid import HipGrid
            self.grid = HipGrid(core, precisions=("float64", "float32") if precision == "mixed" else ("float64",),
                                layout=layout)
        self.device = core.device
        self.visc, self.rou, self.ac, self.ad = visc, rou, ac, ad
        self.precision = precision
        if precision not in ("float64", "mixed"):
            raise ValueError("precision must be 'float64' or 'mixed'")
        if precision == "mixed" and "float32" not in core._coef:
            core._build_coef("float32")
            core._mesh["float32"] = core._make_struct("float32")
        self.mixed_inner_tol = float(mixed_inner_tol)
        self.mixed_max_outer = int(mixed_max_outer)
        self.bc_edge = np.asarray(mesh.bc_edge)
        self.bc_face = np.asarray(mesh.bc_face)
        self.bf = core.bf
        self.nghost = core.nghost
        self.cavity = False
        self._lid = np.zeros(3)
        self.inlet_edges = self.bc_edge == 1
        self.dir_mask = np.isin(self.bc_edge, [3, 4, 5, 6])
        self._install_boundary(ghost_kinds(core.geo, self.bc_face, self.bf, False))
        self.ib = None
        self.ib_alpha = 1.0
        self.ib_implicit = False
        self.motion = None
        self.time = 0.0
        self._step_count = 0
        self.cg_max_iterations = self._positive_cg_limit(cg_max_iterations)
        self.fail_on_cg_nonconvergence = bool(fail_on_cg_nonconvergence)
        self.cg_check_interval = int(cg_check_interval)
        if self.cg_check_interval <= 0:
            raise ValueError("cg_check_interval must be a positive integer")
        self.last_helm_m = 0
        self.helmholtz_iterations_last_step = 0
        self.helmholtz_iterations_sum_last_step = 0
        self.helmholtz_outer_last_step = 0
        self.last_helm_rel_residual = 0.0
        self.last_helm_tolerance_ratio = 0.0
        self.last_helm_converged = True
        self.helmholtz_relative_residual_last_step = 0.0
        self.helmholtz_tolerance_ratio_last_step = 0.0
        self.helmholtz_converged_last_step = True
        self.last_face_flux = None
        self.last_momentum_residual = None
        self.last_step_dt = None
        self._alloc()

    # ------------------------------------------------------------ setup
    @staticmethod
    def _positive_cg_limit(value):
        if isinstance(value, (bool, np.bool_)) or not isinstance(value, (int, np.integer)) or value <= 0:
            raise ValueError("cg_max_iterations must be a positive integer")
        return int(value)

    def _install_boundary(self, gh_kind):
        core = self.core
        inl = np.flatnonzero(self.inlet_edges)
        vals = core.epos[inl, 1] * core.te[inl, 2]
        core.set_boundary(gh_kind, self.dir_mask, inl, vals)
        self.prec_host = core.preconditioner(self.dir_mask)
        self.prec = core.tensor(self.prec_host)
        if "float32" in core._coef:
            self.prec32 = self.prec.float()
        if self.grid is not None:
            self.grid.set_dir_mask(self.dir_mask)
            pi = np.zeros(self.grid.nedge)
            pi[self.grid.lay.edge_map] = self.prec_host
            self.prec_i = core.tensor(pi)
            self.prec_i32 = self.prec_i.float()
            self.grid.set_preconditioner(pi, {"float64": self.prec_i, "float32": self.prec_i32})
            if hasattr(self, "_cgi"):
                self._cgi.prec = ptr(self.prec_i)
                if hasattr(self, "_cgi32"):
                    self._cgi32.prec = ptr(self.prec_i32)

    def _alloc(self):
        dev, c = self.device, self.core
        e = lambda dt=torch.float64: torch.empty(c.nedge, dtype=dt, device=dev)
        self._U = torch.empty(c.nface, dtype=torch.float64, device=dev)
        self._x, self._r, self._p, self._w, self._tr, self._tmp = e(), e(), e(), e(), e(), e()
        self._scal = torch.zeros(N.NSCAL, dtype=torch.float64, device=dev)
        self._out2 = torch.zeros(2, dtype=torch.float64, device=dev)
        self._cg64 = N.CG(ptr(self._x), ptr(self._r), ptr(self._p), ptr(self._w), ptr(self.prec), ptr(self._scal))
        if self.precision == "mixed":
            f = lambda: e(torch.float32)
            self._x32, self._r32, self._p32, self._w32, self._z32 = f(), f(), f(), f(), f()
            self._scal32 = torch.zeros(N.NSCAL, dtype=torch.float64, device=dev)
            self._cg32 = N.CG(ptr(self._x32), ptr(self._r32), ptr(self._p32), ptr(self._w32), ptr(self.prec32),
                              ptr(self._scal32))
        if self.grid is not None:
            gi = lambda dt=torch.float64: torch.zeros(self.grid.nedge, dtype=dt, device=dev)
            self._xi, self._ri, self._pi, self._wi, self._tri, self._tmpi = gi(), gi(), gi(), gi(), gi(), gi()
            self._resi, self._si = gi(), gi()
            self._cgi = N.CG(ptr(self._xi), ptr(self._ri), ptr(self._pi), ptr(self._wi), ptr(self.prec_i),
                             ptr(self._scal))
            if self.precision == "mixed":
                g32 = lambda: gi(torch.float32)
                self._xi32, self._ri32, self._pi32, self._wi32, self._zi32 = g32(), g32(), g32(), g32(), g32()
                self._cgi32 = N.CG(ptr(self._xi32), ptr(self._ri32), ptr(self._pi32), ptr(self._wi32),
                                   ptr(self.prec_i32), ptr(self._scal32))

    def set_cavity(self, U_lid=1.0, spanwise_zerograd=True):
        """ib_driver_np.py:155-177: psi = 0 on every wall edge, ghosts by face position."""
        self.cavity = True
        self.U_lid = float(U_lid)
        self._lid = np.asarray([U_lid, 0.0, 0.0], dtype=np.float64)
        self.inlet_edges = np.zeros(self.core.nedge, dtype=bool)
        self.dir_mask = self.bc_edge > 0
        fp = self.core.fpos
        fy, fz = fp[self.bf, 1], fp[self.bf, 2]
        tol = 1e-6 + 1e-3 * (fy.max() - fy.min())
        lid = fy > fy.max() - tol
        zf = (fz < fz.min() + tol) | (fz > fz.max() - tol)
        if not spanwise_zerograd:
            zf = np.zeros_like(zf, dtype=bool)
        self._install_boundary(ghost_kinds(self.core.geo, self.bc_face, self.bf, True, lid, zf))
        if self.precision == "mixed":
            self._cg32.prec = ptr(self.prec32)
        self._cg64.prec = ptr(self.prec)
        return self

    def set_ib(self, ib, markers, ds, vel_desired, force_mode="replacement"):
        """ib_driver_np.py:179-201 (markers/vel/ds are host arrays; state lives on the device)."""
        if force_mode not in {"replacement", "accumulated_explicit", "fortran_accumulated"}:
            raise ValueError(f"unsupported IB force_mode: {force_mode!r}")
        dev = self.device
        self.ib = ib
        self.ib_markers_host = np.asarray(markers, dtype=np.float64).copy()
        self.ib_markers = torch.from_numpy(self.ib_markers_host.copy()).to(dev)
        self.ib_ds = np.asarray(ds, dtype=np.float64).copy()
        self._ds3 = torch.from_numpy(self.ib_ds ** 3).to(dev)
        self.ib_vel = torch.from_numpy(np.asarray(vel_desired, dtype=np.float64).copy()).to(dev)
        self.ib_force_mode = force_mode
        self.ib_force_E = torch.zeros((self.core.ncell, 3), dtype=torch.float64, device=dev)
        self.ib_force_L = torch.zeros_like(self.ib_vel)
        self.ib_force_update_count = 0
        self.ib_force_updates_last_step = 0
        self._marker_version = 0
        return self

    def set_markers(self, markers, vel):
        self.ib_markers_host = np.asarray(markers, dtype=np.float64).copy()
        self.ib_markers.copy_(torch.from_numpy(self.ib_markers_host))
        self.ib_vel.copy_(torch.from_numpy(np.asarray(vel, dtype=np.float64)))
        self._marker_version += 1

    def ib_state_dict(self):
        if self.ib is None:
            return None
        return {"force_mode": self.ib_force_mode, "force_E": self.ib_force_E.cpu().numpy().copy(),
                "force_L": self.ib_force_L.cpu().numpy().copy(),
                "force_update_count": int(self.ib_force_update_count)}

    def load_ib_state_dict(self, state):
        """ib_driver_np.py:214-239."""
        if self.ib is None:
            raise RuntimeError("set_ib must be called before loading IB state")
        mode = str(state["force_mode"])
        if mode != self.ib_force_mode:
            raise ValueError(f"IB force-mode mismatch: checkpoint={mode!r}, run={self.ib_force_mode!r}")
        fE = np.asarray(state["force_E"], dtype=np.float64)
        fL = np.asarray(state["force_L"], dtype=np.float64)
        if fE.shape != tuple(self.ib_force_E.shape):
            raise ValueError(f"Eulerian IB-force shape mismatch: {fE.shape} != {tuple(self.ib_force_E.shape)}")
        if fL.shape != tuple(self.ib_force_L.shape):
            raise ValueError(f"Lagrangian IB-force shape mismatch: {fL.shape} != {tuple(self.ib_force_L.shape)}")
        self.ib_force_E.copy_(torch.from_numpy(fE))
        self.ib_force_L.copy_(torch.from_numpy(fL))
        self.ib_force_update_count = int(state.get("force_update_count", 0))
        self.ib_force_updates_last_step = 0
        return self

    def set_motion(self, motion, t0=0.0):
        self.motion = motion
        self.time = float(t0)
        return self

    def enable_sparse(self, dt):
        """The HIP path is matrix-free; kept for interface parity (ib_driver_np.py:251-257)."""
        return self

    # ------------------------------------------------------- operators
    def calc_vel(self, s, s_out=None):
        """ib_driver_np.py:267-291 -> (s with inlet values, U, vc[(ncell+nghost),3])."""
        c = self.core
        s_out = torch.empty_like(s) if s_out is None else s_out
        U = torch.empty(c.nface, dtype=torch.float64, device=self.device)
        vc = torch.empty((c.ncell + self.nghost, 3), dtype=torch.float64, device=self.device)
        lid = (C.c_double * 3)(*self._lid)
        N.call("pyib_calc_vel", c.struct(), ptr(s), ptr(s_out), ptr(U), ptr(vc), C.cast(lid, C.c_void_p),
               dev_stream())
        return s_out, U, vc

    def momentum_residual(self, U, vc, base, a_vc, k_conv, k_diff, force=None, k_force=0.0):
        out = torch.empty((self.core.ncell, 3), dtype=torch.float64, device=self.device)
        N.call("pyib_momentum_residual", self.core.struct(), ptr(U), ptr(vc), float(self.visc), ptr(base),
               float(a_vc), float(k_conv), float(k_diff), ptr(force), float(k_force), ptr(out), dev_stream())
        return out

    def conv_diff(self, U, vc):
        """ib_driver_np.py:293-303 -> (conv, diff) (ncell,3)."""
        conv = self.momentum_residual(U, vc, None, 0.0, -1.0, 0.0)
        diff = self.momentum_residual(U, vc, None, 0.0, 0.0, 1.0)
        return conv, diff

    def curl_rhs(self, res_c):
        out = torch.empty(self.core.nedge, dtype=torch.float64, device=self.device)
        N.call("pyib_curl_rhs", self.core.struct(), self.core.work_struct(), ptr(res_c), ptr(out), dev_stream())
        return out

    # ------------------------------------------------------ Helmholtz CG
    def _finish(self, iterations, eta_f, l2_f, l2_init, err_tot, cg_max):
        """ib_driver_np.py:338-377."""
        vals = np.array([eta_f, l2_f, l2_init, err_tot], dtype=float)
        finite = bool(np.all(np.isfinite(vals)) and np.all(vals >= 0.0))
        self.last_helm_m = int(iterations)
        self.last_helm_rel_residual = (float(np.sqrt(l2_f / l2_init)) if finite and l2_init > 0.0
                                       else (0.0 if finite and l2_f == 0.0 else float("inf")))
        self.last_helm_tolerance_ratio = (float(np.sqrt(eta_f / err_tot)) if finite and err_tot > 0.0
                                          else (0.0 if finite and eta_f == 0.0 else float("inf")))
        self.last_helm_converged = bool(finite and eta_f <= err_tot)
        if self.fail_on_cg_nonconvergence and not self.last_helm_converged:
            self.helmholtz_iterations_last_step = max(self.helmholtz_iterations_last_step, int(iterations))
            self.helmholtz_relative_residual_last_step = max(self.helmholtz_relative_residual_last_step,
                                                             self.last_helm_rel_residual)
            self.helmholtz_tolerance_ratio_last_step = max(self.helmholtz_tolerance_ratio_last_step,
                                                           self.last_helm_tolerance_ratio)
            self.helmholtz_converged_last_step = False
            raise RuntimeError(
                "Helmholtz CG did not converge: "
                f"iterations={iterations}/{cg_max}, "
                f"true_relative_residual={self.last_helm_rel_residual:.6e}, "
                f"mixed_tolerance_ratio={self.last_helm_tolerance_ratio:.6e}")

    def _pcg_loop(self, mesh, ws, cg, scal, cg_max, stop_value, dt, grid=False):
        """Run PCG iterations with the host test every ``cg_check_interval``; returns m."""
        check = self.cg_check_interval
        hdt = 0.5 * float(dt)
        timer = self.cg_timer
        prec = "float32" if cg is getattr(self, "_cg32", None) or cg is getattr(self, "_cgi32", None) else "float64"
        it_fn = "pyib_grid_cg_iterate" if grid else "pyib_cg_iterate"
        m = 0
        while m < cg_max:
            n = min(check - (m % check), cg_max - m)
            at_check = (m + n) % check == 0
            ev = timer.begin() if timer is not None else None
            N.call(it_fn, mesh, ws, C.byref(cg), float(self.rou), float(self.visc), hdt, int(n), int(at_check),
                   dev_stream())
            if timer is not None:
                timer.end(ev, (int(n), prec))
            m += n
            if at_check:
                eta_new = float(scal[N.S_ETA_NEW].item())
                if not np.isfinite(eta_new) or eta_new < stop_value():
                    break
                if grid:
                    N.call("pyib_grid_cg_direction", mesh, C.byref(cg), dev_stream())
                else:
                    N.call("pyib_cg_direction", mesh, ws, C.byref(cg), dev_stream())
        return m

    def helm_solve(self, res_e, s, dt, tol1=1e-10, tol2=1e-9, cg_max=None):
        """ib_driver_np.py:311-410 -> (x, iterations) with x a new device tensor."""
        cg_max = self._positive_cg_limit(self.cg_max_iterations if cg_max is None else cg_max)
        if self.grid is not None:
            return self._helm_solve_grid(res_e, s, dt, tol1, tol2, cg_max)
        if self.precision == "mixed":
            return self._helm_solve_mixed(res_e, s, dt, tol1, tol2, cg_max)
        core = self.core
        mesh, ws = core.struct(), core.work_struct()
        hdt = 0.5 * float(dt)
        N.call("pyib_cg_begin", mesh, ws, C.byref(self._cg64), ptr(res_e), ptr(s), float(tol1), float(tol2),
               dev_stream())
        sc = self._scal.cpu().numpy()
        eta0, err_tot, l2_init = float(sc[N.S_ETA]), float(sc[N.S_ERR]), float(sc[N.S_L2I])
        if eta0 < err_tot:
            N.call("pyib_dot", N.F64, ptr(self._r), ptr(self.prec), ptr(self._r), C.c_int64(core.nedge),
                   ptr(core.work["float64"]["partials"]), ptr(self._out2), dev_stream())
            o = self._out2.cpu().numpy()
            self._finish(0, float(o[0]), l2_init, l2_init, err_tot, cg_max)
            return self._x.clone(), 0
        m = self._pcg_loop(mesh, ws, self._cg64, self._scal, cg_max, lambda: err_tot, dt)
        N.call("pyib_cg_true_residual", mesh, ws, C.byref(self._cg64), ptr(res_e), ptr(s), ptr(self._tmp),
               ptr(self._tr), float(self.rou), float(self.visc), hdt, ptr(self._out2), dev_stream())
        o = self._out2.cpu().numpy()
        self._finish(m, float(o[0]), float(o[1]), l2_init, err_tot, cg_max)
        return self._x.clone(), int(m)

    def _helm_solve_mixed(self, res_e, s, dt, tol1, tol2, cg_max):
        """Defect correction: float32 PCG on A32 d = r with float64 reliable updates
        (x += d in float64, r <- res_e - A64 (x - s) in float64, search direction kept)."""
        core = self.core
        m64, w64 = core.struct("float64"), core.work_struct("float64")
        m32, w32 = core.struct("float32"), core.work_struct("float32")
        hdt = 0.5 * float(dt)
        n = C.c_int64(core.nedge)
        st = dev_stream
        N.call("pyib_cg_begin", m64, w64, C.byref(self._cg64), ptr(res_e), ptr(s), float(tol1), float(tol2), st())
        sc = self._scal.cpu().numpy()
        eta0, err_tot, l2_init = float(sc[N.S_ETA]), float(sc[N.S_ERR]), float(sc[N.S_L2I])
        if eta0 < err_tot:
            N.call("pyib_dot", N.F64, ptr(self._r), ptr(self.prec), ptr(self._r), n,
                   ptr(core.work["float64"]["partials"]), ptr(self._out2), st())
            o = self._out2.cpu().numpy()
            self._finish(0, float(o[0]), l2_init, l2_init, err_tot, cg_max)
            return self._x.clone(), 0
        # _x holds s, _r the masked float64 residual; float32 PCG with reliable updates (as
        # in _helm_solve_grid)
        total = 0
        outer = 0
        eta_f, l2_f = eta0, l2_init
        zero32 = self._z32
        zero32.zero_()
        N.call("pyib_convert", N.F64, ptr(self._r), N.F32, ptr(self._r32), n, st())
        N.call("pyib_cg_begin", m32, w32, C.byref(self._cg32), ptr(self._r32), ptr(zero32), 0.0,
               self.mixed_inner_tol, st())
        inner_err = float(self._scal32[N.S_ERR].item())
        while True:
            outer += 1
            k = self._pcg_loop(m32, w32, self._cg32, self._scal32, max(cg_max - total, 1),
                               lambda: max(inner_err, err_tot), dt)
            total += k
            N.call("pyib_add_into_f64", N.F32, ptr(self._x32), ptr(self._x), n, st())
            N.call("pyib_cg_true_residual", m64, w64, C.byref(self._cg64), ptr(res_e), ptr(s), ptr(self._tmp),
                   ptr(self._tr), float(self.rou), float(self.visc), hdt, ptr(self._out2), st())
            o = self._out2.cpu().numpy()
            eta_f, l2_f = float(o[0]), float(o[1])
            if not np.isfinite(eta_f) or eta_f <= err_tot or total >= cg_max or outer >= self.mixed_max_outer:
                break
            self._x32.zero_()
            N.call("pyib_convert", N.F64, ptr(self._tr), N.F32, ptr(self._r32), n, st())
            N.call("pyib_cg_replace_residual", ptr(self._scal32), ptr(self._out2), st())
            N.call("pyib_cg_direction", m32, w32, C.byref(self._cg32), st())
            inner_err = (self.mixed_inner_tol ** 2) * eta_f
        self.helmholtz_outer_last_step += outer
        self._finish(total, eta_f, l2_f, l2_init, err_tot, cg_max)
        return self._x.clone(), int(total)

    def _helm_solve_grid(self, res_e, s, dt, tol1, tol2, cg_max):
        """helm_solve on the structured layout; float64 PCG, or float32 PCG inside float64
        defect correction when precision == 'mixed'."""
        g = self.grid
        m64, w64 = g.struct("float64"), g.work_struct("float64")
        hdt = 0.5 * float(dt)
        st = dev_stream
        n = C.c_int64(g.nedge)
        g.to_internal(res_e, out=self._resi)
        g.to_internal(s, out=self._si)
        N.call("pyib_grid_cg_begin", m64, w64, C.byref(self._cgi), ptr(self._resi), ptr(self._si), float(tol1),
               float(tol2), st())
        sc = self._scal.cpu().numpy()
        eta0, err_tot, l2_init = float(sc[N.S_ETA]), float(sc[N.S_ERR]), float(sc[N.S_L2I])
        if eta0 < err_tot:
            self._finish(0, eta0, l2_init, l2_init, err_tot, cg_max)
            return g.to_reference(self._xi), 0
        if self.precision != "mixed":
            # At the reference's tolerances this is its PCG unchanged.  A solve asked for tighter
            # tolerances (below the recursion residual's drift from the true one) replaces the
            # residual by the float64 true residual whenever the recursion has passed the stop
            # test but the true residual has not, keeps the search direction and goes on.
            # A replacement that does not halve the true residual's eta means the target is below
            # what float64 can resolve; the solve then ends as the reference's would.
            replace = tol1 < self.REFERENCE_TOL1 or tol2 < self.REFERENCE_TOL2
            total, repl, prev = 0, 0, float("inf")
            while True:
                total += self._pcg_loop(m64, w64, self._cgi, self._scal, cg_max - total, lambda: err_tot, dt,
                                        grid=True)
                N.call("pyib_grid_cg_true_residual", m64, w64, C.byref(self._cgi), ptr(self._resi), ptr(self._si),
                       ptr(self._tmpi), ptr(self._tri), float(self.rou), float(self.visc), hdt, ptr(self._out2),
                       st())
                o = self._out2.cpu().numpy()
                eta_f, l2_f = float(o[0]), float(o[1])
                if not (replace and np.isfinite(eta_f) and eta_f > err_tot and total < cg_max
                        and repl < self.max_residual_replacements and eta_f < 0.5 * prev):
                    break
                prev = eta_f
                repl += 1
                self._ri.copy_(self._tri)
                N.call("pyib_cg_replace_residual", ptr(self._scal), ptr(self._out2), st())
                N.call("pyib_grid_cg_direction", m64, C.byref(self._cgi), st())
            self.residual_replacements_last_solve = repl
            self._finish(total, eta_f, l2_f, l2_init, err_tot, cg_max)
            return g.to_reference(self._xi), int(total)
        # float32 PCG with reliable updates: every time the float32 recurrence residual has
        # dropped by mixed_inner_tol (or reached the float64 target) the float32 correction is
        # added into the float64 solution, the float64 true residual is formed and it replaces
        # the float32 residual; the search direction is kept, so the Krylov space is not lost.
        m32, w32 = g.struct("float32"), g.work_struct("float32")
        total, outer = 0, 0
        eta_f, l2_f = eta0, l2_init
        self._zi32.zero_()
        N.call("pyib_convert", N.F64, ptr(self._ri), N.F32, ptr(self._ri32), n, st())
        N.call("pyib_grid_cg_begin", m32, w32, C.byref(self._cgi32), ptr(self._ri32), ptr(self._zi32), 0.0,
               self.mixed_inner_tol, st())
        inner_err = float(self._scal32[N.S_ERR].item())
        while True:
            outer += 1
            k = self._pcg_loop(m32, w32, self._cgi32, self._scal32, max(cg_max - total, 1),
                               lambda: max(inner_err, err_tot), dt, grid=True)
            total += k
            N.call("pyib_add_into_f64", N.F32, ptr(self._xi32), ptr(self._xi), n, st())
            N.call("pyib_grid_cg_true_residual", m64, w64, C.byref(self._cgi), ptr(self._resi), ptr(self._si),
                   ptr(self._tmpi), ptr(self._tri), float(self.rou), float(self.visc), hdt, ptr(self._out2), st())
            o = self._out2.cpu().numpy()
            eta_f, l2_f = float(o[0]), float(o[1])
            if not np.isfinite(eta_f) or eta_f <= err_tot or total >= cg_max or outer >= self.mixed_max_outer:
                break
            self._xi32.zero_()
            N.call("pyib_convert", N.F64, ptr(self._tri), N.F32, ptr(self._ri32), n, st())
            N.call("pyib_cg_replace_residual", ptr(self._scal32), ptr(self._out2), st())
            N.call("pyib_grid_cg_direction", m32, C.byref(self._cgi32), st())
            inner_err = (self.mixed_inner_tol ** 2) * eta_f
        self.helmholtz_outer_last_step += outer
        self._finish(total, eta_f, l2_f, l2_init, err_tot, cg_max)
        return g.to_reference(self._xi), int(total)

    # -------------------------------------------------------------- step
    def _spread_held_force(self):
        ib = self.ib
        ib.stencil(self.ib_markers, self._marker_version)
        self.ib_force_E.zero_()
        ib.spread_into(self.ib_force_E, self.ib_force_L, self._ds3, 1.0, accumulate=False)

    def _lagrangian_force(self, vc, dt):
        """u_d - interp(vc) over dt, through the coupling inverse when the forcing is implicit
        (ib_immersed.py:113-120 / 152-174)."""
        ib = self.ib
        ib.stencil(self.ib_markers, self._marker_version)
        ib.interp(vc)
        if self.ib_implicit:
            return ib.implicit_lagrangian_force(self.ib_vel, dt, self._ds3)
        return ib.lagrangian_force(self.ib_vel, dt)

    def _direct_force_increment(self, vc, dt):
        """Accumulated explicit increment (ib_driver_np.py:511-527)."""
        ib = self.ib
        k = self.rou * self.ib_alpha
        f_lag = self._lagrangian_force(vc, dt)
        ib.spread_into(self.ib_force_E, f_lag, self._ds3, k, accumulate=True)
        N.call("pyib_axpby", N.F64, 1.0, ptr(self.ib_force_L), float(k), ptr(f_lag), ptr(self.ib_force_L),
               C.c_int64(self.ib_force_L.numel()), dev_stream())
        self.ib_force_update_count += 1
        self.ib_force_updates_last_step += 1

    def _direct_force_replace(self, vc, dt):
        """Replacement forcing at every inner iteration (ib_driver_np.py:472-489)."""
        ib = self.ib
        k = self.rou * self.ib_alpha
        f_lag = self._lagrangian_force(vc, dt)
        self.ib_force_E.zero_()
        ib.spread_into(self.ib_force_E, f_lag, self._ds3, k, accumulate=False)
        N.call("pyib_axpby", N.F64, float(k), ptr(f_lag), 0.0, ptr(None), ptr(self.ib_force_L),
               C.c_int64(self.ib_force_L.numel()), dev_stream())
        self.ib_force_update_count += 1
        self.ib_force_updates_last_step += 1

    def step(self, s, s_old, dt, dt_old, inner=3):
        """ib_driver_np.py:412-532 -> (s_new, vc_new, s_old_new) device tensors."""
        core, rou, ac, ad = self.core, self.rou, self.ac, self.ad
        accumulated = self.ib is not None and self.ib_force_mode in {"accumulated_explicit", "fortran_accumulated"}
        if accumulated and inner < 2:
            raise ValueError("accumulated_explicit requires at least 2 fluid inner iterations")
        self._step_count += 1
        self.time += dt
        self.helmholtz_iterations_last_step = 0
        self.helmholtz_iterations_sum_last_step = 0
        self.helmholtz_outer_last_step = 0
        self.helmholtz_relative_residual_last_step = 0.0
        self.helmholtz_tolerance_ratio_last_step = 0.0
        self.helmholtz_converged_last_step = True
        self.last_helm_rel_residual = 0.0
        self.last_helm_tolerance_ratio = 0.0
        self.last_helm_converged = True
        self.last_face_flux = None
        self.last_momentum_residual = None
        self.last_step_dt = None
        if self.ib is not None:
            self.ib_force_updates_last_step = 0
        if self.ib is not None and self.motion is not None:
            mk, vel = self.motion(self.time)
            self.set_markers(mk, vel)
            if accumulated:
                self._spread_held_force()
        cg = self._positive_cg_limit(self.cg_max_iterations)
        _, U, vc = self.calc_vel(s)
        res_c0 = self.momentum_residual(U, vc, None, rou, (1 - ac) * dt * rou, (1 - ad) * dt)
        qe = s.clone()
        guess = torch.empty_like(s)
        N.call("pyib_extrapolate", C.c_int64(core.nedge), ptr(s), ptr(s_old), float(dt), float(dt_old), ptr(guess),
               dev_stream())
        s, U, vc = self.calc_vel(guess)
        s_old = qe
        for i in range(1, inner + 1):
            force = self.ib_force_E if (self.ib is not None and accumulated) else None
            if self.ib is not None and not accumulated:
                self._direct_force_replace(vc, dt)
                force = self.ib_force_E
            res_c = self.momentum_residual(U, vc, res_c0, rou, ac * dt * rou, ad * dt, force, dt)
            if i == inner:
                self.last_momentum_residual = res_c
            res_e = self.curl_rhs(res_c)
            s, helm_m = self.helm_solve(res_e, s, dt, cg_max=cg)
            self.helmholtz_iterations_last_step = max(self.helmholtz_iterations_last_step, int(helm_m))
            self.helmholtz_iterations_sum_last_step += int(helm_m)
            self.helmholtz_relative_residual_last_step = max(self.helmholtz_relative_residual_last_step,
                                                             self.last_helm_rel_residual)
            self.helmholtz_tolerance_ratio_last_step = max(self.helmholtz_tolerance_ratio_last_step,
                                                           self.last_helm_tolerance_ratio)
            self.helmholtz_converged_last_step = self.helmholtz_converged_last_step and self.last_helm_converged
            s, U, vc = self.calc_vel(s, s_out=s)
            if accumulated and i == 1:
                self._direct_force_increment(vc, dt)
        self.last_face_flux = U
        self.last_step_dt = float(dt)
        return s, vc, s_old
