"""Oracle restatement of the Crank-Nicolson stream-function stepper.

TEST INFRASTRUCTURE ONLY.  Follows ``src/pyib/solver/ib_driver_np.py`` in
float64 on the CPU: ``diag_preconditioner`` 78-88, ``Driver`` 94-532
(``set_cavity`` 155-177, ``set_ib`` 179-201, ``calc_vel`` 267-291,
``conv_diff`` 293-303, ``helm_solve`` 311-410, ``step`` 412-532).  The CG
keeps the reference's accepted-iterate contract: warm start, mixed stopping
criterion tested only every ``cg_check_interval`` iterations, guarded
``alpha``/``beta`` and the true-residual epilogue.
"""
from __future__ import annotations

import numpy as np


def diag_preconditioner(core, dir_mask):
    """ib_driver_np.py:78-88."""
    qf = (core.rf * core.nf).sum(axis=1) / (core.Af * core.Af)
    prec = np.bincount(core.lf_edge, np.abs(core.fe_sign_link) * qf[core.lf_face], minlength=core.nedge)
    prec[dir_mask] = 0.0
    inv = np.zeros(core.nedge)
    nz = prec != 0.0
    inv[nz] = 1.0 / prec[nz]
    return inv


class Driver:
    """ib_driver_np.py:94-532, float64, NumPy."""

    def __init__(self, core, mesh, visc, rou=1.0, ac=0.5, ad=0.5, *, cg_max_iterations=1000,
                 fail_on_cg_nonconvergence=False, cg_check_interval=20):
        self.core = core
        self.visc, self.rou, self.ac, self.ad = visc, rou, ac, ad
        self.bc_edge = np.asarray(mesh.bc_edge)
        self.bc_face = np.asarray(mesh.bc_face)
        bf = np.where(~core.m1)[0]
        self.bf = bf
        self.nghost = int(bf.shape[0])
        self.face_ghost = np.full(core.nface, -1, np.int64)
        self.face_ghost[bf] = core.ncell + np.arange(self.nghost)
        self.F2C_ext = core.F2C.copy()
        self.F2C_ext[bf, 1] = self.face_ghost[bf]
        self.bf_fam = self.bc_face[bf]
        self.inlet_edges = self.bc_edge == 1
        self.dir_mask = np.isin(self.bc_edge, [3, 4, 5, 6])
        self.prec = diag_preconditioner(core, self.dir_mask)
        self.Asp = None
        self.ib = None
        self.ib_alpha = 1.0
        self.cavity = False
        self.motion = None
        self.time = 0.0
        self._step_count = 0
        self.cg_max_iterations = int(cg_max_iterations)
        self.fail_on_cg_nonconvergence = bool(fail_on_cg_nonconvergence)
        self.cg_check_interval = int(cg_check_interval)
        self.last_helm_m = 0
        self.helmholtz_iterations_last_step = 0
        self.helmholtz_iterations_sum_last_step = 0
        self.last_helm_rel_residual = 0.0
        self.last_helm_tolerance_ratio = 0.0
        self.last_helm_converged = True
        self.helmholtz_relative_residual_last_step = 0.0
        self.helmholtz_tolerance_ratio_last_step = 0.0
        self.helmholtz_converged_last_step = True
        self.last_face_flux = None
        self.last_momentum_residual = None
        self.last_step_dt = None

    def set_cavity(self, U_lid=1.0, spanwise_zerograd=True):
        """ib_driver_np.py:155-177."""
        self.cavity = True
        self.U_lid = float(U_lid)
        self._lid_vec = np.asarray([U_lid, 0.0, 0.0])
        self.inlet_edges = np.zeros(self.core.nedge, dtype=bool)
        self.dir_mask = self.bc_edge > 0
        self.prec = diag_preconditioner(self.core, self.dir_mask)
        fp = self.core.fpos
        fy, fz = fp[self.bf, 1], fp[self.bf, 2]
        tol = 1e-6 + 1e-3 * (fy.max() - fy.min())
        self._lid_m = fy > fy.max() - tol
        zf = (fz < fz.min() + tol) | (fz > fz.max() - tol)
        if not spanwise_zerograd:
            zf = np.zeros_like(zf, dtype=bool)
        self._zf_m = zf
        return self

    def set_ib(self, ib, markers, ds, vel_desired, force_mode="replacement"):
        """ib_driver_np.py:179-201."""
        if force_mode not in {"replacement", "accumulated_explicit"}:
            raise ValueError(f"unsupported IB force_mode: {force_mode!r}")
        self.ib = ib
        self.ib_markers = np.asarray(markers, float)
        self.ib_ds = np.asarray(ds, float)
        self.ib_vel = np.asarray(vel_desired, float)
        self.ib_force_mode = force_mode
        self.ib_force_E = np.zeros((self.core.ncell, 3))
        self.ib_force_L = np.zeros_like(self.ib_vel)
        self.ib_force_increment_E = np.zeros_like(self.ib_force_E)
        self.ib_force_increment_L = np.zeros_like(self.ib_force_L)
        self.ib_force_update_count = 0
        self.ib_force_updates_last_step = 0
        return self

    def set_motion(self, motion, t0=0.0):
        """ib_driver_np.py:241-249."""
        self.motion = motion
        self.time = float(t0)
        return self

    def enable_sparse(self, dt):
        """ib_driver_np.py:251-257."""
        from .sparse import build_A
        self.Asp = build_A(self.core, alpd=self.ad, visc=self.visc, rou=self.rou, dt=dt)
        return self

    def cgrad_ext(self, col):
        return col[self.F2C_ext[:, 1]] - col[self.F2C_ext[:, 0]]

    def wavg_ext(self, col, wf1, wf2):
        return col[self.F2C_ext[:, 0]] * wf1 + col[self.F2C_ext[:, 1]] * wf2

    def calc_vel(self, s):
        """ib_driver_np.py:267-291: inlet stream BC, U = curl(s), vc = interp(U), ghosts."""
        core = self.core
        s = s.copy()
        s[self.inlet_edges] = core.epos[self.inlet_edges, 1] * core.te[self.inlet_edges, 2]
        U = core.curl(s)
        vc = np.zeros((core.ncell + self.nghost, 3))
        vc[:core.ncell] = core.interp_f2c(U)
        bf = self.bf
        c1 = core.F2C[bf, 0]
        g = self.face_ghost[bf]
        if self.cavity:
            v1 = vc[c1]
            gv = -v1
            gv[self._lid_m] = 2.0 * self._lid_vec - v1[self._lid_m]
            gv[self._zf_m] = v1[self._zf_m]
            vc[g] = gv
            return s, U, vc
        nf = core.nf[bf]
        Af2 = (core.Af[bf] ** 2)[:, None]
        Uf = U[bf][:, None]
        zero = self.bf_fam == 1
        vc[g[zero]] = Uf[zero] * nf[zero] / Af2[zero]
        zg = ~zero
        v1 = vc[c1[zg]]
        vc[g[zg]] = v1 + (Uf[zg] - (v1 * nf[zg]).sum(axis=1, keepdims=True)) * nf[zg] / Af2[zg]
        return s, U, vc

    def conv_diff(self, U, vc):
        """ib_driver_np.py:293-303: central convection + diffusion per component."""
        core = self.core
        conv = np.zeros((core.ncell, 3))
        diff = np.zeros((core.ncell, 3))
        wf = 0.5 * U
        for j in range(3):
            rd = self.cgrad_ext(vc[:, j])
            diff[:, j] = core.vol_ci * core.cdiv(core.AoL * self.visc * rd)
            conv[:, j] = core.vol_ci * core.cdiv(self.wavg_ext(vc[:, j], wf, wf))
        return conv, diff

    def apply_A(self, p, dt):
        if self.Asp is not None:
            return self.Asp @ p
        return self.core.flow_mat(p, alpd=self.ad, visc=self.visc, rou=self.rou, dt=dt)

    def helm_solve(self, res_e, s, dt, tol1=1e-10, tol2=1e-9, cg_max=None):
        """ib_driver_np.py:311-410 (float64 path)."""
        cg_max = self.cg_max_iterations if cg_max is None else int(cg_max)
        r = res_e.copy()
        r[self.dir_mask] = 0.0
        l2_init = r.dot(r)
        x = s.copy()
        pp = r * self.prec
        eta = pp.dot(r)
        rho0 = eta
        xnorm = x.dot(x) + 1e-16
        err_tot = xnorm * (tol1 ** 2) + (tol2 ** 2) * rho0

        def finish(iters, true_res):
            eta_f = (true_res * self.prec).dot(true_res)
            l2_f = true_res.dot(true_res)
            vals = np.array([eta_f, l2_f, l2_init, err_tot], float)
            finite = bool(np.all(np.isfinite(vals)) and np.all(vals >= 0.0))
            self.last_helm_m = int(iters)
            self.last_helm_rel_residual = (float(np.sqrt(l2_f / l2_init)) if finite and l2_init > 0.0
                                           else (0.0 if finite and l2_f == 0.0 else float("inf")))
            self.last_helm_tolerance_ratio = (float(np.sqrt(eta_f / err_tot)) if finite and err_tot > 0.0
                                              else (0.0 if finite and eta_f == 0.0 else float("inf")))
            self.last_helm_converged = bool(finite and eta_f <= err_tot)
            if self.fail_on_cg_nonconvergence and not self.last_helm_converged:
                raise RuntimeError(
                    "Helmholtz CG did not converge: "
                    f"iterations={iters}/{cg_max}, "
                    f"true_relative_residual={self.last_helm_rel_residual:.6e}, "
                    f"mixed_tolerance_ratio={self.last_helm_tolerance_ratio:.6e}")
            return x, int(iters)

        if float(eta) < err_tot:
            return finish(0, r)
        check = self.cg_check_interval
        m = 0
        for m in range(1, cg_max + 1):
            w = self.apply_A(pp, dt)
            w[self.dir_mask] = 0.0
            den = pp.dot(w)
            alf = eta / (den if den != 0.0 else np.inf)
            x += alf * pp
            r -= alf * w
            z = r * self.prec
            eta_new = z.dot(r)
            if m % check == 0:
                ce = float(eta_new)
                if not np.isfinite(ce) or ce < err_tot:
                    break
            beta = eta_new / (eta if eta != 0.0 else np.inf)
            pp = z + beta * pp
            eta = eta_new
        corr_img = self.apply_A(x - s, dt)
        true_res = res_e - corr_img
        true_res[self.dir_mask] = 0.0
        return finish(m, true_res)

    def step(self, s, s_old, dt, dt_old, inner=3):
        """ib_driver_np.py:412-532 (explicit forcing, or implicit forcing when ``ib_implicit``)."""
        core, rou, ac, ad = self.core, self.rou, self.ac, self.ad
        accumulated = self.ib is not None and self.ib_force_mode == "accumulated_explicit"
        if accumulated and inner < 2:
            raise ValueError("accumulated_explicit requires at least 2 fluid inner iterations")
        self._step_count += 1
        self.time += dt
        self.helmholtz_iterations_last_step = 0
        self.helmholtz_iterations_sum_last_step = 0
        self.helmholtz_relative_residual_last_step = 0.0
        self.helmholtz_tolerance_ratio_last_step = 0.0
        self.helmholtz_converged_last_step = True
        self.last_face_flux = None
        self.last_momentum_residual = None
        self.last_step_dt = None
        if self.ib is not None:
            self.ib_force_updates_last_step = 0
        if self.ib is not None and self.motion is not None:
            mk, vel = self.motion(self.time)
            self.ib_markers = np.asarray(mk, float)
            self.ib_vel = np.asarray(vel, float)
            if getattr(self, "ib_implicit", False):
                self.ib.invalidate_implicit()            # ib_driver_np.py:445-446
            if accumulated:
                self.ib_force_E = self.ib.spread(self.ib_force_L, self.ib_markers, self.ib_ds)
        cg = self.cg_max_iterations
        _, U, vc = self.calc_vel(s)
        conv0, diff0 = self.conv_diff(U, vc)
        res_c0 = rou * vc[:core.ncell] - (1 - ac) * dt * rou * conv0 + (1 - ad) * dt * diff0
        qe = s.copy()
        s = s + (s - s_old) * dt / dt_old
        s, U, vc = self.calc_vel(s)
        s_old = qe
        for i in range(1, inner + 1):
            conv, diff = self.conv_diff(U, vc)
            res_c = res_c0 - rou * vc[:core.ncell] - ac * dt * rou * conv + ad * dt * diff
            if self.ib is not None:
                if accumulated:
                    res_c = res_c + dt * self.ib_force_E
                else:
                    force = self.ib.implicit_force if getattr(self, "ib_implicit", False) else self.ib.direct_force
                    fE, fL = force(vc[:core.ncell], self.ib_markers, self.ib_vel, dt, self.ib_ds)
                    self.ib_force_E = (self.rou * self.ib_alpha) * fE
                    self.ib_force_L = (self.rou * self.ib_alpha) * fL
                    self.ib_force_increment_E = self.ib_force_E
                    self.ib_force_increment_L = self.ib_force_L
                    self.ib_force_update_count += 1
                    self.ib_force_updates_last_step += 1
                    res_c = res_c + dt * self.ib_force_E
            if i == inner:
                self.last_momentum_residual = res_c
            res_e = core.rot(core.integr_c2f(res_c))
            s, helm_m = self.helm_solve(res_e, s, dt, cg_max=cg)
            self.helmholtz_iterations_last_step = max(self.helmholtz_iterations_last_step, int(helm_m))
            self.helmholtz_iterations_sum_last_step += int(helm_m)
            self.helmholtz_relative_residual_last_step = max(self.helmholtz_relative_residual_last_step,
                                                             self.last_helm_rel_residual)
            self.helmholtz_tolerance_ratio_last_step = max(self.helmholtz_tolerance_ratio_last_step,
                                                           self.last_helm_tolerance_ratio)
            self.helmholtz_converged_last_step = self.helmholtz_converged_last_step and self.last_helm_converged
            s, U, vc = self.calc_vel(s)
            if accumulated and i == 1:
                force = self.ib.implicit_force if getattr(self, "ib_implicit", False) else self.ib.direct_force
                dE, dL = force(vc[:core.ncell], self.ib_markers, self.ib_vel, dt, self.ib_ds)
                self.ib_force_increment_E = (self.rou * self.ib_alpha) * dE
                self.ib_force_increment_L = (self.rou * self.ib_alpha) * dL
                self.ib_force_E = self.ib_force_E + self.ib_force_increment_E
                self.ib_force_L = self.ib_force_L + self.ib_force_increment_L
                self.ib_force_update_count += 1
                self.ib_force_updates_last_step += 1
        self.last_face_flux = U
        self.last_step_dt = float(dt)
        return s, vc, s_old
