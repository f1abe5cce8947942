"""Adjoint kernels behind torch.autograd.Function (needs a GPU).

Each operator's backward is checked by the dot-product identity
<L dx, g> = <dx, L^T g> against its own forward (linear maps exactly, the
bilinear momentum residual through an exact central difference, the marker
derivatives by central differences), and the K-step gradient of the
period-averaged force with respect to the amplitude and to the marker
positions against central finite differences of the forward run.
"""
import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

import pyib_hip  # noqa: E402,F401
import pyib_hip.autodiff as AD  # noqa: E402
import pyib_hip.case as HC  # noqa: E402
from tests.helpers import progress, small_case  # noqa: E402


def _dot(a, b):
    return float((a.double() * b.double()).sum().item())


def _rel(a, b):
    return abs(a - b) / max(abs(a), abs(b), 1e-300)


@pytest.fixture(scope="module")
def ctx():
    cfg = small_case("oscillating", steps=4)
    rt = HC.build_runtime(cfg)
    HC.initial_state(rt)
    print("[gpu] adjoint context: D8 nested box, 2 levels", flush=True)
    return AD.StepContext(rt, tol1=1e-14, tol2=1e-13, cg_max=20000)


def _vjp(outs, ins, gouts):
    outs = [o for o in outs]
    return torch.autograd.grad(outs, ins, gouts, allow_unused=True)


def test_calc_vel_and_curl_rhs_adjoints(ctx):
    K = ctx
    core = K.core
    g = torch.Generator(device="cuda").manual_seed(0)
    ds = torch.randn(core.nedge, dtype=torch.float64, device="cuda", generator=g)
    gs = torch.randn(core.nedge, dtype=torch.float64, device="cuda", generator=g)
    gU = torch.randn(core.nface, dtype=torch.float64, device="cuda", generator=g)
    gvc = torch.randn((core.ncell + core.nghost, 3), dtype=torch.float64, device="cuda", generator=g)
    x = ds.clone().requires_grad_(True)
    with torch.enable_grad():
        so, U, vc = AD.CalcVel.apply(K, x)
        (gx,) = _vjp([so, U, vc], [x], [gs, gU, gvc])
    lhs = _dot(so, gs) + _dot(U, gU) + _dot(vc, gvc)      # calc_vel is linear here (cavity, lid at rest)
    rhs = _dot(ds, gx)
    print(f"[gpu] calc_vel adjoint: {lhs:.15e} vs {rhs:.15e}")
    assert _rel(lhs, rhs) < 1e-12
    rc = torch.randn((core.ncell, 3), dtype=torch.float64, device="cuda", generator=g)
    ge = torch.randn(core.nedge, dtype=torch.float64, device="cuda", generator=g)
    x = rc.clone().requires_grad_(True)
    with torch.enable_grad():
        e = AD.CurlRhs.apply(K, x)
        (gx,) = _vjp([e], [x], [ge])
    assert _rel(_dot(e, ge), _dot(rc, gx)) < 1e-12


def test_momentum_residual_adjoint(ctx):
    K = ctx
    core = K.core
    g = torch.Generator(device="cuda").manual_seed(1)
    rn = lambda *s: torch.randn(*s, dtype=torch.float64, device="cuda", generator=g)
    U, vc, base, F = rn(core.nface), rn(core.ncell + core.nghost, 3), rn(core.ncell, 3), rn(core.ncell, 3)
    dU, dvc, db, dF = rn(core.nface), rn(core.ncell + core.nghost, 3), rn(core.ncell, 3), rn(core.ncell, 3)
    go = rn(core.ncell, 3)
    coef = (1.0, 0.37, 0.011, 0.01)
    f = lambda a, b, c, d: AD.MomentumResidual.apply(K, a, b, c, d, *coef)
    ins = [t.clone().requires_grad_(True) for t in (U, vc, base, F)]
    with torch.enable_grad():
        out = f(*ins)
        grads = _vjp([out], ins, [go])
    # bilinear in (U, vc), linear in (base, F): the central difference with unit step is exact
    jvp = (f(U + dU, vc + dvc, base + db, F + dF) - f(U - dU, vc - dvc, base - db, F - dF)) / 2.0
    lhs = _dot(jvp, go)
    rhs = sum(_dot(d, gr) for d, gr in zip((dU, dvc, db, dF), grads))
    print(f"[gpu] momentum residual adjoint: {lhs:.15e} vs {rhs:.15e}")
    assert _rel(lhs, rhs) < 1e-11


def test_ib_transfer_adjoints(ctx):
    K = ctx
    core = K.core
    X0 = K.rt.driver.motion.ref
    X = torch.from_numpy(np.ascontiguousarray(X0)).cuda() + 0.3 * K.dx
    m = X.shape[0]
    g = torch.Generator(device="cuda").manual_seed(2)
    rn = lambda *s: torch.randn(*s, dtype=torch.float64, device="cuda", generator=g)
    f, vc, gE, gu, dX = rn(m, 3), rn(core.ncell + core.nghost, 3), rn(core.ncell, 3), rn(m, 3), rn(m, 3)
    st = K.stencil(X)
    fi, Xi = f.clone().requires_grad_(True), X.clone().requires_grad_(True)
    with torch.enable_grad():
        E = AD.Spread.apply(K, st, fi, Xi, 0.7)
        gf, gX = _vjp([E], [fi, Xi], [gE])
    assert _rel(_dot(E, gE), _dot(f, gf)) < 1e-12            # linear in f
    h = 1e-6 * K.dx
    sp, sm = K.stencil(X + h * dX), K.stencil(X - h * dX)
    fd = (_dot(AD.Spread.apply(K, sp, f, X + h * dX, 0.7), gE) - _dot(AD.Spread.apply(K, sm, f, X - h * dX, 0.7), gE)) \
        / (2 * h)
    print(f"[gpu] spread marker derivative: adjoint {_dot(dX, gX):.12e} fd {fd:.12e}")
    assert _rel(fd, _dot(dX, gX)) < 1e-6
    vi, Xi = vc.clone().requires_grad_(True), X.clone().requires_grad_(True)
    with torch.enable_grad():
        u = AD.Interp.apply(K, st, vi, Xi)
        gv, gX = _vjp([u], [vi, Xi], [gu])
    assert _rel(_dot(u, gu), _dot(vc, gv)) < 1e-12            # interpolation = spreading^T
    fd = (_dot(AD.Interp.apply(K, sp, vc, X + h * dX), gu) - _dot(AD.Interp.apply(K, sm, vc, X - h * dX), gu)) / (2 * h)
    print(f"[gpu] interp marker derivative: adjoint {_dot(dX, gX):.12e} fd {fd:.12e}")
    assert _rel(fd, _dot(dX, gX)) < 1e-6


def test_load_adjoints(ctx):
    K = ctx
    core = K.core
    g = torch.Generator(device="cuda").manual_seed(3)
    rn = lambda *s: torch.randn(*s, dtype=torch.float64, device="cuda", generator=g)
    E, g3 = rn(core.ncell, 3), rn(3)
    Ei = E.clone().requires_grad_(True)
    with torch.enable_grad():
        ff = AD.FluidForce.apply(K, Ei)
        (gE,) = _vjp([ff], [Ei], [g3])
    assert _rel(_dot(ff, g3), _dot(E, gE)) < 1e-12
    vc = rn(core.ncell + core.nghost, 3)
    c = K.center0 + torch.tensor([0.3 * K.dx, -0.2 * K.dx, 0.1 * K.dx], dtype=torch.float64, device="cuda")
    vi, ci = vc.clone().requires_grad_(True), c.clone().requires_grad_(True)
    with torch.enable_grad():
        P = AD.InteriorMomentum.apply(K, vi, ci)
        gv, gc = _vjp([P], [vi, ci], [g3])
    assert _rel(_dot(P, g3), _dot(vc, gv)) < 1e-12
    dc = torch.tensor([0.6, -0.3, 0.2], dtype=torch.float64, device="cuda")
    h = 1e-7 * K.dx
    fd = (_dot(AD.InteriorMomentum.apply(K, vc, c + h * dc), g3) - _dot(AD.InteriorMomentum.apply(K, vc, c - h * dc),
                                                                          g3)) / (2 * h)
    print(f"[gpu] interior momentum centre derivative: adjoint {_dot(dc, gc):.12e} fd {fd:.12e}")
    assert _rel(fd, _dot(dc, gc)) < 1e-5


def test_helmholtz_is_self_adjoint(ctx):
    K = ctx
    core = K.core
    g = torch.Generator(device="cuda").manual_seed(4)
    rn = lambda *s: torch.randn(*s, dtype=torch.float64, device="cuda", generator=g)
    r1 = AD.CurlRhs.apply(K, rn(core.ncell, 3))
    r2 = AD.CurlRhs.apply(K, rn(core.ncell, 3))
    z = torch.zeros_like(r1)
    x1 = AD.Helmholtz.apply(K, r1, z)
    x2 = AD.Helmholtz.apply(K, r2, z)
    a, b = _dot(x1, r2), _dot(r1, x2)
    print(f"[gpu] <A^-1 r1, r2> = {a:.15e}, <r1, A^-1 r2> = {b:.15e}")
    assert _rel(a, b) < 1e-9
    # a cotangent outside the operator's range (any edge vector): <G r1, g> = <r1, G^T g>
    gx = rn(core.nedge)
    ri = r1.clone().requires_grad_(True)
    with torch.enable_grad():
        x = AD.Helmholtz.apply(K, ri, z)
        (gr,) = _vjp([x], [ri], [gx])
    a, b = _dot(x, gx), _dot(r1, gr)
    print(f"[gpu] <G r1, g> = {a:.15e}, <r1, G^T g> = {b:.15e}")
    assert _rel(a, b) < 1e-8


def _objective(steps, every, tol1=1e-14, tol2=1e-13):
    cfg = small_case("oscillating", steps=steps)
    rt = HC.build_runtime(cfg)
    HC.initial_state(rt)
    return AD.PeriodForceObjective(rt, steps, checkpoint_every=every, tol1=tol1, tol2=tol2, cg_max=20000,
                                   log=lambda m: progress(f"[gpu] adjoint: {m}")), rt


def test_forward_objective_matches_runner():
    """J from the differentiable forward == mean body_force_x of the product runner (same tolerance)."""
    steps = 4
    obj, _ = _objective(steps, 2, tol1=1e-10, tol2=1e-9)
    J = obj.value()
    cfg = small_case("oscillating", steps=steps)
    rt = HC.build_runtime(cfg)
    s, s_old, vc = HC.initial_state(rt)
    fx = []
    for n in range(1, steps + 1):
        rt.driver.time = (n - 1) * rt.dt
        s, vc, s_old = rt.driver.step(s, s_old, rt.dt, rt.dt, inner=rt.inner_iterations)
        rt.driver.time = n * rt.dt
        fx.append(rt.load_correction.advance(vc, HC.sphere_center(rt.config, rt.driver.time), rt.dt,
                                             rt.driver.ib_force_E).body_force[0])
    ref = float(np.mean(fx))
    print(f"[gpu] J differentiable forward {J:.16e}, runner {ref:.16e}")
    assert abs(J - ref) <= 1e-12 * abs(ref)


def test_gradient_vs_central_differences():
    steps = 6
    obj, rt = _objective(steps, 2)
    J, gA, gX = obj.value_and_grad()
    D = 2.0 * float(rt.config["ib"]["radius"])
    eps = 1e-6 * D
    fdA = AD.central_difference(obj, direction_A=1.0, eps=eps)
    print(f"[gpu] J={J:.12e} dJ/dA adjoint {gA:.12e} fd {fdA:.12e} rel {_rel(gA, fdA):.2e} "
          f"(CG forward {obj.K.cg_iterations}, adjoint {obj.K.adjoint_cg_iterations})", flush=True)
    assert _rel(gA, fdA) <= 1e-5
    d = AD.direction(gX.shape[0], eps)
    fdX = AD.central_difference(obj, direction_X=d / eps, eps=eps)
    adX = _dot(gX, d) / eps
    print(f"[gpu] dJ/dX0 . d adjoint {adX:.12e} fd {fdX:.12e} rel {_rel(adX, fdX):.2e}")
    assert _rel(adX, fdX) <= 1e-5


def test_checkpoint_recomputation_is_exact():
    """Segments of 1 step and one segment of all steps: the same forward (bitwise) and the
    same gradient up to the order in which autograd sums the contributions across segment
    boundaries (the segment recomputation itself repeats the forward bit for bit)."""
    o1, _ = _objective(3, 1, tol1=1e-10, tol2=1e-9)
    o3, _ = _objective(3, 3, tol1=1e-10, tol2=1e-9)
    J1, a1, x1 = o1.value_and_grad()
    J3, a3, x3 = o3.value_and_grad()
    print(f"[gpu] checkpoint every 1 vs 3 steps: dJ/dA rel {_rel(a1, a3):.2e}, "
          f"dJ/dX0 rel {float((x1 - x3).norm() / x3.norm()):.2e}")
    assert J1 == J3
    assert _rel(a1, a3) < 1e-10
    assert float((x1 - x3).norm() / x3.norm()) < 1e-10
