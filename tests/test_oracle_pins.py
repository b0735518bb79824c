"""Pins of the CPU oracle (oracle/refcpu*.hpp) before it is trusted as the checker.

Each check is chosen so that a plausible mistake in the restatement (dropped term, wrong sign,
wrong index or centering, transposed operand) fails:
- reference headers compiled into oracle/_ref (Interpolator, GridLayout/Yee tables, Faraday,
  Ampere, ParticleArray/CellMap partition): bit-exact agreement call for call;
- B-spline weights vs scipy's BSpline.basis_element (what interpolator_test.py uses);
- the deposit collections of tests/core/numerics/interpolator/test_main.cpp:491-748;
- Boris vs the ODE trajectory of tests/core/numerics/pusher/test_pusher.py (1e-5, the reference
  test's tolerance) and vs closed-form gyration / E x B drift;
- Ohm and the electron closure vs polynomial fields for which the discrete operators are exact;
- the seeded Maxwellian loader vs the reference's own draws (maxwellianVelocity compiled from
  maxwellian_particle_initializer.cpp, ParticleDeltaDistribution, GridLayout::indices) bit for bit;
- Pusher::move with box selectors: the partition it leaves and the prePush-only fate of the
  particles the first selector rejects (closed form).
"""
import os

import numpy as np
import pytest

import oracle_lib as O

ORDERS = [1, 2, 3]
DIMS = [1, 2, 3]
# decided from the file, so that collecting this module (as every GPU run does) maps nothing
NEED_REF = pytest.mark.skipif(not os.path.exists(O.REF_SO),
                              reason="oracle/_ref not built (PHARE sources absent)")


# ------------------------------------------------------------------------------------------
# B-spline weights
# ------------------------------------------------------------------------------------------
def _bspline_weights(order, x, start):
    """B-spline weights of the nodes start..start+order at x, from scipy's basis elements
    (cardinal B-spline of degree `order` centred on each node)."""
    from scipy.interpolate import BSpline

    w = []
    for k in range(order + 1):
        node = start + k
        knots = np.arange(order + 2) - (order + 1) / 2.0 + node
        w.append(float(BSpline.basis_element(knots, extrapolate=False)(x)))
    return np.nan_to_num(np.array(w))


@pytest.mark.parametrize("order", ORDERS)
@pytest.mark.parametrize("dual", [0, 1])
def test_weights_match_scipy_bsplines(order, dual):
    # interpolator/test_main.cpp:149-209: icell 3, deltas 0.0 .. 0.9
    icell = 3
    for ipos in range(10):
        delta = ipos * 0.1
        start, w = O.weights(order, dual, icell, delta)
        x = icell + delta - (0.5 if dual else 0.0)
        expect = _bspline_weights(order, x, start)
        np.testing.assert_allclose(w, expect, rtol=0, atol=4e-16 * 4)


@pytest.mark.parametrize("order", ORDERS)
def test_weights_partition_of_unity(order):
    # interpolator/test_main.cpp:92-96 (1e5 random positions, 1e-10)
    rng = np.random.default_rng(12)
    for dual in (0, 1):
        for delta in rng.random(2000):
            _, w = O.weights(order, dual, 7, delta)
            assert abs(w.sum() - 1.0) < 1e-10


@NEED_REF
@pytest.mark.parametrize("order", ORDERS)
def test_weights_bitwise_vs_reference(order):
    rng = np.random.default_rng(3)
    for dual in (0, 1):
        for delta in list(rng.random(500)) + [0.0, 0.5, np.nextafter(0.5, 0), 1 - 2.0**-53]:
            s0, w0 = O.weights(order, dual, 11, delta)
            s1, w1 = O.weights(order, dual, 11, delta, use_ref=True)
            assert s0 == s1
            assert np.array_equal(w0, w1)


# ------------------------------------------------------------------------------------------
# layout, gather, deposit, Faraday, Ampere, laplacian, projections vs reference headers
# ------------------------------------------------------------------------------------------
def _ncells(dim):
    return [9, 7, 6][:dim]


def _fields(dim, order, rng, qids):
    return [rng.standard_normal(O.layout_shape(dim, order, _ncells(dim), q)) for q in qids]


def _particles(dim, order, n, rng, lower):
    ncell = _ncells(dim)
    pg = [1, 2, 2][order - 1]
    ic = np.stack([rng.integers(lower[d] - pg, lower[d] + ncell[d] + pg, n) for d in range(dim)], 1)
    de = rng.random((n, dim))
    return ic.astype(np.int32), de


@NEED_REF
@pytest.mark.parametrize("dim", DIMS)
@pytest.mark.parametrize("order", ORDERS)
def test_layout_shapes_vs_reference(dim, order):
    R = O.ref()
    n = O._i3(_ncells(dim))
    for q in range(10):
        info = np.zeros(12, dtype=np.int32)
        R.ref_layout_info(dim, order, n, q, info)
        shape = O.layout_shape(dim, order, _ncells(dim), q)
        assert tuple(info[0::4][:dim]) == shape


@NEED_REF
@pytest.mark.parametrize("dim", DIMS)
@pytest.mark.parametrize("order", ORDERS)
def test_gather_bitwise_vs_reference(dim, order):
    rng = np.random.default_rng(100 + 10 * dim + order)
    lower = [5, -3, 2][:dim]
    dx = [0.1, 0.2, 0.3][:dim]
    E = _fields(dim, order, rng, [O.Q_EX, O.Q_EY, O.Q_EZ])
    B = _fields(dim, order, rng, [O.Q_BX, O.Q_BY, O.Q_BZ])
    ic, de = _particles(dim, order, 400, rng, lower)

    # reference
    R = O.ref()
    Er = np.zeros((400, 3))
    Br = np.zeros((400, 3))
    R.ref_gather(dim, order, O._i3(_ncells(dim)), O._d3(dx), O._i3(lower), O._ptrs(E), O._ptrs(B),
                 400, ic.reshape(-1), de.reshape(-1), Er.reshape(-1), Br.reshape(-1))
    # oracle
    sim = O.Sim(dim, order, _ncells(dim), dx, lower)
    for c in range(3):
        sim.set(O.F_EAVG + c, E[c])
        sim.set(O.F_BAVG + c, B[c])
    sim.set_particles(0, 0, dict(icell=ic, delta=de, v=np.zeros((400, 3)), weight=np.ones(400),
                                 charge=np.ones(400)))
    Eo, Bo = sim.gather(2)
    assert np.array_equal(Eo, Er)
    assert np.array_equal(Bo, Br)


@NEED_REF
@pytest.mark.parametrize("dim", DIMS)
@pytest.mark.parametrize("order", ORDERS)
def test_deposit_bitwise_vs_reference(dim, order):
    rng = np.random.default_rng(200 + 10 * dim + order)
    lower = [0, 4, -2][:dim]
    dx = [0.2, 0.25, 0.5][:dim]
    np_ = 300
    ic, de = _particles(dim, order, np_, rng, lower)
    v = rng.standard_normal((np_, 3))
    w = rng.random(np_)
    q = 1.0 + rng.random(np_)
    shp = O.layout_shape(dim, order, _ncells(dim), O.Q_MOM)
    outs = {}
    for side in ("ref", "ora"):
        n_ = np.zeros(shp)
        r_ = np.zeros(shp)
        F = [np.zeros(shp) for _ in range(3)]
        args = (dim, order, O._i3(_ncells(dim)), O._d3(dx), O._i3(lower), np_, ic.reshape(-1),
                de.reshape(-1), v.reshape(-1), w, q, 0.7, n_.reshape(-1), r_.reshape(-1), O._ptrs(F))
        (O.ref().ref_deposit if side == "ref" else O.lib().ora_s_deposit)(*args)
        outs[side] = [n_, r_] + F
    for a, b in zip(outs["ref"], outs["ora"]):
        assert np.array_equal(a, b)


@NEED_REF
@pytest.mark.parametrize("dim", DIMS)
@pytest.mark.parametrize("order", ORDERS)
def test_faraday_ampere_bitwise_vs_reference(dim, order):
    rng = np.random.default_rng(300 + 10 * dim + order)
    n = O._i3(_ncells(dim))
    dx = O._d3([0.1, 0.2, 0.3][:dim])
    lo = O._i3([0] * dim)
    B = _fields(dim, order, rng, [O.Q_BX, O.Q_BY, O.Q_BZ])
    E = _fields(dim, order, rng, [O.Q_EX, O.Q_EY, O.Q_EZ])
    res = {}
    for side in ("ref", "ora"):
        Bn = [np.full_like(b, 7.0) for b in B]
        J = _fields(dim, order, np.random.default_rng(1), [O.Q_JX, O.Q_JY, O.Q_JZ])
        if side == "ref":
            O.ref().ref_faraday(dim, order, n, dx, lo, O._ptrs(B), O._ptrs(E), O._ptrs(Bn), 0.37)
            O.ref().ref_ampere(dim, order, n, dx, lo, O._ptrs(B), O._ptrs(J))
        else:
            O.lib().ora_s_faraday(dim, order, n, dx, lo, O._ptrs(B), O._ptrs(E), O._ptrs(Bn), 0.37)
            O.lib().ora_s_ampere(dim, order, n, dx, lo, O._ptrs(B), O._ptrs(J))
        res[side] = Bn + J
    for a, b in zip(res["ref"], res["ora"]):
        assert np.array_equal(a, b)


# projection ids -> (source quantity, destination quantity)
PROJ = {0: (O.Q_MOM, O.Q_EX), 1: (O.Q_MOM, O.Q_EY), 2: (O.Q_MOM, O.Q_EZ), 3: (O.Q_BX, O.Q_EX),
        4: (O.Q_BY, O.Q_EX), 5: (O.Q_BZ, O.Q_EX), 6: (O.Q_BX, O.Q_EY), 7: (O.Q_BY, O.Q_EY),
        8: (O.Q_BZ, O.Q_EY), 9: (O.Q_BX, O.Q_EZ), 10: (O.Q_BY, O.Q_EZ), 11: (O.Q_BZ, O.Q_EZ),
        12: (O.Q_JX, O.Q_MOM), 13: (O.Q_JY, O.Q_MOM), 14: (O.Q_JZ, O.Q_MOM)}


@NEED_REF
@pytest.mark.parametrize("dim", DIMS)
@pytest.mark.parametrize("order", ORDERS)
def test_projections_and_laplacian_bitwise_vs_reference(dim, order):
    rng = np.random.default_rng(400 + 10 * dim + order)
    n = O._i3(_ncells(dim))
    dx = O._d3([0.1, 0.2, 0.3][:dim])
    lo = O._i3([0] * dim)
    for which, (sq, dq) in PROJ.items():
        src = _fields(dim, order, rng, [sq])[0]
        outs = []
        for f in (O.ref().ref_project, O.lib().ora_s_project):
            out = np.zeros(O.layout_shape(dim, order, _ncells(dim), dq))
            f(dim, order, n, dx, lo, which, sq, dq, src.reshape(-1), out.reshape(-1))
            outs.append(out)
        assert np.array_equal(outs[0], outs[1]), which
    for q in (O.Q_JX, O.Q_JY, O.Q_JZ):
        src = _fields(dim, order, rng, [q])[0]
        outs = []
        for f in (O.ref().ref_laplacian, O.lib().ora_s_laplacian):
            out = np.zeros(src.shape)
            f(dim, order, n, dx, lo, q, src.reshape(-1), out.reshape(-1))
            outs.append(out)
        assert np.array_equal(outs[0], outs[1])


@NEED_REF
@pytest.mark.parametrize("dim", DIMS)
def test_partition_order_vs_reference(dim):
    rng = np.random.default_rng(500 + dim)
    lo = np.array([-3] * dim, dtype=np.int32)
    hi = np.array([8] * dim, dtype=np.int32)
    np_ = 700
    ic0 = rng.integers(-2, 8, (np_, dim)).astype(np.int32)
    move = rng.integers(-1, 2, (np_, dim)) * (rng.random((np_, dim)) < 0.3)
    ic1 = (ic0 + move).astype(np.int32)
    plo = np.array([0] * dim, dtype=np.int32)
    phi = np.array([5] * dim, dtype=np.int32)
    res = []
    for f in (O.ref().ref_partition, O.lib().ora_partition):
        order = np.zeros(np_, dtype=np.int64)
        piv = f(dim, O._i3(lo), O._i3(hi), np_, ic0.reshape(-1), ic1.reshape(-1), O._i3(plo),
                O._i3(phi), order)
        res.append((piv, order))
    assert res[0][0] == res[1][0]
    assert np.array_equal(res[0][1], res[1][1])


# ------------------------------------------------------------------------------------------
# deposit golden collections (interpolator/test_main.cpp:491-748)
# ------------------------------------------------------------------------------------------
COLLECTIONS_1D = {
    1: [(19, 0.5, 1.0, 2.0), (20, 0.5, 0.4, 1.85), (20, 0.5, 0.6, 2.1)],
    2: [(19, 0.0, 1.0, 2.0), (20, 0.0, 0.2, 3.2), (20, 0.0, 0.8, 1.7), (21, 0.0, 1.0, 2.0)],
    3: [(18, 0.5, 1.0, 2.0), (19, 0.5, 1.0, 2.0), (20, 0.5, 1.0, 2.0), (21, 0.5, 0.1, 3.35),
        (21, 0.5, 0.9, 1.85)],
}


def _ulps_equal(a, b, n=4):
    return abs(a - b) <= n * np.spacing(max(abs(a), abs(b)))


@pytest.mark.parametrize("order", ORDERS)
def test_deposit_collection_1d(order):
    parts = COLLECTIONS_1D[order]
    npart = len(parts)
    ic = np.array([[p[0]] for p in parts], dtype=np.int32)
    de = np.array([[p[1]] for p in parts])
    w = np.array([p[2] for p in parts])
    q = np.array([p[3] for p in parts])
    v = np.tile([2.0, -1.0, 1.0], (npart, 1))
    shp = O.layout_shape(1, order, [30], O.Q_MOM)
    n_, r_ = np.zeros(shp), np.zeros(shp)
    F = [np.zeros(shp) for _ in range(3)]
    O.lib().ora_s_deposit(1, order, O._i3([30]), O._d3([0.1]), O._i3([0]), npart, ic.reshape(-1),
                          de.reshape(-1), v.reshape(-1), w, q, 1.0, n_, r_, O._ptrs(F))
    g = [2, 4, 4][order - 1]
    idx = 20 + g
    assert _ulps_equal(n_[idx], 1.0)
    assert _ulps_equal(r_[idx], 2.0)
    assert _ulps_equal(F[0][idx], 2.0)
    assert _ulps_equal(F[1][idx], -1.0)
    assert _ulps_equal(F[2][idx], 1.0)


@pytest.mark.parametrize("order", ORDERS)
def test_deposit_collection_2d(order):
    ij = [(i, j) for i in range(5) for j in range(5)]
    npart = len(ij)
    ic = np.array(ij, dtype=np.int32)
    de = np.full((npart, 2), 0.5)
    w = np.ones(npart)
    q = np.zeros(npart)
    v = np.tile([2.0, -1.0, 1.0], (npart, 1))
    shp = O.layout_shape(2, order, [15, 15], O.Q_MOM)
    n_, r_ = np.zeros(shp), np.zeros(shp)
    F = [np.zeros(shp) for _ in range(3)]
    O.lib().ora_s_deposit(2, order, O._i3([15, 15]), O._d3([0.1, 0.1]), O._i3([0, 0]), npart,
                          ic.reshape(-1), de.reshape(-1), v.reshape(-1), w, q, 1.0, n_.reshape(-1),
                          r_.reshape(-1), O._ptrs(F))
    g = [2, 4, 4][order - 1]
    idx = (2 + g, 2 + g)
    assert _ulps_equal(n_[idx], 1.0)
    assert _ulps_equal(F[0][idx], 2.0)
    assert _ulps_equal(F[1][idx], -1.0)
    assert _ulps_equal(F[2][idx], 1.0)


# ------------------------------------------------------------------------------------------
# Boris
# ------------------------------------------------------------------------------------------
def _trajectory(dim, nsteps, dt, dx, E, B, v0, ic0, de0, q=1.0, m=1.0):
    out = np.zeros(nsteps * dim)
    O.lib().ora_boris_trajectory(dim, nsteps, dt, O._d3(dx), m, q, np.array(E, float),
                                 np.array(B, float), O._i3(ic0), np.array(list(de0) + [0] * (3 - dim), float),
                                 np.array(v0, float), out)
    return out.reshape(nsteps, dim)


@pytest.mark.slow
@pytest.mark.parametrize("dim", DIMS)
def test_boris_trajectory_vs_ode(dim):
    """tests/core/numerics/pusher/test_pusher.py:35-67 + test_pusher.cpp:180-243:
    E=(0.01,-0.05,0.05), B=(1,1,1), q=m=1, dx=0.05, dt=1e-4, 100001 steps, x0=0.25, v0=(0,10,0);
    positions within 1e-5 of the ODE solution."""
    from scipy.integrate import odeint

    E = np.array([0.01, -0.05, 0.05])
    B = np.array([1.0, 1.0, 1.0])

    def deriv(y, t):
        v = y[3:]
        a = np.cross(v, B) + E
        return np.concatenate([v, a])

    dt, tend = 1e-4, 10.0
    nt = int(tend / dt) + 1
    t = np.arange(0, nt * dt, dt)
    sol = odeint(deriv, [0.25, 0.25, 0.25, 0.0, 10.0, 0.0], t, rtol=1e-12, atol=1e-12)
    traj = _trajectory(dim, nt, dt, [0.05] * dim, E, B, [0.0, 10.0, 0.0], [5] * dim, [0.0] * dim)
    # positions are stored as float in the reference test
    traj32 = traj.astype(np.float32).astype(np.float64)
    assert np.max(np.abs(traj32 - sol[:, :dim])) < 1e-5


def test_boris_gyration_closed_form():
    """E=0, B=Bz: |v| conserved to round-off and the rotation angle per step is
    2*atan(q B dt / 2m) (exact property of the Boris rotation)."""
    dt, Bz = 0.05, 2.0
    traj = _trajectory(3, 400, dt, [1.0, 1.0, 1.0], [0, 0, 0], [0, 0, Bz], [0.3, 0.0, 0.1],
                       [100, 100, 100], [0.5, 0.5, 0.5])
    x, y = traj[:, 0], traj[:, 1]
    theta = 2 * np.arctan(Bz * dt / 2)
    # centre of the circle and radius from the exact discrete orbit: v rotates by theta per step
    vx = np.diff(x) / dt
    vy = np.diff(y) / dt
    speed = np.hypot(vx, vy)
    # positions ~100 cells: differencing them costs ~1e-12 of relative precision
    np.testing.assert_allclose(speed, speed[0], rtol=1e-10)
    ang = np.unwrap(np.arctan2(vy, vx))
    np.testing.assert_allclose(np.diff(ang), -theta, rtol=1e-10)


def test_boris_exb_drift():
    """E=(0,Ey,0), B=(0,0,Bz): the guiding centre drifts at E x B / B^2 = (Ey/Bz, 0, 0)."""
    dt, Bz, Ey = 0.01, 1.0, 0.05
    nper = int(round(2 * np.pi / (2 * np.arctan(Bz * dt / 2))))
    n = 20 * nper
    traj = _trajectory(2, n + 1, dt, [1.0, 1.0], [0, Ey, 0], [0, 0, Bz], [0.0, 0.0, 0.0], [50, 50],
                       [0.0, 0.0])
    # average over whole gyro-periods
    xs = traj[:, 0]
    drift = (np.mean(xs[-nper:]) - np.mean(xs[:nper])) / ((n - nper) * dt)
    np.testing.assert_allclose(drift, Ey / Bz, rtol=2e-3)


# ------------------------------------------------------------------------------------------
# Ohm and electrons on polynomial fields (discrete operators exact)
# ------------------------------------------------------------------------------------------
def _coords(sim, q_centering, shape):
    """physical coordinates of every node of a field of the given centering tuple"""
    g = [2, 4, 4][sim.order - 1]
    axes = []
    for d in range(sim.dim):
        idx = np.arange(shape[d])
        off = 0.5 if q_centering[d] == "d" else 0.0
        axes.append((idx - g + off) * sim.dx[d])
    return np.meshgrid(*axes, indexing="ij")


CENT = {"Bx": "pdd", "By": "dpd", "Bz": "ddp", "Ex": "dpp", "Ey": "pdp", "Ez": "ppd", "Jx": "dpp",
        "Jy": "pdp", "Jz": "ppd", "M": "ppp"}


def _poly_setup(dim, order, hyper_mode, rng):
    eta, nu = 0.3, 0.02
    sim = O.Sim(dim, order, [8, 7, 6][:dim], [0.3, 0.25, 0.2][:dim], eta=eta, nu=nu, hyper_mode=hyper_mode)
    a = rng.uniform(-1, 1, (20, 4))

    def lin(k, X):
        r = a[k, 0] + sum(a[k, d + 1] * X[d] for d in range(dim))
        return r

    def quad(k, X):
        return lin(k, X) + sum((0.5 + a[k, d + 1] ** 2) * X[d] ** 2 for d in range(dim))

    fns = {
        "Bx": lambda X: lin(0, X), "By": lambda X: lin(1, X), "Bz": lambda X: lin(2, X),
        "Vx": lambda X: lin(3, X), "Vy": lambda X: lin(4, X), "Vz": lambda X: lin(5, X),
        "n": lambda X: 3.0 + 0.1 * lin(6, X), "Pe": lambda X: lin(7, X),
        "Jx": lambda X: quad(8, X), "Jy": lambda X: quad(9, X), "Jz": lambda X: quad(10, X),
    }
    # fields on the grid
    for c, name in enumerate(["Bx", "By", "Bz"]):
        fid = O.F_B + c
        sim.set(fid, fns[name](_coords(sim, CENT[name], sim.shape(fid))))
    for c, name in enumerate(["Jx", "Jy", "Jz"]):
        fid = O.F_J + c
        sim.set(fid, fns[name](_coords(sim, CENT[name], sim.shape(fid))))
    for c, name in enumerate(["Vx", "Vy", "Vz"]):
        fid = O.F_VE + c
        sim.set(fid, fns[name](_coords(sim, CENT["M"], sim.shape(fid))))
    sim.set(O.F_PE, fns["Pe"](_coords(sim, CENT["M"], sim.shape(O.F_PE))))
    sim.set(O.F_RHO, fns["n"](_coords(sim, CENT["M"], sim.shape(O.F_RHO))))
    return sim, fns, eta, nu


def _grad(fn, X, d, h=1e-6):
    Xp = [x.copy() for x in X]
    Xm = [x.copy() for x in X]
    Xp[d] = Xp[d] + h
    Xm[d] = Xm[d] - h
    return (fn(Xp) - fn(Xm)) / (2 * h)


def _lap_exact(k_quad_coefs_fn, X, dim):
    # second derivative of the quadratic part: 2*(0.5 + a^2) per direction
    return k_quad_coefs_fn(X)


@pytest.mark.parametrize("dim", DIMS)
@pytest.mark.parametrize("order", [1, 2])
@pytest.mark.parametrize("hyper_mode", [0, 1])
def test_ohm_exact_on_polynomials(dim, order, hyper_mode):
    rng = np.random.default_rng(600 + dim * 10 + order + 100 * hyper_mode)
    sim, fns, eta, nu = _poly_setup(dim, order, hyper_mode, rng)
    sim.ohm(0, 0)
    g = [2, 4, 4][order - 1]
    for c, name in enumerate(["Ex", "Ey", "Ez"]):
        E = sim.get(O.F_E + c)
        X = _coords(sim, CENT[name], E.shape)
        V = [fns["Vx"](X), fns["Vy"](X), fns["Vz"](X)]
        Bv = [fns["Bx"](X), fns["By"](X), fns["Bz"](X)]
        n = fns["n"](X)
        vxb = np.cross(np.stack(V, -1), np.stack(Bv, -1))[..., c]
        ideal = -vxb
        press = -_grad(fns["Pe"], X, c) / n if c < dim else 0.0 * n
        Jc = fns[["Jx", "Jy", "Jz"][c]]
        lap = sum((_grad(lambda Y: _grad(Jc, Y, d, 1e-4), X, d, 1e-4)) for d in range(dim))
        if hyper_mode == 0:
            hyper = -nu * lap
        else:
            b = np.sqrt(sum(x * x for x in Bv))
            hyper = -nu * (b / (n + 0.1) + 1) * lap
        expect = ideal + press + eta * Jc(X) + hyper
        sl = tuple(slice(g, E.shape[d] - g) for d in range(dim))
        np.testing.assert_allclose(E[sl], expect[sl], rtol=0, atol=1e-6 * np.abs(expect[sl]).max())


@pytest.mark.parametrize("dim", DIMS)
def test_electrons_exact_on_polynomials(dim):
    rng = np.random.default_rng(700 + dim)
    sim, fns, _, _ = _poly_setup(dim, 1, 0, rng)
    # linear J so that the dual->primal average is exact
    a = rng.uniform(-1, 1, (3, 4))
    for c, name in enumerate(["Jx", "Jy", "Jz"]):
        fid = O.F_J + c
        X = _coords(sim, CENT[name], sim.shape(fid))
        sim.set(fid, a[c, 0] + sum(a[c, d + 1] * X[d] for d in range(dim)))
    vi = [rng.standard_normal(sim.shape(O.F_VI)) for _ in range(3)]
    for c in range(3):
        sim.set(O.F_VI + c, vi[c])
    sim.electrons()
    X = _coords(sim, CENT["M"], sim.shape(O.F_RHO))
    n = fns["n"](X)
    g = 2
    sl = tuple(slice(g, sim.shape(O.F_RHO)[d] - g) for d in range(dim))
    for c in range(3):
        J = a[c, 0] + sum(a[c, d + 1] * X[d] for d in range(dim))
        expect = vi[c] - J / n
        np.testing.assert_allclose(sim.get(O.F_VE + c)[sl], expect[sl], rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(sim.get(O.F_PE), sim.get(O.F_RHO) * 0.1, rtol=0, atol=0)


# ------------------------------------------------------------------------------------------
# Maxwellian loader: the draw stream
# ------------------------------------------------------------------------------------------
@NEED_REF
@pytest.mark.parametrize("dim", DIMS)
def test_loader_bitwise_vs_reference_draws(dim):
    """the oracle's loader against the reference's own pieces: the cell order of
    GridLayout::indices(AMRBox), maxwellianVelocity (three fresh normal distributions per
    particle), then dim ParticleDeltaDistribution<double> draws, one mt19937_64(seed); profiles
    varying per cell and one cell under the density cut-off. Drawing delta before v, sharing one
    normal distribution or visiting cells column-major fails this."""
    ncells = [[23], [6, 5], [4, 3, 5]][dim - 1]
    lower = [[-2], [3, 0], [1, -1, 2]][dim - 1]
    rng = np.random.default_rng(900 + dim)
    nc = int(np.prod(ncells))
    n = 0.5 + rng.random(nc)
    n[nc // 3] = 1e-20
    V = [0.3 * rng.standard_normal(nc) for _ in range(3)]
    T = [0.1 + rng.random(nc) for _ in range(3)]
    ppc, seed = 7, 4242 + dim
    ref = O.ref_load_maxwellian(ncells, lower, n, V, T, 1.5, ppc, seed)
    sim = O.Sim(dim, 1, ncells, [0.3] * dim, lower=lower)
    sim.load_maxwellian(0, n, V, T, 1.5, ppc, seed)
    got = sim.get_particles()
    assert len(ref["weight"]) == (nc - 1) * ppc
    for k in ("icell", "delta", "v", "weight", "charge"):
        assert np.array_equal(got[k], ref[k]), k


# ------------------------------------------------------------------------------------------
# Pusher::move with box selectors
# ------------------------------------------------------------------------------------------
def test_push_selectors_partition_and_prepush_only():
    """a particle the first selector rejects is only prePushed (half step of position, velocity
    unchanged); those it keeps are accelerated; the array is partitioned [both | first only |
    neither] with the returned ends (boris.hpp:93-138)"""
    rng = np.random.default_rng(77)
    nc = [30]
    s = O.Sim(1, 1, nc, [0.2])
    for k in range(6):
        s.set(O.F_EAVG + k, (0.05 if k < 3 else 1.0) * np.ones(s.shape(O.F_EAVG + k)))
    n = 400
    p = dict(icell=rng.integers(0, 30, (n, 1)).astype(np.int32), delta=rng.random((n, 1)),
             v=rng.standard_normal((n, 3)), weight=np.ones(n), charge=np.ones(n))
    s.set_particles(0, 0, p)
    dt = 0.02
    first = [((5,), (20,))]
    second = [((10,), (25,))]
    e2, e1 = s.push(dt, first, second)
    q = s.get_particles()
    assert 0 < e2 < e1 < n
    c = q["icell"][:, 0]
    assert np.all((c[:e2] >= 10) & (c[:e2] <= 25))
    # prePush moved every particle by dt/2 v / dx; the first selector sees that position
    x0 = p["icell"][:, 0] + p["delta"][:, 0] + 0.5 * dt * p["v"][:, 0] / 0.2
    pre_in = (np.floor(x0) >= 5) & (np.floor(x0) <= 20)
    assert e1 == int(pre_in.sum())
    # the rejected ones: prePush only, velocity untouched
    rej = q["weight"][e1:]
    key = lambda d: np.round(d["v"][:, 0], 15)
    orig = {float(a): (b, cc) for a, b, cc in zip(key(p), x0, pre_in)}
    for i in range(e1, n):
        x, inside = orig[float(np.round(q["v"][i, 0], 15))]
        assert not inside
        assert abs(q["icell"][i, 0] + q["delta"][i, 0] - x) < 1e-12
    # the accelerated ones changed velocity
    moved = {float(a) for a in key(p)}
    assert sum(float(np.round(v, 15)) not in moved for v in q["v"][:e1, 0]) == e1
