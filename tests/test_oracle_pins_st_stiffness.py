"""Known answers for the oracle's surface-tension terms and DF/CD stiffness.

Each expected value is worked out by hand from the reference's equations on a small
configuration whose geometry makes the sums short (two or three particles on the axes, or a
symmetric star), using only the smoothing kernels W, G, STW, STG (pinned on their own in
test_oracle_kernels.py: normalisation, derivative relations, support):

- the surface-tension factor Delta_i = sum_j m_j v_ij . grad W_ij
  (src/MPS_Computer/MPSSPHNonPressureUtil.cuh:425-465);
- the pair factor g_ij and the CG right-hand side b_i with viscosity and XSPH switched off
  (:467-521);
- the surface-tension part of the CG operator A.p (:523-585, with ComputeFinalCGAp :635-648);
- the DF stiffness kappa = min(delta dt, rho/rho0 + dt delta - 0.8) alpha / dt^2 on both sides
  of the min, and the CD stiffness kappa = (rho/rho0 + dt delta - 1) alpha / dt^2 for e > 0 and
  e <= 0, with the reported mean error (src/MPS_Computer/MPSSPHDFSPHUtil.cuh:259-340,
  MPSSPHDFSPHUtil.cu:331-333,417-419).

Every test also evaluates the same closed form with one plausible slip at a time (a sign, a
factor 1/2, 1/rho_bar, dt against dt^2, 0.8 against 1, the wrong branch of the min, m_i for
m_j) and asserts that the slip moves the answer by far more than the tolerance, so the oracle
could not pass with that slip.
"""
import numpy as np

from mps_cuda_amd import _capi as capi
from oracle.binding import OracleSim, kernel_eval

R = 0.01
H = 4 * R
RHO0 = 1000.0
V = H ** 3 / 48.0 * 3.141592          # SPHMaterial::SetParam volume (MPSSPHMaterial.cpp:16)
M = RHO0 * V                           # material mass


def kern(d):
    """(W, G, STW, STG) of the cubic family at distance d, support H, FP64."""
    w, g, stw, stg = kernel_eval(64, 1, np.atleast_1d(np.asarray(d, np.float64)), 1.0 / H)
    return w[0], g[0], stw[0], stg[0]


def make(pos, vel, mass=None, sigma=0.0, mu=0.0, xsph=0.0, dt=1e-3, **cfgkw):
    cfg = capi.default_config()
    for i in range(3):
        cfg.hash_size[i] = 64
    cfg.cell_size = H
    cfg.kernel = 1
    for k, v in cfgkw.items():
        setattr(cfg, k, v)
    o = OracleSim(cfg, 64)
    o.set_physics(dt, (0.0, 0.0, 0.0), (-10, -10, -10), (10, 10, 10))
    o.set_fluid_material(H, RHO0, mu, xsph, sigma)
    pos = np.asarray(pos, np.float64)
    o.upload_fluid(pos, np.full(len(pos), R), np.asarray(vel, np.float64), mass)
    for ph in ["HASH", "NEIGHBORS", "BOUNDARY_VOLUME", "DENSITY", "ALPHA", "GRAVITY"]:
        o.run_phase(ph)
    return o


def assert_slips_detected(want, slips, rtol):
    for name, v in slips.items():
        err = np.max(np.abs(np.asarray(v) - want)) / np.max(np.abs(want))
        assert err > 1e3 * rtol, f"slip '{name}' is not distinguishable ({err:.2e})"


# ------------------------------------------------------------------ surface tension


def test_surface_tension_factor_three_particles():
    """Delta_0 with neighbours on +x and +y, distinct masses and velocities:
    Delta_0 = m1 (v0-v1).(G1 x01) + m2 (v0-v2).(G2 x02) = -m1 (a-u) d1 G1 + m2 w d2 G2."""
    d1, d2 = 0.021, 0.027                 # 1 and 2 are neighbours of 0 (and of each other)
    a, u, w = 0.3, -0.7, 0.45
    m = np.array([M, 1.5 * M, 0.6 * M])
    o = make([[0, 0, 0], [d1, 0, 0], [0, d2, 0]], [[a, 0, 0], [u, 0, 0], [0, w, 0]], mass=m, sigma=0.2)
    o.cg_operator(np.zeros((3, 3)), 0.0)   # runs the factor and the CG set-up on this state
    W1, G1, _, _ = kern(d1)
    W2, G2, _, _ = kern(d2)
    want = -m[1] * (a - u) * d1 * G1 + m[2] * w * d2 * G2
    got = o.get("FLUID_FACTOR_ST")[0]
    np.testing.assert_allclose(got, want, rtol=1e-12)
    assert_slips_detected(want, {
        "m_i for m_j": -m[0] * (a - u) * d1 * G1 + m[0] * w * d2 * G2,
        "volume for mass (Q24 reversed)": -V * (a - u) * d1 * G1 + V * w * d2 * G2,
        "v_ji for v_ij": m[1] * (a - u) * d1 * G1 - m[2] * w * d2 * G2,
        "W for G": -m[1] * (a - u) * d1 * W1 + m[2] * w * d2 * W2,
    }, 1e-12)


def _st_pair(d, u, m0, m1, dt, sigma):
    """Two particles: 0 at rest at the origin, 1 at (d,0,0) moving with (u,0,0).
    Returns the hand-derived (Delta_0, Delta_1, rho, g_01) and the pieces of g."""
    W0 = kern(0.0)[0]
    Wd, Gd, w, stg = kern(d)
    rho = RHO0 * V * (W0 + Wd)               # both particles: rho0 (V W(0) + V W(d))
    inv_d = 2.0 / (rho + rho)                # 1/rho_bar
    D0 = m1 * u * Gd * d                     # m1 (v0 - v1).(G x01) = m1 (-u) G (-d)
    D1 = m0 * u * Gd * d                     # m0 (v1 - v0).(G x10) = m0 u G d
    vs = u * d * stg                         # v01 . (STG x01) = (-u)(STG)(-d)
    g = (w + dt * (vs - (D0 + D1) * inv_d * 0.5 * w)) * inv_d
    parts = dict(w=w, vs=vs, D=D0 + D1, inv_d=inv_d)
    return D0, D1, rho, g, parts


def test_surface_tension_rhs_two_particles():
    """b_i = v_i + (dt sigma / m_i)(-sum_j m_bar g_ij x_ij) with mu = XSPH = 0, particles
    approaching at q = 0.65 (STG != 0) with distinct masses:
    b_0 = (dt sigma/m0) m_bar g d x^,  b_1 = (u,0,0) - (dt sigma/m1) m_bar g d x^."""
    d, u, dt, sigma = 0.026, -0.8, 1e-3, 0.3
    m0, m1 = M, 1.5 * M
    o = make([[0, 0, 0], [d, 0, 0]], [[0, 0, 0], [u, 0, 0]], mass=[m0, m1], sigma=sigma, dt=dt)
    o.cg_operator(np.zeros((2, 3)), 0.0)
    D0, D1, rho, g, p = _st_pair(d, u, m0, m1, dt, sigma)
    np.testing.assert_allclose(o.get("FLUID_DENSITY"), [rho, rho], rtol=1e-12)
    np.testing.assert_allclose(o.get("FLUID_FACTOR_ST"), [D0, D1], rtol=1e-12)
    mb = 0.5 * (m0 + m1)
    want = np.array([[dt * sigma / m0 * mb * g * d, 0, 0], [u - dt * sigma / m1 * mb * g * d, 0, 0]])
    np.testing.assert_allclose(o.get("FLUID_CG_B"), want, rtol=1e-12, atol=1e-15)

    def b_with(gv, scale0=dt * sigma / m0 * mb, scale1=dt * sigma / m1 * mb):
        return np.array([[scale0 * gv * d, 0, 0], [u - scale1 * gv * d, 0, 0]])

    w, vs, D, inv_d = p["w"], p["vs"], p["D"], p["inv_d"]
    assert_slips_detected(want, {
        "+(Delta_i+Delta_j)": b_with((w + dt * (vs + D * inv_d * 0.5 * w)) * inv_d),
        "missing 0.5": b_with((w + dt * (vs - D * inv_d * w)) * inv_d),
        "missing inner 1/rho_bar": b_with((w + dt * (vs - D * 0.5 * w)) * inv_d),
        "missing outer 1/rho_bar": b_with(w + dt * (vs - D * inv_d * 0.5 * w)),
        "dt^2 inside g": b_with((w + dt * dt * (vs - D * inv_d * 0.5 * w)) * inv_d),
        "m_i for m_bar": b_with(g, dt * sigma, dt * sigma),
        "dt^2 outside": b_with(g, dt * dt * sigma / m0 * mb, dt * dt * sigma / m1 * mb),
    }, 1e-12)


def test_surface_tension_operator_two_particles():
    """Surface-tension part of A.p (mu = XSPH = 0):
    (A p)_0 = (1+f) p_0 + dt (dt sigma/m0) m_bar g_01 (p_0 - p_1)."""
    d, u, dt, sigma, f = 0.026, -0.8, 1e-3, 0.3, 0.01
    m0, m1 = M, 1.5 * M
    o = make([[0, 0, 0], [d, 0, 0]], [[0, 0, 0], [u, 0, 0]], mass=[m0, m1], sigma=sigma, dt=dt)
    p = np.array([[0.3, -0.2, 0.5], [-0.4, 0.6, 0.1]])
    out = o.cg_operator(p, f)
    _, _, _, g, _ = _st_pair(d, u, m0, m1, dt, sigma)
    mb = 0.5 * (m0 + m1)
    off0 = dt * (dt * sigma / m0) * mb * g * (p[0] - p[1])
    off1 = dt * (dt * sigma / m1) * mb * g * (p[1] - p[0])
    got = out - (1 + f) * p
    np.testing.assert_allclose(got, [off0, off1], rtol=1e-9)
    assert_slips_detected(np.array([off0, off1]), {
        "dt^2 dropped to dt": [off0 / dt, off1 / dt],
        "dt^3": [off0 * dt, off1 * dt],
        "sigma/m_j": [off0 * m0 / m1, off1 * m1 / m0],
        "p_j - p_i": [-off0, -off1],
    }, 1e-9)


# ------------------------------------------------------------------ DF / CD stiffness


def _first_eval(o, df):
    """One evaluation of the DF (or CD) loop: tolerance huge and minimum 0 so the loop ends
    at l = 1 right after evaluating; kappa and delta are that evaluation's."""
    st = o.run_phase("DIVERGENCE" if df else "CONSTANT_DENSITY")
    assert (st.l_df if df else st.l_cd) == 1
    return st


def _pair_state(d, a, b, dt):
    """Two particles at (0,0,0) and (d,0,0) with velocities (a,0,0), (b,0,0): both have
    rho/rho0 = V (W(0) + W(d)), delta = V G d (b - a), alpha = 1 / (2 V^2 G^2 d^2)."""
    W0 = kern(0.0)[0]
    Wd, Gd, _, _ = kern(d)
    ratio = V * (W0 + Wd)
    delta = V * Gd * d * (b - a)
    alpha = 1.0 / (2.0 * (V * Gd * d) ** 2)
    return ratio, delta, alpha


def test_cd_stiffness_two_particles_compressing_and_separating():
    d, dt = 0.01, 0.01
    for (a, b), positive in (((5.0, -5.0), True), ((-5.0, 5.0), False)):
        o = make([[0, 0, 0], [d, 0, 0]], [[a, 0, 0], [b, 0, 0]], dt=dt, cd_tol=1e30, cd_min_iter=0)
        st = _first_eval(o, df=False)
        ratio, delta, alpha = _pair_state(d, a, b, dt)
        np.testing.assert_allclose(o.get("FLUID_DELTA"), [delta, delta], rtol=1e-12)
        np.testing.assert_allclose(o.get("FLUID_ALPHA"), [alpha, alpha], rtol=1e-12)
        e = ratio + dt * delta - 1.0
        assert (e > 0) == positive
        want = e * alpha / dt ** 2 if e > 0 else 0.0
        np.testing.assert_allclose(o.get("FLUID_KAPPA"), [want, want], rtol=1e-12)
        np.testing.assert_allclose(st.err_cd[0], max(e, 0.0), rtol=1e-12)   # mean of two equal errors
        if positive:
            assert_slips_detected(want, {
                "0.8 for 1": (ratio + dt * delta - 0.8) * alpha / dt ** 2,
                "dt for dt^2": e * alpha / dt,
                "alpha divides": e / alpha / dt ** 2,
                "DF min branch": min(dt * delta, ratio + dt * delta - 0.8) * alpha / dt ** 2,
            }, 1e-12)


def test_df_stiffness_two_particles_density_side_of_min():
    """rho/rho0 = 0.29 < 0.8: the min takes rho/rho0 + dt delta - 0.8."""
    d, dt, a, b = 0.01, 0.01, 5.0, -5.0
    o = make([[0, 0, 0], [d, 0, 0]], [[a, 0, 0], [b, 0, 0]], dt=dt, df_tol=1e30, df_min_iter=0)
    st = _first_eval(o, df=True)
    ratio, delta, alpha = _pair_state(d, a, b, dt)
    assert ratio < 0.8
    e = min(delta * dt, ratio + dt * delta - 0.8)
    assert e == ratio + dt * delta - 0.8 and e > 0
    want = e * alpha / dt ** 2
    np.testing.assert_allclose(o.get("FLUID_KAPPA"), [want, want], rtol=1e-12)
    np.testing.assert_allclose(st.err_df[0], e, rtol=1e-12)
    assert_slips_detected(want, {
        "1 for 0.8": (ratio + dt * delta - 1.0) * alpha / dt ** 2,
        "other branch": delta * dt * alpha / dt ** 2,
        "dt for dt^2": e * alpha / dt,
    }, 1e-12)


def _star(d, s):
    """A centre particle at rest and six at distance d on the axes moving radially with
    speed s. Centre: sum_j grad W = 0, so alpha = 1/(6 V^2 G^2 d^2); rho/rho0 =
    V (W(0) + 6 W(d)); delta = sum_j V G (x_cj . v_cj) = 6 V G d s."""
    dirs = np.array([[1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1], [0, 0, -1]], np.float64)
    pos = np.vstack([np.zeros((1, 3)), d * dirs])
    vel = np.vstack([np.zeros((1, 3)), s * dirs])
    W0 = kern(0.0)[0]
    Wd, Gd, _, _ = kern(d)
    return pos, vel, V * (W0 + 6 * Wd), 6 * V * Gd * d * s, 1.0 / (6.0 * (V * Gd * d) ** 2)


def test_df_stiffness_star_both_sides_of_min():
    dt = 1e-3
    # d = h/4: the centre sits above 0.8 rho0, so the min takes the divergence term delta dt
    pos, vel, ratio, delta, alpha = _star(0.01, -1.0)
    o = make(pos, vel, dt=dt, df_tol=1e30, df_min_iter=0)
    _first_eval(o, df=True)
    assert ratio > 0.8 and delta > 0
    np.testing.assert_allclose(o.get("FLUID_ALPHA")[0], alpha, rtol=1e-12)
    np.testing.assert_allclose(o.get("FLUID_DELTA")[0], delta, rtol=1e-12)
    want = delta * dt * alpha / dt ** 2
    np.testing.assert_allclose(o.get("FLUID_KAPPA")[0], want, rtol=1e-12)
    assert_slips_detected(want, {
        "density branch": (ratio + dt * delta - 0.8) * alpha / dt ** 2,
        "max for min": max(delta * dt, ratio + dt * delta - 0.8) * alpha / dt ** 2,
    }, 1e-12)
    # d = 0.3 h: below 0.8 rho0, the min takes rho/rho0 + dt delta - 0.8 (still > 0)
    pos, vel, ratio, delta, alpha = _star(0.012, -1.0)
    o = make(pos, vel, dt=dt, df_tol=1e30, df_min_iter=0)
    _first_eval(o, df=True)
    e = ratio + dt * delta - 0.8
    assert ratio < 0.8 and 0 < e < delta * dt
    want = e * alpha / dt ** 2
    np.testing.assert_allclose(o.get("FLUID_KAPPA")[0], want, rtol=1e-12)
    assert_slips_detected(want, {"divergence branch": delta * dt * alpha / dt ** 2,
                                 "1 for 0.8": (ratio + dt * delta - 1.0) * alpha / dt ** 2}, 1e-12)
    # expanding star: e <= 0 on both branches, kappa clamps to 0
    pos, vel, ratio, delta, alpha = _star(0.012, 1.0)
    o = make(pos, vel, dt=dt, df_tol=1e30, df_min_iter=0)
    _first_eval(o, df=True)
    assert o.get("FLUID_KAPPA")[0] == 0.0
