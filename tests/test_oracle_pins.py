"""Pins of the CPU oracle against what the mathematics and independent
libraries fix: closed forms, library functions (numpy / scipy / mpmath),
brute force on small inputs, invariants and published known-answer vectors.
The reference ships no tests or golden vectors for this path, so these are
the anchors of every parity claim."""
import ctypes as C
import math

import numpy as np
import pytest
import scipy.special as sp

from hpsim_amd import configs as cfg
from tests import helpers as H

ULP = np.finfo(np.float64).eps


@pytest.fixture(scope="module")
def L(oracle_lib):
    return oracle_lib.lib()


def vec(f, xs):
    return np.array([f(float(v)) for v in xs])


# --------------------------------------------------------------------------
# elementary functions
# --------------------------------------------------------------------------

def test_exp_within_one_ulp(L):
    x = np.random.default_rng(0).uniform(-700, 700, 20000)
    got = vec(L.orc_t_exp, x)
    assert np.max(np.abs(got - np.exp(x)) / np.exp(x)) <= 2 * ULP
    assert L.orc_t_exp(0.0) == 1.0
    assert L.orc_t_exp(800.0) == math.inf and L.orc_t_exp(-800.0) == 0.0
    assert math.isnan(L.orc_t_exp(math.nan))


def test_log_within_one_ulp(L):
    rng = np.random.default_rng(1)
    x = np.concatenate([10.0 ** rng.uniform(-300, 300, 10000), rng.uniform(0.5, 2.0, 10000),
                        [5e-324, 2.2e-308]])
    got = vec(L.orc_t_log, x)
    ref = np.log(x)
    assert np.all(np.abs(got - ref) <= 2 * ULP * np.maximum(np.abs(ref), 1e-300) + 1e-300)
    assert L.orc_t_log(1.0) == 0.0
    assert L.orc_t_log(0.0) == -math.inf
    assert math.isnan(L.orc_t_log(-1.0))


def test_sincos_against_mpmath(L):
    import mpmath
    mpmath.mp.prec = 120
    rng = np.random.default_rng(2)
    x = np.concatenate([rng.uniform(-1e4, 1e4, 3000), rng.uniform(-4, 4, 2000),
                        [0.0, 1e-300, np.pi / 2, np.pi, 1e6]])
    s, c = C.c_double(), C.c_double()
    for v in x:
        L.orc_t_sincos(float(v), C.byref(s), C.byref(c))
        es, ec = float(mpmath.sin(mpmath.mpf(float(v)))), float(mpmath.cos(mpmath.mpf(float(v))))
        assert abs(s.value - es) <= 2 * ULP * max(abs(es), 1e-300) + 1e-300 or abs(s.value - es) <= 2e-16
        assert abs(c.value - ec) <= 2 * ULP * max(abs(ec), 1e-300) + 1e-300 or abs(c.value - ec) <= 2e-16


def test_sinh_cosh(L):
    x = np.random.default_rng(3).uniform(-20, 20, 20000)
    assert np.max(np.abs(vec(L.orc_t_sinh, x) - np.sinh(x)) / np.maximum(np.abs(np.sinh(x)), 1e-300)) <= 3 * ULP
    assert np.max(np.abs(vec(L.orc_t_cosh, x) - np.cosh(x)) / np.cosh(x)) <= 3 * ULP
    assert L.orc_t_sinh(0.0) == 0.0 and L.orc_t_cosh(0.0) == 1.0


def test_fmod_is_exact(L):
    rng = np.random.default_rng(4)
    x = np.concatenate([rng.uniform(0, 1e5, 20000), [0.0, 2 * np.pi, 4 * np.pi, 1e-30]])
    got = vec(lambda v: L.orc_t_fmod_pos(v, 2 * np.pi), x)
    assert np.array_equal(got.view(np.uint64), np.fmod(x, 2 * np.pi).view(np.uint64))


# --------------------------------------------------------------------------
# Bessel and transit-time factor (simulation_kernel.cu:15-73, 290-308)
# --------------------------------------------------------------------------

def test_bessel_polynomials_within_abramowitz_stegun_bounds(L):
    x = np.concatenate([np.linspace(-3.74, 3.74, 801), np.linspace(3.76, 30, 400)])
    i0 = vec(L.orc_t_bessi0, x)
    i1 = vec(L.orc_t_bessi1, x)
    # A&S 9.8.1-9.8.4: relative errors below 1.6e-7 / 1.9e-7 / 8e-9 / 2.2e-7
    assert np.max(np.abs(i0 / sp.i0(x) - 1)) < 3e-7
    nz = np.abs(x) > 1e-3
    assert np.max(np.abs(i1[nz] / sp.i1(x[nz]) - 1)) < 3e-7
    # I1' = I0 - I1/x; limit 0.5 at 0
    xp = x[nz]
    assert np.max(np.abs(vec(L.orc_t_bessi1p, xp) - sp.ivp(1, xp)) / np.abs(sp.ivp(1, xp))) < 1e-6
    assert L.orc_t_bessi1p(0.0) == 0.5


def _gap(**kw):
    from oracle.oracle import GapParam
    g = GapParam()
    for k, v in kw.items():
        setattr(g, k, v)
    return g


def test_transit_time_factor_polynomial_and_clamp(L):
    t = [0.8, 0.05, -0.2, 0.3, 0.01, -0.002]
    s = [0.1, 0.02, -0.03, 0.004, 0.0005]
    g = _gap(fit_beta_center=0.1, fit_beta_min=0.05, **{f"fit_t{i}": v for i, v in enumerate(t)},
             **{f"fit_s{i + 1}": v for i, v in enumerate(s)})
    T, Tp, Sp = C.c_double(), C.c_double(), C.c_double()
    for beta in (0.06, 0.08, 0.1, 0.109):
        L.orc_t_ttf(beta, C.byref(g), C.byref(T), C.byref(Tp), C.byref(Sp))
        poly = np.polynomial.Polynomial(t)
        assert abs(T.value - poly(beta)) < 1e-15
        # T' and S' are beta^2 / (2 pi beta_g) times the beta-derivatives
        coef = beta * beta / (2 * np.pi * 0.1)
        assert abs(Tp.value - coef * poly.deriv()(beta)) < 1e-15
        spoly = np.polynomial.Polynomial([0.0] + s)
        assert abs(Sp.value + coef * spoly.deriv()(beta)) < 1e-15
    # clamp: above 1.1 beta_g the factor is evaluated at beta_g; below beta_min at beta_min
    L.orc_t_ttf(0.2, C.byref(g), C.byref(T), C.byref(Tp), C.byref(Sp))
    assert T.value == pytest.approx(np.polynomial.Polynomial(t)(0.1), abs=1e-15)
    L.orc_t_ttf(0.01, C.byref(g), C.byref(T), C.byref(Tp), C.byref(Sp))
    assert T.value == pytest.approx(np.polynomial.Polynomial(t)(0.05), abs=1e-15)


# --------------------------------------------------------------------------
# canonical reduction
# --------------------------------------------------------------------------

def test_canonical_sum_exact_on_integers_and_close_to_fsum(L):
    rng = np.random.default_rng(5)
    for n in (1, 255, 1024, 1025, 70001):
        v = rng.integers(-1000, 1000, n).astype(np.float64)
        assert L.orc_t_canon_sum(v.ctypes.data_as(C.POINTER(C.c_double)), n) == float(v.sum())
        w = rng.standard_normal(n)
        s = L.orc_t_canon_sum(w.ctypes.data_as(C.POINTER(C.c_double)), n)
        assert abs(s - math.fsum(w)) <= 4 * ULP * math.fsum(np.abs(w))


# --------------------------------------------------------------------------
# SCHEFF: ring Green's function and field sum (space_charge_kernel.cu)
# --------------------------------------------------------------------------

def test_green_table_matches_exact_ring_field(L):
    """fld1 = c1 * (E_r, E_z) of a charged ring; compare the Hastings
    polynomials of the reference with scipy's complete elliptic integrals."""
    nr, nz, dr, dz, freq, beta, num, cur = 8, 16, 1e-4, 2e-4, 201.25, 0.05, 1000, 0.01 / 939.294
    f1 = np.empty((nr + 1) * nr * nz * 2)
    L.orc_t_green_table(nr, nz, dr, dz, freq, beta, num, cur, 0, f1.ctypes.data_as(C.POINTER(C.c_double)))
    f1 = f1.reshape(nr + 1, nr, nz, 2)
    gm = 1 / math.sqrt(1 - beta * beta)
    dzg = dz * gm
    c1 = 5721.67 * cur / (freq * 1e6 * num)
    # A&S 17.3.34 / 17.3.36: |error| of the polynomial K, E below 3e-5 / 4e-5
    eK, eE = 3e-5, 4e-5
    for rp_i in range(nr + 1):
        for k in range(nr):
            for t in range(nz):
                rs = math.sqrt(0.5 * (k * k + (k + 1) ** 2)) * dr
                rp = rp_i * dr
                d = (t + 0.5) * dzg
                b = (rp + rs) ** 2
                c = (rp - rs) ** 2
                m = 4 * rs * rp / (b + d * d)
                K, E = sp.ellipk(m), sp.ellipe(m)
                a = math.sqrt(b + d * d)
                ez = d * E / (a * (c + d * d))
                assert abs(f1[rp_i, k, t, 1] / c1 - ez) <= eE * d / (a * (c + d * d)) * 1.01
                if rp_i == 0:
                    assert f1[rp_i, k, t, 0] == 0.0
                    # on axis: E_z of a ring is (pi/2) d / (rs^2 + d^2)^1.5
                    assert abs(ez - (math.pi / 2) * d / (rs * rs + d * d) ** 1.5) < 1e-12 * ez
                else:
                    R = (rs * rs - rp * rp + d * d) / (c + d * d)
                    er = (K - R * E) / (2 * rp * a)
                    bound = (eK + eE * abs(R)) / (2 * rp * a) * 1.01
                    assert abs(f1[rp_i, k, t, 0] / c1 - er) <= bound


def _ring_field(rs, rp, d):
    """E_r, E_z of a unit charged ring (radius rs) at (rp, axial offset d),
    from scipy's complete elliptic integrals, and error bounds of the
    Hastings polynomials (A&S 17.3.34 / 17.3.36) for the same terms."""
    eK, eE = 3e-5, 4e-5
    b = (rp + rs) ** 2
    c = (rp - rs) ** 2
    m = 4 * rs * rp / (b + d * d)
    K, E = sp.ellipk(m), sp.ellipe(m)
    a = math.sqrt(b + d * d)
    ez = d * E / (a * (c + d * d))
    bz = eE * abs(d) / (a * (c + d * d)) * 1.01
    if rp == 0:
        return 0.0, ez, 0.0, bz
    R = (rs * rs - rp * rp + d * d) / (c + d * d)
    er = (K - R * E) / (2 * rp * a)
    return er, ez, (eK + eE * abs(R)) / (2 * rp * a) * 1.01, bz


def test_green_table_adjacent_bunch_images(L):
    """With adj_bunch = k the table adds the rings of k neighbouring bunches
    on each side, one RF period beta*gamma*lambda apart in the bunch frame
    (space_charge_kernel.cu:74-94): the sum of 2k + 1 exact ring fields."""
    nr, nz, dr, dz, freq, beta, num, cur = 6, 12, 1.5e-3, 4e-3, 201.25, 0.04, 1000, 0.01 / 939.294
    adj = 2
    f1 = np.empty((nr + 1) * nr * nz * 2)
    L.orc_t_green_table(nr, nz, dr, dz, freq, beta, num, cur, adj,
                        f1.ctypes.data_as(C.POINTER(C.c_double)))
    f1 = f1.reshape(nr + 1, nr, nz, 2)
    gm = 1 / math.sqrt(1 - beta * beta)
    dzg = dz * gm
    pl = beta * gm * 299.792458 / freq
    c1 = 5721.67 * cur / (freq * 1e6 * num)
    for rp_i in range(nr + 1):
        for k in range(nr):
            for t in range(nz):
                rs = math.sqrt(0.5 * (k * k + (k + 1) ** 2)) * dr
                rp = rp_i * dr
                d = (t + 0.5) * dzg
                er = ez = ber = bez = 0.0
                for img in range(-adj, adj + 1):
                    e_r, e_z, b_r, b_z = _ring_field(rs, rp, d - img * pl)
                    er, ez, ber, bez = er + e_r, ez + e_z, ber + b_r, bez + b_z
                assert abs(f1[rp_i, k, t, 1] / c1 - ez) <= bez + 1e-13 * abs(ez)
                if rp_i == 0:
                    assert f1[rp_i, k, t, 0] == 0.0
                else:
                    assert abs(f1[rp_i, k, t, 0] / c1 - er) <= ber + 1e-13 * abs(er)
    # the images matter at this spacing: they change E_z by more than the bound
    f0 = np.empty_like(f1.ravel())
    L.orc_t_green_table(nr, nz, dr, dz, freq, beta, num, cur, 0,
                        f0.ctypes.data_as(C.POINTER(C.c_double)))
    assert np.max(np.abs(f1.ravel() - f0)[1::2] / c1) > 100 * 4e-5


def test_offmesh_field_is_gradient_of_bunch_chain_potential(L):
    """Off-mesh particles see the bunch, and with adj_bunch = k also its k
    neighbours on each side, as point charges one period pl apart
    (space_charge_kernel.cu:642-662): (r/d^3, z/d^3) summed over the chain is
    minus the gradient of the chain's potential sum 1/d, checked here by
    central differences."""
    rng = np.random.default_rng(12)
    pl = 0.0596
    ro, zo = C.c_double(), C.c_double()

    def phi(r, z, k):
        return sum(1.0 / math.hypot(r, z + j * pl) for j in range(-k, k + 1))

    for k in (0, 1, 3):
        for _ in range(40):
            r = rng.uniform(1e-3, 8e-3)
            z = rng.uniform(-0.025, 0.025)
            L.orc_t_offmesh_field(z, r, pl, k, C.byref(ro), C.byref(zo))
            h = 1e-6
            gr = -(phi(r + h, z, k) - phi(r - h, z, k)) / (2 * h)
            gz = -(phi(r, z + h, k) - phi(r, z - h, k)) / (2 * h)
            assert ro.value == pytest.approx(gr, rel=1e-6)
            assert zo.value == pytest.approx(gz, rel=1e-6, abs=1e-6 * abs(gr))
    # a wrong pl or a one-sided chain breaks the symmetry z -> -z
    L.orc_t_offmesh_field(0.01, 2e-3, pl, 3, C.byref(ro), C.byref(zo))
    a = (ro.value, zo.value)
    L.orc_t_offmesh_field(-0.01, 2e-3, pl, 3, C.byref(ro), C.byref(zo))
    assert ro.value == pytest.approx(a[0], rel=1e-14) and zo.value == pytest.approx(-a[1], rel=1e-14)


def test_adjacent_bunch_mesh_and_energy_cutoff(oracle_lib):
    """With adjacent bunches the remesh spans the largest radius about the
    beam's last all-particle centroid (x_avg, y_avg: stale until a getter or
    an aperture updates them; particles with loss != 0 excluded, lloss
    flags not) and one RF period beta*lambda (space_charge.cu:124-128,
    beam.cu:829-851); above the cutoff energy the bunch is alone again
    (space_charge.cpp:102-105) and the mesh goes back to 3 sigma."""
    O = oracle_lib
    n = 4096
    wl = cfg.config(0, num=n, periods=1)
    dist = [np.array(a) for a in cfg.make_distribution(wl.beam)]
    loss = np.zeros(n, np.uint32)
    lloss = np.zeros(n, np.uint32)
    dist[0][100], dist[2][100] = 0.03, 0.0      # transversely lost, far out
    loss[100] = 7
    dist[0][200], dist[2][200] = 0.0, -0.012    # longitudinally flagged, still counted
    lloss[200] = 1

    def run(adj, cutoff):
        b = O.Beam(n, 939.294, -1.0, 0.01)
        b.set_distribution(*dist, loss=loss, lloss=lloss)
        b.set_frequency(201.25)
        b.set_ref_w(0.75)
        b.get_avg("x")                    # x_avg of the untranslated beam
        b.translate("x", 2e-3)            # leaves x_avg stale
        st = b.stats()
        sc = O.SpaceCharge(32, 128, 0.05, adj)
        sc.set_adj_bunch_cutoff_w(cutoff)
        x0 = b.arrays()
        sc.start(b, 0.05)
        return b, sc.state(), st, x0

    b, s, st, a = run(2, 0.0)
    assert st["x_avg"] != pytest.approx(np.mean(a["x"][loss == 0]), abs=1e-4)
    ok = a["loss"] == 0
    r2 = (a["x"][ok] - st["x_avg"]) ** 2 + (a["y"][ok] - st["y_avg"]) ** 2
    assert s["dr"] * 32 == pytest.approx(math.sqrt(r2.max()), rel=1e-15)
    assert np.argmax(r2) == np.flatnonzero(ok).tolist().index(200)
    bst = b.stats()
    gamma = bst["w_avg_good"] / 939.294 + 1
    beta = math.sqrt(1 - 1 / gamma ** 2)
    assert s["dz"] * 128 == pytest.approx(beta * 299.792458 / 201.25, rel=1e-15)
    # cutoff above the beam energy: images on; below: off
    _, s2, _, _ = run(2, 0.8)
    assert s2["dr"] == s["dr"]
    b3, s3, _, _ = run(2, 0.7)
    assert s3["dr"] * 32 == pytest.approx(3 * b3.stats()["y_sig_good"], rel=1e-15)
    b4, s4, _, _ = run(0, 0.0)
    assert s4["dr"] == s3["dr"] and s4["dz"] == s3["dz"]
    assert H.compare_arrays(b3.arrays(), b4.arrays()) is None


def test_mesh_size_cutoff_follows_beam_energy(oracle_lib):
    """mesh_size_cutoff_w (space_charge.cpp:107-117): at each kick, a beam
    whose good mean energy is above the cutoff gets half the mesh in each
    direction and four times the kick interval; a drift reads the interval
    when it starts (simulation_engine.cu:128-130).  Observed element by
    element: the kick count of every drift, the mesh after every element."""
    O = oracle_lib
    n, mw, iv = 2048, 1.27, 0.01
    wl = cfg.config(0, num=n, periods=6)
    els = cfg.build_lattice(wl.lattice)
    b = O.Beam(n, 939.294, -1.0, 0.01)
    b.set_distribution(*cfg.make_distribution(wl.beam))
    b.set_frequency(201.25)
    b.set_ref_w(0.75)
    sc = O.SpaceCharge(32, 128, iv)
    sc.set_mesh_size_cutoff_w(mw)
    sim = O.Simulator(b, els, sc)
    ke = None           # good mean energy at the last kick
    crossed = False
    for i, e in enumerate(els):
        before = sc.state()
        sim.simulate(i, i)
        after = sc.state()
        kicks = after["kick_count"] - before["kick_count"]
        if e.kind == "Drift":
            want_iv = iv * 4 if (ke is not None and ke > mw) else iv
            assert before["interval"] == want_iv
            assert kicks == math.ceil(e.length / want_iv)
        else:
            assert kicks == 1
        ke = b.stats()["w_avg_good"]
        big = ke > mw
        crossed |= big
        assert (after["nx"], after["nz"]) == ((16, 64) if big else (32, 128))
    assert crossed and ke > mw


@pytest.mark.parametrize("nr,nz", [(4, 8), (7, 5), (70, 3)])
def test_field_sum_brute_force(L, nr, nz):
    """Canonical field sum vs a plain double loop of charge x Green's value
    (mesh widths below, between and above the 64 partial sums)."""
    rng = np.random.default_rng(6 + nr)
    bins = rng.integers(0, 2 ** 40, nr * nz).astype(np.int64)
    bins[rng.random(nr * nz) < 0.3] = 0
    f1 = rng.standard_normal((nr + 1) * nr * nz * 2)
    f2 = np.empty((nz + 1) * (nr + 1) * 2)
    L.orc_t_field(nr, nz, bins.ctypes.data_as(C.POINTER(C.c_int64)),
                  f1.ctypes.data_as(C.POINTER(C.c_double)), f2.ctypes.data_as(C.POINTER(C.c_double)))
    G = f1.reshape(nr + 1, nr, nz, 2)
    q = bins.astype(np.float64) * 2.0 ** -36
    for zp in range(nz + 1):
        for rp in range(nr + 1):
            ex, ez = [], []
            for zs in range(nz):
                for rs in range(nr):
                    # field point z index zp sits (zp - zs - 1/2) bins from source zs
                    off = zp - zs - 0.5
                    dzi = int(abs(off) - 0.5)
                    sgn = 1.0 if off > 0 else -1.0
                    ex.append(q[zs * nr + rs] * G[rp, rs, dzi, 0])
                    ez.append(sgn * q[zs * nr + rs] * G[rp, rs, dzi, 1])
            i = zp * (nr + 1) + rp
            assert f2[2 * i] == pytest.approx(math.fsum(ex), rel=1e-13, abs=1e-300)
            assert f2[2 * i + 1] == pytest.approx(math.fsum(ez), rel=1e-12, abs=1e-12 * max(map(abs, ez)))


# --------------------------------------------------------------------------
# element maps through the oracle engine
# --------------------------------------------------------------------------

def _single(O, elements, x=0.0, xp=0.0, y=0.0, yp=0.0, phi=0.0, w=10.0, n=2, sc=None):
    """Beam of n identical particles (index 0 is forced on axis)."""
    b = O.Beam(n, 939.294, -1.0, 0.0)
    b.set_distribution(*[np.full(n, v) for v in (x, xp, y, yp, phi, w)])
    b.set_frequency(201.25)
    b.set_ref_w(w)
    sim = O.Simulator(b, elements, sc)
    sim.simulate()
    return b.arrays()


def test_drift_closed_form(oracle_lib):
    O = oracle_lib
    L_ = 0.37
    a = _single(O, [cfg.Element("D", "Drift", length=L_)], x=1e-3, xp=2e-3, y=-1e-3, yp=5e-4)
    g = 10.0 / 939.294
    beta = math.sqrt(g * (g + 2)) / (g + 1)
    lam = 299.792458 / 201.25
    assert a["x"][1] == pytest.approx(1e-3 + 2e-3 * L_, rel=1e-15)
    assert a["y"][1] == pytest.approx(-1e-3 + 5e-4 * L_, rel=1e-15)
    assert a["xp"][1] == 2e-3 and a["w"][1] == 10.0
    assert a["phi"][1] == pytest.approx(2 * math.pi * L_ / (lam * beta), rel=1e-14)


def test_drift_aperture_loss_code_is_element_index(oracle_lib):
    O = oracle_lib
    el = [cfg.Element("D0", "Drift", length=0.1, aperture=0.01),
          cfg.Element("D1", "Drift", length=1.0, aperture=0.01)]
    a = _single(O, el, x=5e-3, xp=9e-3, n=3)
    # first half of D0 ends at 5e-3 + 9e-3 * 0.05 < 1 cm; D1 pushes it out: code 1;
    # a particle lost in element 0 would carry code 0 (= not lost, reference quirk)
    assert list(a["loss"]) == [0, 1, 1]
    a0 = _single(O, el[:1], x=9.9e-3, xp=1e-1, n=3)
    assert list(a0["loss"]) == [0, 0, 0] and a0["x"][1] == 9.9e-3


def test_quad_thick_lens_matrices(oracle_lib):
    """Two quarter-length steps per half equal the textbook thick-lens
    matrix; det = 1 (symplectic) in both planes."""
    O = oracle_lib
    G, Lq, w = 12.0, 0.08, 5.0
    el = [cfg.Element("Q", "Quad", length=Lq, gradient=G)]
    rays = {}
    for name, (x, xp) in {"a": (1e-6, 0.0), "b": (0.0, 1e-9)}.items():
        a = _single(O, el, x=x, xp=xp, y=x, yp=xp, w=w)
        rays[name] = (a["x"][1], a["xp"][1], a["y"][1], a["yp"][1])
    g1 = w / 939.294
    k = math.sqrt(abs(G * -1.0) / (math.sqrt(g1 * (g1 + 2)) * 939.294) * 299.792458)
    # charge -1 and G > 0: qG < 0 -> defocusing in x, focusing in y
    Mx = np.array([[math.cosh(k * Lq), math.sinh(k * Lq) / k], [k * math.sinh(k * Lq), math.cosh(k * Lq)]])
    My = np.array([[math.cos(k * Lq), math.sin(k * Lq) / k], [-k * math.sin(k * Lq), math.cos(k * Lq)]])
    gx = np.array([[rays["a"][0] / 1e-6, rays["b"][0] / 1e-9], [rays["a"][1] / 1e-6, rays["b"][1] / 1e-9]])
    gy = np.array([[rays["a"][2] / 1e-6, rays["b"][2] / 1e-9], [rays["a"][3] / 1e-6, rays["b"][3] / 1e-9]])
    assert np.allclose(gx, Mx, rtol=1e-9, atol=1e-12)
    assert np.allclose(gy, My, rtol=1e-9, atol=1e-12)
    assert abs(np.linalg.det(gx) - 1) < 1e-9 and abs(np.linalg.det(gy) - 1) < 1e-9
    # G = 0 is a drift
    a = _single(O, [cfg.Element("Q", "Quad", length=Lq, gradient=0.0)], x=1e-3, xp=1e-3)
    assert a["x"][1] == pytest.approx(1e-3 + 1e-3 * Lq, rel=1e-15)


def test_synchronous_particle_follows_design(oracle_lib):
    """Particle 0 (on axis, never kicked) gains V*T*cos(phi_s) per gap
    (Panofsky), so its energy equals the lattice's design output energy."""
    wl = cfg.config(0, num=64, periods=6)
    elems = cfg.build_lattice(wl.lattice)
    from tests.helpers import run_oracle
    beam, _, _ = run_oracle(wl, elems=elems)
    a = beam.arrays()
    assert a["w"][0] == pytest.approx(elems[-1].gap["energy_out"], rel=1e-12)
    gaps = [e for e in elems if e.kind == "RFGap-DTL"]
    g0 = gaps[0].gap
    assert g0["energy_out"] - 0.75 == pytest.approx(
        0.3 * cfg._ttf(g0["beta_center"], g0) * math.cos(math.radians(-30)), rel=1e-12)


def test_space_charge_kick_skips_every_quarter_of_nextpow2(oracle_lib):
    """Q3: ApplySpchKickKernel's thread 0 exits, so particle indices that are
    multiples of NextPow2(N)/4 are never kicked (space_charge_kernel.cu:586-588)."""
    O = oracle_lib
    n = 3000
    wl = cfg.config(0, num=n, periods=1)
    dist = cfg.make_distribution(wl.beam)
    b = O.Beam(n, 939.294, -1.0, 0.05)
    b.set_distribution(*dist)
    b.set_frequency(201.25)
    b.set_ref_w(0.75)
    before = b.arrays()
    sc = O.SpaceCharge(32, 128, 0.05)
    sc.start(b, 0.05)
    after = b.arrays()
    stride = 4096 // 4
    changed = after["xp"] != before["xp"]
    skip = np.arange(n) % stride == 0
    assert not changed[skip].any()
    assert changed[~skip].mean() > 0.5


def test_zero_current_kick_is_near_identity(oracle_lib):
    """Q9: with zero current the field is zero, but the kick recomputes
    gamma* non-relativistically, a relative change of order (beta gamma*)^4."""
    O = oracle_lib
    n = 2048
    wl = cfg.config(0, num=n, periods=1)
    b = O.Beam(n, 939.294, -1.0, 0.0)
    b.set_distribution(*cfg.make_distribution(wl.beam))
    b.set_frequency(201.25)
    b.set_ref_w(0.75)
    before = b.arrays()
    O.SpaceCharge(32, 128, 0.05).start(b, 0.05)
    after = b.arrays()
    assert np.max(np.abs(after["xp"] - before["xp"]) / np.maximum(np.abs(before["xp"]), 1e-30)) < 1e-9
    assert np.max(np.abs(after["w"] / before["w"] - 1)) < 1e-12


def test_deposit_conserves_charge(oracle_lib):
    """The fixed-point histogram holds one unit of weight per particle on
    the mesh (each particle's four weights sum to 1) up to truncation."""
    O = oracle_lib
    n = 4096
    wl = cfg.config(0, num=n, periods=1)
    b = O.Beam(n, 939.294, -1.0, 0.01)
    b.set_distribution(*cfg.make_distribution(wl.beam))
    b.set_frequency(201.25)
    b.set_ref_w(0.75)
    sc = O.SpaceCharge(32, 128, 0.05)
    sc.start(b, 0.05)
    _, _, bins = sc.tables()
    on = sc.state()["num_on_mesh"]
    assert 0 < on <= n
    assert abs(int(bins.sum()) - on * 2 ** 36) <= 4 * on


def test_stale_loss_denominator(oracle_lib):
    """Q1: the SCHEFF moments divide by N - num_loss, where num_loss only
    changes in UpdateLoss (getters, apertures); calling a getter between two
    simulate() calls therefore changes the later result."""
    O = oracle_lib
    wl = cfg.config(0, num=2048, periods=4)
    elems = cfg.build_lattice(wl.lattice)
    dist = cfg.make_distribution(wl.beam)
    out = []
    for call_getter in (False, True):
        b = O.Beam(2048, 939.294, -1.0, 0.01)
        b.set_distribution(*dist)
        b.set_frequency(201.25)
        b.set_ref_w(0.75)
        sim = O.Simulator(b, elems, O.SpaceCharge(32, 128, 0.05))
        sim.simulate(0, 7)
        if call_getter:
            assert b.get_loss_num("t") > 0
        sim.simulate(8, 15)
        out.append(b.arrays()["x"])
    assert not np.array_equal(out[0], out[1])


# --------------------------------------------------------------------------
# synthetic workload generator
# --------------------------------------------------------------------------

def test_philox_known_answer_vectors():
    """Random123 philox4x32-10 known-answer vectors."""
    cases = [((0, 0, 0, 0), (0, 0), (0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8)),
             ((0xffffffff,) * 4, (0xffffffff, 0xffffffff), (0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd)),
             ((0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344), (0xa4093822, 0x299f31d0),
              (0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1))]
    for c, k, want in cases:
        r = cfg.philox4x32(*[np.array([v], np.uint32) for v in c], *k)
        assert tuple(int(v[0]) for v in r) == want


def test_truncated_gaussian_moments():
    z = cfg.truncated_normal(200000, 3, seed=11)
    assert np.max(np.abs(z)) <= 3.0
    # variance of a unit normal truncated at 3 sigma: 1 - 6 phi(3) / (2 Phi(3) - 1)
    var = 1 - 6 * math.exp(-4.5) / math.sqrt(2 * math.pi) / math.erf(3 / math.sqrt(2))
    assert abs(z.mean()) < 0.01 and abs(z.var() - var) < 0.01
    assert np.array_equal(z, cfg.truncated_normal(200000, 3, seed=11))


# --------------------------------------------------------------------------
# waterbag / DC initial beams (beam.cpp:391-537)
# --------------------------------------------------------------------------

def _rms_twiss(u, v):
    cuu = np.mean(u * u) - np.mean(u) ** 2
    cvv = np.mean(v * v) - np.mean(v) ** 2
    cuv = np.mean(u * v) - np.mean(u) * np.mean(v)
    e = math.sqrt(cuu * cvv - cuv * cuv)
    return e, -cuv / e, cuu / e


def test_libc_rand_is_the_glibc_generator():
    """The samplers draw from the C library's rand(); glibc's srand(1) stream
    starts 1804289383, 846930886, 1681692777 (the reference's Linux build
    draws the same stream)."""
    libc = C.CDLL(None)
    libc.srand(1)
    assert [libc.rand() for _ in range(3)] == [1804289383, 846930886, 1681692777]


def test_waterbag_moments_match_twiss_closed_form(oracle_lib):
    O = oracle_lib
    """A uniform 6-ball has <r_i^2> = 1/8, so each plane's rms emittance is
    e/8 with the input alpha and beta (x in cm, phase mapped back through
    beta*lambda, W through 2 W_s)."""
    n = 200001
    ax, bx, ex, ay, by, ey, az, bz, ez = -1.2, 20.0, 0.003, 0.8, 15.0, 0.004, 0.3, 30.0, 0.02
    phis, ws, f, m = -0.5, 0.75, 201.25, 939.294
    b = O.Beam(n, m, -1.0, 0.0)
    b.set_waterbag(ax, bx, ex, ay, by, ey, az, bz, ez, phis, ws, f, 7)
    a = b.arrays()
    assert a["x"][0] == 0 and a["xp"][0] == 0 and a["y"][0] == 0 and a["yp"][0] == 0
    assert b.get_frequency() == f
    gm = ws / m + 1
    k = 2 * math.pi / (math.sqrt(1 - 1 / gm ** 2) * 29979.2458 / f)
    planes = [(a["x"][1:] * 100, a["xp"][1:], ex, ax, bx),
              (a["y"][1:] * 100, a["yp"][1:], ey, ay, by),
              (-(a["phi"][1:] - phis) / k, (a["w"][1:] - ws) / (2 * ws), ez, az, bz)]
    for u, v, e, al, bt in planes:
        er, alr, btr = _rms_twiss(u, v)
        assert abs(er * 8 / e - 1) < 0.01
        assert abs(alr - al) < 0.02 and abs(btr / bt - 1) < 0.01
    # every point lies inside the joint 6-ball: sum of the normalised
    # Courant-Snyder invariants (gamma u^2 + 2 alpha u v + beta v^2) / e <= 1
    inv = sum(((1 + al * al) / bt * u * u + 2 * al * u * v + bt * v * v) / e
              for u, v, e, al, bt in planes)
    assert np.all(inv <= 1 + 1e-9) and inv.max() > 0.99
    # same seed, same beam; another seed, another beam
    b2 = O.Beam(n, m, -1.0, 0.0)
    b2.set_waterbag(ax, bx, ex, ay, by, ey, az, bz, ez, phis, ws, f, 7)
    assert np.array_equal(b2.arrays()["phi"], a["phi"])
    b2.set_waterbag(ax, bx, ex, ay, by, ey, az, bz, ez, phis, ws, f, 8)
    assert not np.array_equal(b2.arrays()["phi"], a["phi"])


def test_dc_beam_ramp_and_4ball_moments(oracle_lib):
    O = oracle_lib
    n = 100000
    ax, bx, ex, ay, by, ey = 0.5, 12.0, 0.002, -0.4, 18.0, 0.0025
    dphi, phis, ws = 1.5, 0.2, 0.75
    b = O.Beam(n, 939.294, -1.0, 0.0)
    b.set_dc(ax, bx, ex, ay, by, ey, dphi, phis, ws, 3)
    a = b.arrays()
    i = np.arange(n, dtype=np.float64)
    assert np.array_equal(a["phi"], (2.0 * (i + 1.0) / n - 1.0) * dphi + phis)
    assert np.all(a["w"] == ws)
    for u, v, e, al, bt in ((a["x"][1:] * 100, a["xp"][1:], ex, ax, bx),
                            (a["y"][1:] * 100, a["yp"][1:], ey, ay, by)):
        er, alr, btr = _rms_twiss(u, v)
        # uniform 4-ball: <r_i^2> = 1/6
        assert abs(er * 6 / e - 1) < 0.01
        assert abs(alr - al) < 0.02 and abs(btr / bt - 1) < 0.01


def test_openmp_oracle_is_bit_identical_to_serial(oracle_lib, tmp_path):
    """liboracle_omp.so (OpenMP loops, used for the all-core CPU baseline)
    reproduces the serial checker bit for bit on a space-charge run."""
    import os
    import subprocess
    import sys
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    out = tmp_path / "omp.npz"
    code = (
        "import sys, numpy as np\n"
        f"sys.path.insert(0, {root!r})\n"
        "from oracle import oracle as O\n"
        "O.use_openmp()\n"
        "from tests.helpers import workload, run_oracle\n"
        "b, _, _ = run_oracle(workload(0, num=6000, periods=3))\n"
        f"np.savez({str(out)!r}, **b.arrays())\n")
    env = dict(os.environ, OMP_NUM_THREADS="4")
    subprocess.run([sys.executable, "-c", code], check=True, env=env, cwd=root)
    from tests.helpers import run_oracle, workload
    ser, _, _ = run_oracle(workload(0, num=6000, periods=3))
    a, o = ser.arrays(), np.load(out)
    for k in a:
        assert np.array_equal(a[k].view(np.uint8), o[k].view(np.uint8)), k


# --------------------------------------------------------------------------
# dipole and buncher maps (simulation_kernel.cu:603-874)
# --------------------------------------------------------------------------

def test_tan_within_a_few_ulp(L):
    xs = np.linspace(-1.4, 1.4, 20001)
    got = vec(L.orc_t_tan, xs)
    assert np.max(np.abs(got - np.tan(xs)) / np.abs(np.spacing(np.tan(xs)))) <= 4


def _track(O, els, x, xp, y, yp, phi, w, freq=201.25):
    b = O.Beam(len(x), 939.294, -1.0, 0.0)
    b.set_distribution(x, xp, y, yp, phi, w)
    b.set_frequency(freq)
    b.set_ref_w(w[0])
    O.Simulator(b, els, None).simulate()
    return b.arrays()


def _dipole_matrices(O, dip, w):
    """Linear transfer blocks from four small unit offsets (the maps are
    linear in x, x', y, y' at the design energy)."""
    eps = 1e-7
    z = np.zeros(5)
    x, xp, y, yp = z.copy(), z.copy(), z.copy(), z.copy()
    x[1], xp[2], y[3], yp[4] = eps, eps, eps, eps
    a = _track(O, [dip], x, xp, y, yp, z, np.full(5, w))
    mx = np.array([[a["x"][1], a["x"][2]], [a["xp"][1], a["xp"][2]]]) / eps
    my = np.array([[a["y"][3], a["y"][4]], [a["yp"][3], a["yp"][4]]]) / eps
    return mx, my, a


def test_sector_dipole_matrices(oracle_lib):
    """Field index n: x focuses with k^2 = h^2 (1 - n), y with h^2 n; n = 0 is
    the textbook sector bend; the two halves compose to the full arc."""
    rho, th, w = 1.5, math.radians(30), 0.75
    for n in (0.0, 0.5):
        dip = cfg.Element("B", "Dipole", angle=th,
                          params=dict(rho=rho, kinetic_energy=w, field_index=n))
        mx, my, a = _dipole_matrices(oracle_lib, dip, w)
        L_ = rho * th
        kx, ky = math.sqrt(1 - n) / rho, math.sqrt(n) / rho
        ex = [[math.cos(kx * L_), math.sin(kx * L_) / kx],
              [-kx * math.sin(kx * L_), math.cos(kx * L_)]]
        ey = ([[1, L_], [0, 1]] if n == 0 else
              [[math.cos(ky * L_), math.sin(ky * L_) / ky], [-ky * math.sin(ky * L_), math.cos(ky * L_)]])
        assert np.allclose(mx, ex, atol=1e-8) and np.allclose(my, ey, atol=1e-8)
        assert abs(np.linalg.det(mx) - 1) < 1e-8 and abs(np.linalg.det(my) - 1) < 1e-8
        g = w / 939.294 + 1
        beta = math.sqrt(1 - 1 / g ** 2)
        assert a["phi"][0] == pytest.approx(2 * math.pi * L_ / (beta * 299.792458 / 201.25),
                                            rel=1e-14)


def test_dipole_edge_focusing_thin_lens(oracle_lib):
    """A pole-face rotation b with no fringe field (K1 = 0) is a thin lens:
    x' += tan(b)/rho x, y' -= tan(b)/rho y, at entrance and exit."""
    rho, th, w, b = 2.0, 1e-9, 0.75, math.radians(12)
    dip = cfg.Element("B", "Dipole", angle=th,
                      params=dict(rho=rho, kinetic_energy=w, edge_angle_in=b))
    mx, my, _ = _dipole_matrices(oracle_lib, dip, w)
    assert mx[1, 0] == pytest.approx(math.tan(b) / rho, rel=1e-6)
    assert my[1, 0] == pytest.approx(-math.tan(b) / rho, rel=1e-6)


def test_buncher_energy_gain_and_adiabatic_damping(oracle_lib):
    """On axis the gap gives dW = |q| V cos(h phi + phase); a particle with
    x = 0 keeps beta*gamma*x' (x' scales by bg_in/bg_out)."""
    w, v, ph0 = 0.75, 0.1, 0.3
    bun = cfg.Element("U", "Buncher", params=dict(voltage=v, frequency=201.25, phase=ph0, on=True))
    phi = np.array([0.0, 0.2, -0.5, 1.1])
    xp = np.array([0.0, 1e-3, -2e-3, 5e-4])
    a = _track(oracle_lib, [bun], np.zeros(4), xp, np.zeros(4), np.zeros(4), phi, np.full(4, w))
    dw = v * np.cos(phi + ph0)
    assert np.allclose(a["w"] - w, dw, rtol=1e-12, atol=1e-15)
    bg = lambda e: np.sqrt((e / 939.294) * (e / 939.294 + 2))
    assert np.allclose(a["xp"][1:], xp[1:] * bg(w) / bg(w + dw[1:]), rtol=1e-12)
    off = cfg.Element("U", "Buncher", params=dict(voltage=v, frequency=201.25, phase=ph0, on=False))
    b = _track(oracle_lib, [off], np.zeros(4), xp, np.zeros(4), np.zeros(4), phi, np.full(4, w))
    assert np.array_equal(b["w"], np.full(4, w))
