"""Pins of the CPU FP64 oracle against what the mathematics fixes.

The reference ships no tests, golden vectors or fixtures (its only hook, TEST_MULTIGRID_MASK,
writes CSVs at run time and needs FFTW-MPI), and it cannot be built here (no fftw3.h). So each
oracle function is pinned here by a closed form, an invariant, a brute-force computation or a
special case chosen so that a plausible mistake (dropped term, wrong sign, wrong index,
transposed operand) fails.
"""
import numpy as np
import pytest

from oracle import oracle as O
from qsf_amd import qsf_types as T

RNG = np.random.default_rng(7)


def crand(*shape):
    return RNG.normal(size=shape) + 1j * RNG.normal(size=shape)


# ---------------------------------------------------------------------------------- FFT
@pytest.mark.parametrize("n", [8, 12, 16, 24, 48, 64, 96, 768])
def test_fft_matches_long_double_dft(n):
    x = crand(n)
    ref = O.dft_bruteforce(x, -1)
    assert np.max(np.abs(O.fft1d(x, -1) - ref)) <= 1e-13 * np.max(np.abs(ref))
    refb = O.dft_bruteforce(x, +1)
    assert np.max(np.abs(O.fft1d(x, +1) - refb)) <= 1e-13 * np.max(np.abs(refb))


def test_fft_sign_convention_single_mode():
    # forward transform of e^{+2πi k0 j/n} is n·δ_{k,k0}: fixes the sign of the exponent
    n, k0 = 16, 3
    x = np.exp(2j * np.pi * k0 * np.arange(n) / n)
    X = O.fft1d(x, -1)
    e = np.zeros(n, complex)
    e[k0] = n
    assert np.allclose(X, e, atol=1e-12)


def test_fft_parseval_and_linearity():
    n = 1536
    x, y = crand(n), crand(n)
    X = O.fft1d(x)
    assert np.isclose(np.sum(np.abs(X) ** 2), n * np.sum(np.abs(x) ** 2), rtol=1e-13)
    assert np.allclose(O.fft1d(2 * x - 3j * y), 2 * X - 3j * O.fft1d(y), atol=1e-10)


def test_grid_fourier_roundtrip_and_transposed_p_layout():
    # 2D main region: P layout is (n1, n0) (FFTW_MPI_TRANSPOSED_OUT, grid.h:242-248)
    d = T.grid_desc(2, [8, 12], [0.3, 0.5])
    g = O.GridSet(d)[0]
    x = crand(8, 12)
    g.set_psi(x)
    g.fourier(T.REP_P)
    P = g.psi().reshape(12, 8)
    assert np.allclose(P, np.fft.fft2(x).T, atol=1e-12)
    g.fourier(T.REP_X)
    assert np.allclose(g.psi().reshape(8, 12), x, atol=1e-14)


# ---------------------------------------------------------------------------------- borders
def test_box_border_1d_points():
    # BOX<s> in 1D is ±1 exactly at centred index s-1 and -s (borders.h:48-52, borders_g.h:8)
    s = 5
    vals = {i: O.border_vec_d(T.BORDER_BOX, s, 1, [i])[0] for i in range(-8, 8)}
    assert vals[s - 1] == 1.0 and vals[-s] == -1.0
    assert all(v == 0.0 for i, v in vals.items() if i not in (s - 1, -s))


def test_box_border_2d_faces_and_corners():
    s = 4
    b = lambda i, j: O.border_vec_d(T.BORDER_BOX, s, 2, [i, j])
    assert list(b(s - 1, 0)) == [1.0, 0.0]
    assert list(b(-s, 2)) == [-1.0, 0.0]
    assert list(b(1, -s)) == [0.0, -1.0]
    r = 1 / np.sqrt(2.0)
    assert np.array_equal(b(s - 1, s - 1), [r, r])      # corner: (1,1)/√2 (borders_t.h:7-13)
    assert np.array_equal(b(-s, s - 1), [-r, r])
    assert list(b(s, 0)) == [0.0, 0.0] and list(b(0, 0)) == [0.0, 0.0]


def test_box_border_3d_face():
    s = 3
    v = O.border_vec_d(T.BORDER_BOX, s, 3, [s - 1, 0, 1])
    assert list(v) == [1.0, 0.0, 0.0]
    v = O.border_vec_d(T.BORDER_BOX, s, 3, [0, 0, -s])
    assert list(v) == [0.0, 0.0, -1.0]


def test_region_border_constants():
    # N2S sits at p_NS=25 nodes for |other| < p_SD=14 (borders.h:30-36, 111-120)
    assert list(O.border_vec_d(T.BORDER_N2S, 0, 2, [24, 13])) == [1.0, 0.0]
    assert list(O.border_vec_d(T.BORDER_N2S, 0, 2, [24, 14])) == [0.0, 0.0]
    assert list(O.border_vec_d(T.BORDER_N2S, 0, 2, [-25, -14])) == [-1.0, 0.0]
    # S2CAP at p_CAP=40
    assert list(O.border_vec_d(T.BORDER_S2CAP, 0, 2, [3, 39])) == [0.0, 1.0]
    # D2CAP spans p_SD..p_CAP along the other axis
    assert list(O.border_vec_d(T.BORDER_D2CAP, 0, 2, [39, 20])) == [1.0, 0.0]
    assert list(O.border_vec_d(T.BORDER_D2CAP, 0, 2, [39, 5])) == [0.0, 0.0]


# ---------------------------------------------------------------------------------- absorbers
def test_cap_closed_form():
    nCAP = 16
    g = O.GridSet(T.grid_desc(2, 64, 0.5, absorber=T.ABS_CAP, nCAP=nCAP))[0]
    eta = (3.0 / nCAP) ** 4
    # at the edge (neg_dist 0) nodes = nCAP, S·η = 81 (absorbers.h:36-54)
    assert np.isclose(g.absorber_mask([0, -40]), np.exp(-81.0), rtol=1e-14)
    assert g.absorber_mask([-nCAP, -nCAP]) == 1.0
    assert g.absorber_mask([-20, -30]) == 1.0
    a, b = 3, 5  # nodes = nCAP - a etc.
    want = np.exp(-(((nCAP - a) ** 4 + (nCAP - b) ** 4)) * eta)
    assert np.isclose(g.absorber_mask([-a, -b]), want, rtol=1e-14)
    assert np.isclose(g.absorber_absorb(0.05, [-a, -b]),
                      np.exp(-0.05 * ((nCAP - a) ** 4 + (nCAP - b) ** 4) * eta), rtol=1e-14)


def test_cos18_closed_form():
    w, tau = 10, 0.05
    g = O.GridSet(T.grid_desc(1, 128, 0.1, absorber=T.ABS_COS18, nCAP=w, absorber_tau=tau))[0]
    for dist in range(0, 14):
        node = w - dist
        want = np.cos(0.5 * np.pi * node / w) ** 0.125 if node > 0 else 1.0
        if node == w:
            want = 0.0                     # the edge node: cos(π/2) is 0
        assert np.isclose(g.absorber_mask([-dist]), want, rtol=1e-14, atol=1e-300)
    # widths whose rounded π/2·node/nCAP lands past π/2 still give a mask in [0, 1], never NaN
    for w2 in range(1, 64):
        g2 = O.GridSet(T.grid_desc(1, 256, 0.1, absorber=T.ABS_COS18, nCAP=w2, absorber_tau=tau))[0]
        vals = [g2.absorber_mask([-d]) for d in range(0, w2 + 1)]
        assert all(0.0 <= v <= 1.0 for v in vals), w2
        assert vals[0] == 0.0
    # two half-substeps apply exactly one full mask per tau
    m = g.absorber_mask([-2])
    assert np.isclose(g.absorber_absorb(tau / 2, [-2]) ** 2, m, rtol=1e-14)


# ---------------------------------------------------------------------------------- potentials
def test_potential_closed_forms():
    e = O.GridSet(T.grid_desc(2, 8, 0.2, potential=T.POT_EBERLY, Ncharge=2, Echarge=-1,
                              Nsoft=0.5, Esoft=0.33))[0]
    x, y = 0.7, -1.3
    want = 1 / np.sqrt((x - y) ** 2 + 0.33) - 2 / np.sqrt(x * x + 0.5) - 2 / np.sqrt(y * y + 0.5)
    assert np.isclose(e.potential([x, y]), want, rtol=1e-14)
    assert np.isclose(e.potential_partial(T.AXIS_Y, [x, y]), -2 / np.sqrt(x * x + 0.5), rtol=1e-14)
    assert e.potential_partial(T.AXIS_XY, [x, y]) == 0.0
    es = O.GridSet(T.grid_desc(3, 4, 0.3, potential=T.POT_ECKHARDT_SACHA, Ncharge=3,
                               Echarge=-1, Nsoft=1, Esoft=1))[0]
    c = [0.4, -0.9, 1.7]
    ee = lambda a, b: 1 / np.sqrt((a - b) ** 2 + a * b + 1)
    want = ee(c[0], c[1]) + ee(c[1], c[2]) + ee(c[2], c[0]) - sum(3 / np.sqrt(v * v + 1) for v in c)
    assert np.isclose(es.potential(c), want, rtol=1e-14)


# ---------------------------------------------------------------------------------- fields
def test_field_closed_form_and_vector_potential():
    E0, w, nc = 0.0534, 0.057, 4
    f = T.field_desc(T.AXIS_X, E0, w, nc, phase_in_pi=0.25)
    T_end = 2 * np.pi * nc / w
    assert np.isclose(O.max_pulse_duration([f]), T_end, rtol=1e-15)
    t = np.array([0.0, 10.0, 123.4, T_end * 0.5, T_end + 1.0])
    F = O.coupling_sequence([f], T.GAUGE_LENGTH, t)[:, 0]
    want = E0 * np.sin(w * t + 0.25 * np.pi) * np.sin(w * t / (2 * nc)) ** 2
    want[t > T_end] = 0.0
    assert np.allclose(F, want, rtol=1e-14, atol=1e-17)
    # velocity gauge: right-Riemann A(t) = -∫F converges to the closed-form integral
    dt = 0.01
    ts = np.arange(0, T_end, dt)
    A = O.coupling_sequence([f], T.GAUGE_VELOCITY, ts)[:, 0]
    from scipy.integrate import quad
    Fq = lambda s: E0 * np.sin(w * s + 0.25 * np.pi) * np.sin(w * s / (2 * nc)) ** 2
    for k in (len(ts) // 3, len(ts) - 1):
        assert abs(A[k] - (-quad(Fq, 0, ts[k], limit=200)[0])) < 5 * dt * E0
    # field on XY acts on both coordinates
    fxy = T.field_desc(T.AXIS_XY, E0, w, nc)
    C = O.coupling_sequence([fxy], T.GAUGE_LENGTH, np.array([50.0]))[0]
    assert C[0] == C[1] and C[2] == 0.0


# ---------------------------------------------------------------------------------- evolution
def free_gaussian(x, t, x0=0.0, p0=0.0, s=1.0):
    # exact free evolution of exp(-(x-x0)^2/(2s^2)+i p0 x), normalised
    a = s * s + 1j * t
    norm = (np.pi * s * s) ** -0.25 * s / np.sqrt(a)
    return norm * np.exp(-((x - x0 - p0 * t) ** 2) / (2 * a) + 1j * p0 * (x - x0) - 0.5j * p0 * p0 * t)


def test_free_gaussian_spreads_exactly():
    # V=0: the kinetic phase is exact; this pins dp, the P coordinates, the FFT sign and 1/m
    n, dx, dt, steps = 512, 0.1, 0.05, 40
    gs = O.GridSet(T.grid_desc(1, n, dx))
    x = -0.5 * dx * (n - 1) + dx * np.arange(n)
    psi0 = free_gaussian(x, 0.0, p0=1.5)
    prop = O.Propagator(gs, T.prop_desc(T.MODE_IM, 1, dt), [])
    g = gs[0]
    g.set_psi(psi0)
    # real-time free steps through the single-grid primitives (X: identity phase)
    for _ in range(steps):
        g.fourier(T.REP_P)
        g.evolve(T.MODE_RE, T.REP_P, dt)
        g.fourier(T.REP_X)
    want = free_gaussian(x, steps * dt, p0=1.5)
    assert np.max(np.abs(g.psi() - want)) < 1e-10


def test_ground_state_matches_dense_diagonalisation():
    # IM propagation converges to the lowest eigenvector of the same Fourier-grid H
    n, dx, dt = 128, 0.4, 0.01
    d = T.grid_desc(1, n, dx, potential=T.POT_COULOMB, Ncharge=1, Echarge=-1, Nsoft=1)
    gs = O.GridSet(d)
    x = -0.5 * dx * (n - 1) + dx * np.arange(n)
    gs[0].set_psi(np.exp(-x * x / 2) + 0j)
    pd = T.prop_desc(T.MODE_IM, 1, dt, max_steps=3000,
                     outputs=T.OUT_NORM | T.OUT_EPOT | T.OUT_EKIN | T.OUT_ETOT | T.OUT_EDIFF
                     | T.OUT_NORMALIZE)
    prop = O.Propagator(gs, pd, [])
    rec = prop.run()
    E_im = rec[-1, T.REC_ETOT]
    p = 2 * np.pi * np.fft.fftfreq(n, d=dx)
    F = np.fft.fft(np.eye(n), axis=0)
    Tm = np.linalg.solve(F, np.diag(0.5 * p * p) @ F)
    H = Tm.real + np.diag(-1 / np.sqrt(x * x + 1))
    E0 = np.linalg.eigvalsh(0.5 * (H + H.T))[0]
    assert abs(E_im - E0) < 2e-5
    assert abs(E0 - (-0.6698)) < 2e-3   # literature value of the 1D soft-core atom
    assert abs(rec[-1, T.REC_NORM] - 1.0) < 1e-6


def test_norm_conservation_and_reversibility():
    n, dx, dt = 64, 0.3, 0.05
    d = T.grid_desc(2, n, dx, potential=T.POT_EBERLY, Ncharge=2, Echarge=-1, Nsoft=0.5, Esoft=0.33)
    g = O.GridSet(d)[0]
    psi0 = crand(n * n) * 0.01
    g.set_psi(psi0)
    n0 = g.average(T.REP_X, 0)
    C = [0.03, 0.03, 0.0]
    for sgn in (+1, -1):
        for _ in range(10):
            g.evolve(T.MODE_RE, T.REP_X, sgn * dt / 2)
            g.fourier(T.REP_P)
            g.evolve(T.MODE_RE, T.REP_P, sgn * dt, C)
            g.fourier(T.REP_X)
            g.evolve(T.MODE_RE, T.REP_X, sgn * dt / 2)
        if sgn > 0:
            assert abs(g.average(T.REP_X, 0) / n0 - 1) < 1e-13
            C = C  # the backward pass reverses the same substeps in mirrored order
    assert np.max(np.abs(g.psi() - psi0)) < 1e-12 * np.max(np.abs(psi0))


def test_gauge_equivalence_free_electron():
    # length and velocity gauge densities agree for a free electron in a pulse (signs of x·F
    # and p·A, A = -∫F): a sign error moves the packet the wrong way.
    n, dx, dt = 1024, 0.2, 0.02
    f = T.field_desc(T.AXIS_X, 0.05, 0.3, 2)
    x = -0.5 * dx * (n - 1) + dx * np.arange(n)
    dens = []
    for gauge in (T.GAUGE_LENGTH, T.GAUGE_VELOCITY):
        gs = O.GridSet(T.grid_desc(1, n, dx))
        gs[0].set_psi(free_gaussian(x, 0.0))
        prop = O.Propagator(gs, T.prop_desc(T.MODE_RE, 1, dt, gauge=gauge), [f])
        prop.run()
        dens.append(np.abs(gs[0].psi()) ** 2)
    xs = [np.sum(x * d) / np.sum(d) for d in dens]
    assert abs(xs[0] - xs[1]) < 2e-2
    # classical drift: x(T) = ∫_0^T A dt' (p0 = 0), A = -∫F
    ts = np.arange(0, 2 * np.pi * 2 / 0.3 + 2 * dt, dt / 8)
    F = 0.05 * np.sin(0.3 * ts) * np.sin(0.3 * ts / 4) ** 2
    A = -np.cumsum(F) * (dt / 8)
    drift = np.sum(A) * (dt / 8)
    assert abs(xs[0] - drift) < 0.05 * abs(drift) + 2e-2
    assert abs(drift) > 0.1


# ---------------------------------------------------------------------------------- splitting
def test_split_tables_and_mps_quirk():
    # Order-2 MultiProductSplit returns S(dt/2)^2 ψ (propagator.h:389 quirk): compare with two
    # Strang steps of dt/2 driven through the same coupling sequence.
    n, dx, dt = 64, 0.3, 0.04
    d = T.grid_desc(2, n, dx, potential=T.POT_EBERLY, Ncharge=2, Echarge=-1, Nsoft=0.5, Esoft=0.33)
    f = T.field_desc(T.AXIS_XY, 0.1, 0.057, 3)
    psi0 = crand(n * n) * 0.01
    a = O.GridSet(d)
    a[0].set_psi(psi0)
    pa = O.Propagator(a, T.prop_desc(T.MODE_RE, 2, dt), [f])
    b = O.GridSet(d)
    b[0].set_psi(psi0)
    pb = O.Propagator(b, T.prop_desc(T.MODE_RE, 1, dt / 2), [f])
    for _ in range(5):
        pa.make_step()
        pb.make_step()
        pb.make_step()
    assert np.isclose(pa.time, pb.time, rtol=1e-14)
    assert np.allclose(pa.coupling(), pb.coupling(), rtol=1e-13, atol=0)
    assert np.max(np.abs(a[0].psi() - b[0].psi())) < 1e-13


# ---------------------------------------------------------------------------------- multigrid
def test_multigrid_partition_of_unity_2d():
    # One postCompute with only the main region populated splits ψ exactly into the regions:
    # m(a,b) + m_b·u(1−v/2) + m_a·v(1−u/2) + uv(2−u/2−v/2) = 1 (u=1−m_a, v=1−m_b).
    n, nCAP = 64, 8
    d = T.grid_desc(2, n, 0.5, absorber=T.ABS_CAP, nCAP=nCAP, multigrid=1)
    gs = O.GridSet(d)
    assert len(gs) == 4
    psi = crand(n, n)
    gs[0].set_psi(psi)
    gs.post_compute()
    total = gs[0].psi().reshape(n, n).copy()
    total += np.fft.ifft(gs[1].psi().reshape(n, n), axis=0)        # X free: axis 0 in P
    total += np.fft.ifft(gs[2].psi().reshape(n, n), axis=1)        # Y free: axis 1 in P
    total += np.fft.ifft2(gs[3].psi().reshape(n, n))               # XY free
    assert np.max(np.abs(total - psi)) < 1e-13
    # the main region keeps exactly m(a,b)ψ (bit-exact multiply by the CAP mask)
    nd = lambda i: -i if i < n // 2 else i - n + 1
    mrow = np.array([[gs[0].absorber_mask([nd(i), nd(j)]) for j in range(n)] for i in range(n)])
    assert np.array_equal(gs[0].psi().reshape(n, n), psi * mrow)
