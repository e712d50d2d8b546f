"""Pins of the oracle's probability current and surface flux by mathematics alone.

The current (wf/grid.h:467-528), its velocity-gauge term (propagator.h:249-257) and the flux
sum (wf/grid.h:559-573) are checked against facts that do not come from re-reading those lines:

* closed forms of the central-difference current of a modulated plane wave psi = a_i e^{ikx}:
  j_i = a_i (a_{i+1} + a_{i-1}) sin(k dx) / (2 dx), zero where the strict linear-index rule
  (c - s > 0 and c + s < m) removes the stencil, and the exact value where the inner axis
  wraps into the next row;
* the flux through a BOX of known geometry (2s - 2 face nodes of weight 1 and two corners of
  weight 1/sqrt(2) per face) times dV/dx0, on a grid with dx0 != dx1;
* continuity: for a free packet the norm that leaves a closed BOX equals the time integral of
  the flux through it (split-operator propagation is exact for a free particle; what remains
  is the O((k dx)^2) difference between the central-difference current and the spectral
  kinetic operator);
* gauge invariance: the same packet driven by the same field in length and in velocity gauge
  gives the same flux only if the velocity-gauge current carries +A|psi|^2.

Each pin fails for a flipped sign of Im, a factor 2 in 1/(2 dx), a wrong dV/dx0 or a wrong sign
(or absence) of the A|psi|^2 term.
"""
import numpy as np

from oracle import oracle as O
from qsf_amd import qsf_types as T


def plane_wave_1d(n, dx, k, amp):
    x = -0.5 * dx * (n - 1) + dx * np.arange(n)
    return amp * np.exp(1j * k * x)


def test_current_of_modulated_plane_wave_1d():
    n, dx = 128, 0.3
    k = 2 * np.pi * 9 / (n * dx)
    i = np.arange(n)
    a = 1.0 + 0.3 * np.cos(2 * np.pi * i / n) + 0.1 * np.sin(6 * np.pi * i / n)
    g = O.GridSet(T.grid_desc(1, n, dx))
    g[0].set_psi(plane_wave_1d(n, dx, k, a))
    j = g[0].current_map(1)[:, 0]
    expect = np.zeros(n)
    c = np.arange(2, n - 1)      # strict rule: c - 1 > 0 and c + 1 < n
    expect[c] = a[c] * (a[c + 1] + a[c - 1]) * np.sin(k * dx) / (2 * dx)
    assert np.allclose(j, expect, rtol=1e-13, atol=1e-14)
    assert j[0] == 0.0 and j[1] == 0.0 and j[n - 1] == 0.0


def test_current_of_plane_wave_2d_with_row_wrap():
    n0, n1, dx0, dx1 = 16, 24, 0.3, 0.2
    k0 = 2 * np.pi * 3 / (n0 * dx0)
    k1 = 2 * np.pi * 5 / (n1 * dx1)
    x0 = -0.5 * dx0 * (n0 - 1) + dx0 * np.arange(n0)
    x1 = -0.5 * dx1 * (n1 - 1) + dx1 * np.arange(n1)
    psi = np.exp(1j * (k0 * x0[:, None] + k1 * x1[None, :])).ravel()
    g = O.GridSet(T.grid_desc(2, [n0, n1], [dx0, dx1]))
    g[0].set_psi(psi)
    j = g[0].current_map(2).reshape(n0, n1, 2)
    m = n0 * n1
    c = np.arange(m).reshape(n0, n1)
    # axis 0 (stride n1): sin(k0 dx0)/dx0 wherever c - n1 > 0 and c + n1 < m
    ok0 = (c - n1 > 0) & (c + n1 < m)
    assert np.allclose(j[..., 0][ok0], np.sin(k0 * dx0) / dx0, rtol=1e-13)
    assert np.all(j[..., 0][~ok0] == 0.0)
    # axis 1 (stride 1): interior of a row sin(k1 dx1)/dx1; at the row ends the neighbour is the
    # next (previous) row's first (last) node: Im(e^{i(k0 dx0 + k1 dx1)} - e^{-i k1 dx1}) / (2 dx1)
    inner = j[1:-1, 1:-1, 1]
    assert np.allclose(inner, np.sin(k1 * dx1) / dx1, rtol=1e-13)
    wrap = (np.sin(k0 * dx0 + k1 * dx1) + np.sin(k1 * dx1)) / (2 * dx1)
    assert np.allclose(j[1:-1, n1 - 1, 1], wrap, rtol=1e-12)
    assert np.allclose(j[1:-1, 0, 1], wrap, rtol=1e-12)


def test_flux_through_box_closed_form_2d():
    n0, n1, dx0, dx1, s = 64, 48, 0.3, 0.2, 6
    k0 = 2 * np.pi * 4 / (n0 * dx0)
    i0 = np.arange(n0)
    a = 1.0 + 0.5 * np.cos(2 * np.pi * i0 / n0 + 0.3)      # depends on axis 0 only
    x0 = -0.5 * dx0 * (n0 - 1) + dx0 * i0
    psi = (a * np.exp(1j * k0 * x0))[:, None] * np.ones(n1)[None, :]
    g = O.GridSet(T.grid_desc(2, [n0, n1], [dx0, dx1]))
    g[0].set_psi(psi.ravel())
    F = g[0].flux([(T.BORDER_BOX, s)])[0]

    def jx(ic):   # axis-0 current at centred index ic (i0 = ic + n0/2)
        i = ic + n0 // 2
        return a[i] * (a[i + 1] + a[i - 1]) * np.sin(k0 * dx0) / (2 * dx0)

    # +x face at centred index s-1 and -x face at -s, each 2s nodes along axis 1: 2s-2 of
    # weight 1 and two corners of weight 1/sqrt(2); the axis-1 current is zero inside the box
    weight = (2 * s - 2) + 2 / np.sqrt(2.0)
    expect = (dx0 * dx1 / dx0) * weight * (jx(s - 1) - jx(-s))
    assert abs(F - expect) <= 1e-13 * max(abs(expect), 1e-300)
    assert abs(expect) > 1e-3


def test_flux_through_box_closed_form_1d():
    n, dx, s = 256, 0.1, 30
    k = 2 * np.pi * 11 / (n * dx)
    i = np.arange(n)
    a = np.exp(-((i - n / 2 - 10) / 40.0) ** 2)
    g = O.GridSet(T.grid_desc(1, n, dx))
    g[0].set_psi(plane_wave_1d(n, dx, k, a))
    F = g[0].flux([(T.BORDER_BOX, s)])[0]
    j = lambda c: a[c] * (a[c + 1] + a[c - 1]) * np.sin(k * dx) / (2 * dx)
    expect = j(n // 2 + s - 1) - j(n // 2 - s)        # dV/dx0 = 1 in 1D
    assert abs(F - expect) <= 1e-13 * abs(expect)


def test_continuity_free_packet_through_closed_box():
    """Norm leaving BOX<s> over K steps equals the time integral of the flux through it."""
    n, dx, s, dt, steps = 2048, 0.05, 60, 0.004, 1000
    x = -0.5 * dx * (n - 1) + dx * np.arange(n)
    sigma = 0.5
    psi0 = np.exp(-x * x / (4 * sigma * sigma)) + 0j
    psi0 /= np.sqrt(np.sum(np.abs(psi0) ** 2) * dx)
    d = T.grid_desc(1, n, dx)
    f = [T.field_desc(T.AXIS_X, 0.0, 1.0, 1.0)]        # a zero field: the free particle
    pd = T.prop_desc(T.MODE_RE, 1, dt, gauge=T.GAUGE_LENGTH, outputs=T.OUT_NORM | T.OUT_FLUX,
                     borders=[(T.BORDER_BOX, s)])
    g = O.GridSet(d)
    g[0].set_psi(psi0)
    p = O.Propagator(g, pd, f)
    p.start()
    dens = []
    for _ in range(steps):
        p.step()
    rec = p.records()
    # the border nodes sit at centred s-1 and -s: the inside counts the nodes strictly between
    # them and half of each border node (the current there is the flux through that node)
    lo, hi = n // 2 - s, n // 2 + s - 1
    def inside(ps):
        w = np.abs(ps) ** 2
        return (np.sum(w[lo + 1:hi]) + 0.5 * (w[lo] + w[hi])) * dx
    n_start = inside(psi0)
    n_end = inside(g[0].psi())
    flux = rec[:, T.REC_FLUX0]
    k = min(len(flux), steps + 1)
    integral = np.sum(0.5 * (flux[1:k] + flux[:k - 1])) * dt
    lost = n_start - n_end
    assert lost > 0.3                       # most of the packet has left the box
    assert abs(lost - integral) <= 5e-3 * lost, (lost, integral)


def test_velocity_gauge_current_term_from_gauge_invariance():
    """The flux of one packet under one field, in length and in velocity gauge (psi_L =
    e^{iA x} psi_V with A(0) = 0): equal only with the +A|psi|^2 term of the velocity gauge."""
    n, dx, s, dt, steps = 1024, 0.05, 20, 0.005, 400
    x = -0.5 * dx * (n - 1) + dx * np.arange(n)
    # off centre, so |psi|^2 differs between the two faces and the A term does not cancel
    psi0 = np.exp(-(x - 1.0) ** 2 / 8.0) + 0j
    psi0 /= np.sqrt(np.sum(np.abs(psi0) ** 2) * dx)
    f = [T.field_desc(T.AXIS_X, 0.5, 1.0, 1)]
    out = {}
    term = np.zeros(steps + 1)
    for gauge in (T.GAUGE_LENGTH, T.GAUGE_VELOCITY):
        g = O.GridSet(T.grid_desc(1, n, dx))
        g[0].set_psi(psi0)
        pd = T.prop_desc(T.MODE_RE, 1, dt, gauge=gauge, outputs=T.OUT_NORM | T.OUT_FLUX,
                         borders=[(T.BORDER_BOX, s)])
        p = O.Propagator(g, pd, f)
        p.start()
        for k in range(1, steps + 1):
            p.step()
            if gauge == T.GAUGE_VELOCITY:
                # the A|psi|^2 part of this step's flux: A times the density difference of the
                # outward (+, centred s-1) and inward (-, centred -s) border nodes
                w = np.abs(g[0].psi()) ** 2
                term[k] = p.coupling()[0] * (w[n // 2 + s - 1] - w[n // 2 - s])
        out[gauge] = p.records()
    fl = out[T.GAUGE_LENGTH][:, T.REC_FLUX0]
    fv = out[T.GAUGE_VELOCITY][:, T.REC_FLUX0]
    assert fl.shape == fv.shape == (steps + 1,)
    big = np.max(np.abs(term))
    assert big > 0.02
    # a missing term leaves |term| of difference, a flipped one 2|term|
    assert np.max(np.abs(fl - fv)) <= 0.05 * big, (np.max(np.abs(fl - fv)), big)
