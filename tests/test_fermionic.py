"""operation<AntiSymmetrize<...>> (wf/grid.h:611-723): the oracle pinned by the reference's loops
and by the algebra of antisymmetrisers; the device entry point bit-exact against the oracle."""
import itertools

import numpy as np
import pytest

from oracle import fermionic as F


def crand(m, seed):
    r = np.random.default_rng(seed)
    return r.normal(size=m) + 1j * r.normal(size=m)


@pytest.mark.parametrize("dim,particles,n", [(2, 2, 7), (3, 2, 5), (3, 3, 6), (1, 2, 9), (2, 3, 4),
                                             (3, 2, (4, 6, 6)), (3, 2, (7, 3, 3))])
def test_oracle_equals_reference_loops(dim, particles, n):
    m = int(np.prod(n)) if isinstance(n, tuple) else n ** dim
    psi = crand(m, 1)
    assert np.array_equal(F.antisymmetrize(psi, n, dim, particles), F.antisymmetrize_loops(psi, n, dim, particles))


def test_oracle_fp32_order_is_float32_arithmetic():
    # the complex64 form adds in float32, in the kernel's order: (b + c) - d - e - f, then a + t;
    # check one node against explicit float32 scalar arithmetic
    n = 5
    psi = crand(n ** 3, 11).astype(np.complex64)
    got = F.antisymmetrize(psi, n, 3, 3, dtype=np.complex64).reshape(n, n, n)
    a = psi.reshape(n, n, n)
    i, j, k = 4, 2, 1
    f32 = np.float32
    for part in ("real", "imag"):
        g = lambda *ix: f32(getattr(a[ix], part))
        t = f32(g(j, k, i) + g(k, i, j))
        t = f32(t - g(j, i, k))
        t = f32(t - g(i, k, j))
        t = f32(t - g(k, j, i))
        assert f32(g(i, j, k) + t) == getattr(got[i, j, k], part)


def test_antisymmetry_and_projection():
    n = 6
    a = F.antisymmetrize(crand(n ** 3, 2), n, 3, 3).reshape(n, n, n)
    for perm in itertools.permutations(range(3)):
        sign = np.linalg.det(np.eye(3)[list(perm)])
        assert np.array_equal(a.transpose(perm), sign * a)
    # the unnormalised antisymmetriser A satisfies A(A psi) = 6 A psi (3 particles), 2 A psi (2)
    np.testing.assert_allclose(F.antisymmetrize(a, n, 3, 3).reshape(n, n, n), 6 * a, rtol=1e-14, atol=1e-14)
    b = F.antisymmetrize(crand(n * n, 3), n, 2, 2).reshape(n, n)
    assert np.array_equal(b.T, -b) and np.array_equal(np.diag(b), np.zeros(n))
    np.testing.assert_allclose(F.antisymmetrize(b, n, 2, 2).reshape(n, n), 2 * b, rtol=1e-15)
    # a symmetric state is annihilated
    s = crand(n * n, 4).reshape(n, n)
    assert not np.any(F.antisymmetrize(s + s.T, n, 2, 2))


@pytest.mark.gpu
@pytest.mark.parametrize("dim,particles,n", [(2, 2, 64), (2, 2, 48), (3, 2, 16), (3, 3, 16), (3, 3, 12), (1, 2, 64),
                                             (3, 2, (8, 16, 16)), (3, 2, (24, 12, 12))])
@pytest.mark.parametrize("prec", [0, 1])
def test_device_antisymmetrize_bit_exact(dim, particles, n, prec):
    from qsf_amd import hip
    from qsf_amd import qsf_types as T
    print(f"[case] antisymmetrize dim={dim} particles={particles} n={n} prec={prec}", flush=True)
    shape = list(n) if isinstance(n, tuple) else [n] * dim
    precision = T.FP64 if prec == 0 else T.FP32_MIXED
    ctx = hip.Context(0)
    gs = hip.GridSet(ctx, T.grid_desc(dim, shape, 0.3, precision=precision))
    psi = crand(int(np.prod(shape)), 5)
    if precision == T.FP64:
        want = F.antisymmetrize(psi, n, dim, particles)
    else:   # the FP32 state: float32 additions in the kernel's order
        psi = psi.astype(np.complex64).astype(np.complex128)
        want = F.antisymmetrize(psi, n, dim, particles, dtype=np.complex64).astype(np.complex128)
    gs[0].set_psi(psi)
    gs[0].antisymmetrize(particles)
    assert np.array_equal(gs[0].psi(), want)
    gs.close()
    ctx.close()


@pytest.mark.parametrize("shape,au", [((96, 96), 10.0), ((64, 64), 0.0), ((24, 24, 24), -50.0), ((12, 16, 8), -55.0)])
def test_cross_out_oracle_equals_filter_folds(shape, au):
    w = F.cross_out_weights(shape, au)
    for idx in itertools.product(*[range(n) for n in shape]):
        nodes = [int(F._neg_dist(np.int64(i), n)) for i, n in zip(idx, shape)]
        assert w[idx] == F.multi_ionization_filter_node(au + 2 * F.RAMP, F.RAMP, nodes, shape[0])
    # weights lie in [0, 1]; once the radius fits inside the grid, nodes with exactly two
    # coordinates beyond it keep weight 1
    assert w.min() >= 0.0 and w.max() <= 1.0
    if int(au + 2 * F.RAMP) < min(shape) // 2 - 1:
        assert w.max() == 1.0


@pytest.mark.gpu
@pytest.mark.parametrize("dim,n,au", [(2, 256, 10.0), (2, 192, 0.0), (3, 64, -50.0), (1, 256, 10.0)])
@pytest.mark.parametrize("prec", [0, 1])
def test_device_cross_out(dim, n, au, prec):
    from qsf_amd import hip
    from qsf_amd import qsf_types as T
    print(f"[case] cross_out dim={dim} n={n} au={au} prec={prec}", flush=True)
    ctx = hip.Context(0)
    precision = T.FP64 if prec == 0 else T.FP32_MIXED
    gs = hip.GridSet(ctx, T.grid_desc(dim, n, 0.3, precision=precision))
    psi = crand(n ** dim, 6)
    if precision != T.FP64:
        psi = psi.astype(np.complex64).astype(np.complex128)
    gs[0].set_psi(psi)
    gs[0].cross_out(au)
    want = F.cross_out(psi, (n,) * dim, au)
    if precision == T.FP64:
        assert np.array_equal(gs[0].psi(), want)
    else:   # the FP64 product rounded once to FP32
        assert np.array_equal(gs[0].psi(), want.astype(np.complex64).astype(np.complex128))
    gs.close()
    ctx.close()


def test_orthogonalize_oracle_projects_out():
    n, dx = 32, 0.3
    dV = dx * dx
    phi = crand(n * n, 7)
    phi /= np.sqrt(np.sum(np.abs(phi) ** 2) * dV)      # <phi|phi> dV = 1
    psi = crand(n * n, 8)
    out, s = F.orthogonalize(psi, phi, dV)
    assert abs(np.sum(np.conj(phi) * out) * dV) < 1e-12 * abs(s)
    # a state already orthogonal is left alone; phi itself is removed completely
    again, s2 = F.orthogonalize(out, phi, dV)
    assert abs(s2) < 1e-12 * abs(s) and np.allclose(again, out, rtol=0, atol=1e-13)
    gone, s3 = F.orthogonalize(2.5 * phi, phi, dV)
    assert abs(s3 - 2.5) < 1e-13 and np.max(np.abs(gone)) < 1e-13


@pytest.mark.gpu
@pytest.mark.parametrize("prec", [0, 1])
def test_device_orthogonalize_with_file(tmp_path, prec):
    from qsf_amd import hip, psi_file as PF
    from qsf_amd import qsf_types as T
    n, dx = 96, 0.3
    print(f"[case] orthogonalize prec={prec}", flush=True)
    precision = T.FP64 if prec == 0 else T.FP32_MIXED
    phi = crand(n * n, 9)
    phi /= np.sqrt(np.sum(np.abs(phi) ** 2) * dx * dx)
    psi = crand(n * n, 10)
    if precision != T.FP64:
        psi = psi.astype(np.complex64).astype(np.complex128)
    name = PF.write(tmp_path / "ground", phi, [n, n], [dx, dx])
    ctx = hip.Context(0)
    gs = hip.GridSet(ctx, T.grid_desc(2, n, dx, precision=precision))
    gs[0].set_psi(psi)
    s = PF.orthogonalize_with(gs, name)
    want, s_ref = F.orthogonalize(psi, phi, dx * dx)
    tol = 1e-13 if precision == T.FP64 else 1e-6
    assert abs(s - s_ref) <= 1e-13 * abs(s_ref)
    got = gs[0].psi()
    assert np.linalg.norm(got - want) <= tol * np.linalg.norm(want)
    gs.close()
    ctx.close()
