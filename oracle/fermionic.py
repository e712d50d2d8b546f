"""Oracle for operation<AntiSymmetrize<...>> (wf/grid.h:611-723), one rank, X representation.
Test infrastructure only: the tests compare the device entry point qsf_antisymmetrize with it.

`antisymmetrize` is the vectorised restatement; `antisymmetrize_loops` transcribes the
reference's loops node by node (pure Python, tiny grids) and pins it.
"""
import numpy as np


def _shape(n, dim):
    return tuple(int(v) for v in n) if hasattr(n, "__len__") else (int(n),) * dim


def antisymmetrize(psi, n, dim, particles, dtype=np.complex128):
    """psi: complex array of prod(shape) nodes, row-major; `n` is the axis length or the shape
    (AntiSymmetrize<2_D> on a 3D grid allows n0 != n1 = n2). Returns the new array (input
    untouched), computed in `dtype` (complex64: every addition a float32 operation, in the
    order below)."""
    shape = _shape(n, dim)
    a = np.array(psi, dtype=dtype).reshape(shape)
    if dim < particles:
        return a.ravel()
    if particles == 2 and dim == 2:                     # :700-716 (psi_total = original)
        return (a - a.T).ravel()
    if particles == 2 and dim == 3:                     # :686-699, last two axes
        n = shape[1]
        out = a - a.transpose(0, 2, 1)
        j, k = np.indices((n, n))
        lower = k < j
        # (i,k,j) = -(i,j,k) for k < j: copy the negated lower triangle over the upper one
        up = out.transpose(0, 2, 1).copy()
        out = np.where(lower[None], out, -up)
        out[:, np.arange(n), np.arange(n)] = 0
        return out.ravel()
    # particles == 3, dim == 3 (:646-677)
    n = shape[0]
    t = a.transpose(1, 2, 0) + a.transpose(2, 0, 1)     # (j,k,i) + (k,i,j) seen from (i,j,k)
    t = t - a.transpose(1, 0, 2)
    t = t - a.transpose(0, 2, 1)
    t = t - a.transpose(2, 1, 0)
    v = a + t
    i, j, k = np.indices((n, n, n))
    out = np.zeros_like(a)
    s = (i > j) & (j > k)
    I, J, K = i[s], j[s], k[s]
    val = v[s]
    for (p, q, r), sign in (((I, J, K), 1), ((J, K, I), 1), ((K, I, J), 1),
                            ((J, I, K), -1), ((I, K, J), -1), ((K, J, I), -1)):
        out[p, q, r] = sign * val
    return out.ravel()


def antisymmetrize_loops(psi, n, dim, particles):
    """The reference's loops, node by node (tiny grids only)."""
    p = [complex(x) for x in np.asarray(psi).ravel()]
    shape = _shape(n, dim)
    n = shape[-1]
    off = (lambda i, j: i * n + j) if dim == 2 else (lambda i, j, k: (i * shape[1] + j) * n + k)
    if dim < particles:
        return np.array(p)
    if particles == 2 and dim == 2:
        total = list(p)
        for i in range(n):
            for j in range(n):
                p[off(i, j)] -= total[off(j, i)]
    elif particles == 2:
        for i in range(shape[0]):
            for j in range(shape[1]):
                for k in range(j):
                    p[off(i, j, k)] -= p[off(i, k, j)]
                    p[off(i, k, j)] = -p[off(i, j, k)]
                p[off(i, j, j)] = 0j
    else:
        for i in range(n):
            for j in range(i):
                for k in range(j):
                    r, r2, r3 = off(i, j, k), off(j, k, i), off(k, i, j)
                    s1, s2, s3 = off(j, i, k), off(i, k, j), off(k, j, i)
                    p[r] += p[r2] + p[r3] - p[s1] - p[s2] - p[s3]
                    p[r2] = p[r]
                    p[r3] = p[r]
                    p[s1] = -p[r]
                    p[s2] = -p[r]
                    p[s3] = -p[r]
            for j in range(n):
                p[off(i, j, j)] = 0j
                p[off(j, i, j)] = 0j
                p[off(j, j, i)] = 0j
    return np.array(p)


# ---- croossOut / multiIonizationFilter<2_D, 2_D> (wf/grid.h:862-907) ----
RAMP = 30


def _neg_dist(i, n):
    return np.where(i < n // 2, -i, i - n + 1)


def cross_out_weights(shape, au_distance=50.0, ramp=RAMP):
    """The factor croossOut multiplies each node by (vectorised restatement)."""
    dim = len(shape)
    from_center = int(au_distance + 2 * ramp)          # (ind) of a double argument
    from_boundary = shape[0] // 2 - from_center          # n2[0] - nodesFromCenter
    idx = np.indices(shape)
    nodes = np.stack([_neg_dist(idx[a], shape[a]) + from_boundary for a in range(dim)])
    count = (nodes > 0).sum(axis=0)
    take_max = (count == 0) & (dim != 2)
    mv = np.where(take_max, nodes.max(axis=0), nodes.min(axis=0))
    mv = np.minimum(np.abs(mv), ramp)
    table = np.array([0.5 * (1. + np.cos(np.pi * k / ramp)) for k in range(ramp + 1)])
    return np.where(count == 2, 1.0, table[mv])


def cross_out(psi, shape, au_distance=50.0):
    return (np.asarray(psi, dtype=np.complex128).reshape(shape) * cross_out_weights(shape, au_distance)).ravel()


def multi_ionization_filter_node(nodes_from_center, ramp, nodes, n_axis0, frm=2, to=2):
    """multiIonizationFilter<from, to>(nodesFromCenter, ramp, nodes...) for one node, with the
    reference's folds written out (tiny grids)."""
    import math
    dim = len(nodes)
    from_boundary = n_axis0 // 2 - int(nodes_from_center)
    nodes = [x + from_boundary for x in nodes]
    count = sum(1 for x in nodes if x > 0)
    if frm <= count <= to:
        return 1.0
    min_val = float(nodes[-1])                          # (nodes, ...): the last one
    for x in nodes:
        if (x > min_val) if (count == 0 and frm != dim) else (x < min_val):
            min_val = x
    min_val = ramp if abs(min_val) > ramp else abs(min_val)
    return 0.5 * (1. + math.cos(math.pi * min_val / ramp))


def orthogonalize(psi, phi, dV):
    """orthogonalizeWith (wf/grid.h:909-946): s = sum conj(phi) psi * dV, psi - s phi."""
    psi = np.asarray(psi, dtype=np.complex128).ravel()
    phi = np.asarray(phi, dtype=np.complex128).ravel()
    s = np.sum(np.conj(phi) * psi) * dV
    return psi - s * phi, s
