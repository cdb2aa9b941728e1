"""Dense-matrix assembly of the cell-centred Poisson operator, used to pin the
oracle (and through it the GPU path) by brute force on tiny grids.

The operator is built from its mathematical definition: on each axis the
metric-weighted second difference  c2*(x+ - 2x + x-) + c1*(x+ - x-)  with
c2 = xi_x^2 / h^2 and c1 = xi_xx / (2h)  (lib/poisson/poisson_d3.cc:81-87),
and the ghost cell replaced by the boundary rule of that axis:
    periodic   x(-1) = x(n-1)
    Neumann    x(-1) = x(0)            (poisson_d3.cc:892-895)
    Dirichlet  x(-1) = -x(0)           (poisson_d3.cc:896-900)
    Dirichlet with wall value w: x(-1) = 2w - x(0) -> moves 2 w c to the RHS.
"""
import numpy as np


def axis_coeffs(xix2, xixx, h):
    """c2, c1 over the core (inputs over [-1, n])."""
    return xix2[1:-1] / h ** 2, xixx[1:-1] * (0.5 / h)


def assemble(shape, coeffs, kinds):
    """shape: core extents per active axis; coeffs: list of (c2, c1) per axis;
    kinds: per axis 'P', 'N' or 'D', or a (lo, hi) pair such as 'ND'.  Returns
    the dense matrix A (row-major flattening, last axis fastest)."""
    shape = tuple(shape)
    N = int(np.prod(shape))
    A = np.zeros((N, N))
    idx = np.arange(N).reshape(shape)
    for ax, ((c2, c1), kind) in enumerate(zip(coeffs, kinds)):
        n = shape[ax]
        for flat in range(N):
            pos = list(np.unravel_index(flat, shape))
            i = pos[ax]
            a2, a1 = c2[i], c1[i]
            # centre
            A[flat, flat] += -2.0 * a2
            for off, w in ((+1, a2 + a1), (-1, a2 - a1)):
                j = i + off
                if 0 <= j < n:
                    q = list(pos)
                    q[ax] = j
                    A[flat, idx[tuple(q)]] += w
                else:
                    kk = kind if len(kind) == 1 else kind[0 if off < 0 else 1]
                    if kk == "P":
                        q = list(pos)
                        q[ax] = j % n
                        A[flat, idx[tuple(q)]] += w
                    elif kk == "N":
                        A[flat, flat] += w
                    elif kk == "D":
                        A[flat, flat] -= w
                    else:
                        raise ValueError(kind)
    return A


def wall_rhs_shift(shape, coeffs, kinds, walls):
    """RHS contribution of inhomogeneous Dirichlet ghosts 2w - x(0):
    A x + sum(w_face * 2 * coef) = b  ->  move to RHS.  walls[ax] = (lo, hi)
    arrays of the wall value on the face (shape of the other axes)."""
    shape = tuple(shape)
    shift = np.zeros(shape)
    for ax, ((c2, c1), kind) in enumerate(zip(coeffs, kinds)):
        if walls[ax] is None:
            continue
        lo, hi = walls[ax]
        n = shape[ax]
        sl_lo = [slice(None)] * len(shape)
        sl_lo[ax] = 0
        sl_hi = [slice(None)] * len(shape)
        sl_hi[ax] = n - 1
        shift[tuple(sl_lo)] += 2.0 * (c2[0] - c1[0]) * lo
        shift[tuple(sl_hi)] += 2.0 * (c2[n - 1] + c1[n - 1]) * hi
    return shift
