"""Shared helpers of the test suite: build oracle arguments from a problem."""
import numpy as np

from oracle import oracle as O


def oracle_args(pb, **solver):
    g = O.geom(pb.dim, pb.n[: pb.dim], pb.dx)
    kw = dict(pb.solver)
    kw.update(solver)
    p = O.params(**kw)
    cf = O.coef(pb.alpha, pb.beta, pb.beta_ed, pb.gamma)
    return g, p, cf


def relerr(a, b):
    a = np.concatenate([np.ravel(x) for x in a]) if isinstance(a, (list, tuple)) else np.ravel(a)
    b = np.concatenate([np.ravel(x) for x in b]) if isinstance(b, (list, tuple)) else np.ravel(b)
    den = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (den if den > 0 else 1.0)


def bit_equal(a, b):
    if isinstance(a, (list, tuple)):
        return all(np.array_equal(x, y) for x, y in zip(a, b))
    return np.array_equal(a, b)
