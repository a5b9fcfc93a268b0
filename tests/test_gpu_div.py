"""compute_div (div_and_grad.f90:16-455, periodic) through the C ABI against
the oracle: the divergence of a face field, written or added to a cell field
(increment).  Each thread forms an x-pair of cells with 2-wide loads; the
last pair of a row takes its x neighbour across the periodic wrap."""
import numpy as np
import pytest

from lowmachmixtures_amd import stokes as S
from oracle import oracle as O
from tests.helpers import bit_equal, relerr

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("dim,n", [(2, (16, 8)), (2, (2, 4)), (3, (8, 4, 4)), (3, (64, 32, 16)), (3, (2, 2, 2))])
@pytest.mark.parametrize("increment", [False, True])
def test_compute_div_bit_exact(dim, n, increment):
    shape = tuple(reversed(n))
    rng = np.random.default_rng(41)
    u = [rng.standard_normal(shape) for _ in range(dim)]
    d0 = rng.standard_normal(shape)
    dx = 0.37
    want = O.compute_div(O.geom(dim, n, dx), u, div=d0 if increment else None, increment=increment)
    with S.Context(dim, n, dx) as ctx:
        uf = ctx.faces(u)
        div = ctx.field(S.CELL, 1, d0)
        S.compute_div(ctx, uf, div, increment=increment)
        got = div.download()
    assert bit_equal(got, want), relerr(got, want)
