"""lmm_diffusive_mass_fluxdiv_noalloc: the species-diffusion entry point whose
flux scratch is the library's file-scope device array instead of a per-call
allocation.  Called through the C ABI; its results must equal
lmm_diffusive_mass_fluxdiv's bit for bit, on two calls in a row, the second
one smaller (fewer cells and species), and a request beyond the array's
size must be refused before any device work."""
import ctypes as C

import numpy as np
import pytest

from lowmachmixtures_amd import stokes as S
from tests.helpers import bit_equal

pytestmark = pytest.mark.gpu


def _case(dim, n, ns, seed):
    shape = tuple(reversed(n))
    rng = np.random.default_rng(seed)
    return dict(
        x=rng.random((ns,) + shape), rw=rng.standard_normal((ns * ns,) + shape),
        gm=rng.standard_normal((ns * ns,) + shape), rho=rng.random((ns,) + shape),
        ft0=[rng.standard_normal((ns,) + shape) for _ in range(dim)],
    )


def _run(entry, ctx, dim, ns, c):
    ft = [ctx.field(S.FACE_X + d, ns, c["ft0"][d]) for d in range(dim)]
    div = ctx.field(S.CELL, ns)
    rt = c["rho"].sum(axis=0)
    fn = getattr(S.lib(), entry)
    rc = fn(ctx.h, ns, S._h(ctx.field(S.CELL, ns, c["rho"])), S._h(ctx.field(S.CELL, 1, rt)),
            S._h(ctx.field(S.CELL, ns, c["x"])), S._h(ctx.field(S.CELL, ns * ns, c["rw"])),
            S._h(ctx.field(S.CELL, ns * ns, c["gm"])), S._h(div), None, None, S._a3([None] * 3), S._a3(ft), 1)
    return rc, div.download(), [f.download() for f in ft]


@pytest.mark.parametrize("dim,big,small", [(3, ((64, 32, 32), 4), ((16, 8, 8), 2)),
                                           (2, ((256, 128), 5), ((32, 16), 3))])
def test_noalloc_equals_the_allocating_entry_point(dim, big, small):
    for k, (n, ns) in enumerate((big, small)):
        c = _case(dim, n, ns, 40 + k)
        with S.Context(dim, n, 0.25) as ctx:
            rc0, div0, ft0 = _run("lmm_diffusive_mass_fluxdiv", ctx, dim, ns, c)
            rc1, div1, ft1 = _run("lmm_diffusive_mass_fluxdiv_noalloc", ctx, dim, ns, c)
        assert rc0 == 0 and rc1 == 0
        assert bit_equal(div1, div0)
        assert bit_equal(ft1, ft0)


def test_noalloc_refuses_sizes_beyond_its_scratch():
    dim, n, ns = 3, (128, 64, 64), 3  # 3 * 3 * 524288 values > 4194304
    c = _case(dim, n, ns, 50)
    with S.Context(dim, n, 0.25) as ctx:
        rc, _, _ = _run("lmm_diffusive_mass_fluxdiv_noalloc", ctx, dim, ns, c)
        msg = S.lib().lmm_ctx_last_error(ctx.h).decode()
    assert rc == 1 and "scratch" in msg
