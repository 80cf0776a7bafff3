"""st_solve_norms_static, the allocation-free variant of st_solve_norms, through
the C ABI: the same norms as st_solve_norms on two calls in a row, the second on
a smaller tank (fewer workgroups in its reduction, so the fixed array's earlier
contents are left over where the second call does not write)."""
import pytest

from splashtank_amd import capi, scenes
from splashtank_amd.capi import FlipSimC
from tests.helpers import opts_for, upload_scene

pytestmark = pytest.mark.gpu


def _stepped(sc, n):
    g = FlipSimC(opts_for(sc))
    upload_scene(g, sc)
    prm = capi.make_params(**sc.params)
    for _ in range(n):
        g.step(prm)
    return g


def test_static_norms_equal_allocating_norms():
    big = scenes.dam_break(d=1.0 / 96, iters=20)   # 48^3 cells
    small = scenes.dam_break(d=1.0 / 32, iters=20)  # 16^3 cells
    gb = _stepped(big, 3)
    gs = _stepped(small, 3)
    try:
        print(f"[no-alloc] {big.n} and {small.n} particles, 3 substeps each", flush=True)
        for g in (gb, gs):  # the large tank first, then the smaller one on the same array
            want = g.solve_norms()
            got = g.solve_norms_static()
            print(f"[no-alloc] allocating {want}\n[no-alloc] static     {got}", flush=True)
            assert want["fluid"] > 100
            assert got == want  # same kernels, same fixed reduction order: equal bits
        with pytest.raises(capi.StError):
            FlipSimC(opts_for(small)).solve_norms_static()  # nothing solved yet
    finally:
        gb.close()
        gs.close()
