"""cup_diagnostics_noalloc (the diagnostics over caller-owned device fields, partial sums in
the library's static device array) against cup_diagnostics on the same fields: two calls in
a row on one GPU, the second on a smaller grid in the other precision, so the second reads a
scratch array that still holds the first call's partial sums."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _random_fields(G, rng):
    shape = G.shape
    G.set("chi", rng.uniform(0.0, 1.0, shape))
    for n in ("u", "v", "w", "p"):
        G.set(n, rng.standard_normal(shape))


def _noalloc(G, nu):
    import torch
    from cubismup_3d_amd import _lib as L
    out = torch.full((20,), float("nan"), dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    d = L.GridDesc((C.c_int * 3)(*G.bpd), float(G.extent.max()), G.precision, 1, 0)
    ptrs = (C.c_void_p * 5)(*[G.device_ptr(n) for n in ("chi", "u", "v", "w", "p")])
    G.synchronize()
    rc = L.lib().cup_diagnostics_noalloc(C.byref(d), ptrs, float(nu), C.c_void_p(G.stream()),
                                         C.c_void_p(out.data_ptr()))
    assert rc == 0, L.lib().cup_last_error(None)
    G.synchronize()
    return out.cpu().numpy()


@pytest.mark.parametrize("order", ["large_then_small"])
def test_noalloc_diagnostics_match_the_allocating_entry_point(cup, order):
    rng = np.random.default_rng(7)
    for bpd, prec, nu in (((8, 8, 8), 8, 0.01), ((4, 2, 2), 4, 0.003)):
        print(f"noalloc diagnostics: bpd {bpd}, precision {prec}", flush=True)
        G = cup.Grid(bpd, precision=prec, decomposition="fused")
        _random_fields(G, rng)
        ref = G.diagnostics(nu)
        got = _noalloc(G, nu)
        print(f"  max |noalloc - cup_diagnostics| = {np.abs(got - ref).max():.3e}", flush=True)
        assert np.all(np.isfinite(got))
        # the same kernels over the same partial-sum layout: identical bits
        assert np.array_equal(got, ref)
        G.close()
