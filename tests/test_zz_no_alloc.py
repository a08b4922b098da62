"""The no-allocation TVME primal step against the existing two-call path (primal + reduction).

tr_tvme_primal_noalloc keeps its stopping-sum partials in a library-owned device array; its
results must be bit-identical to tr_tvme_primal + tr_reduce_partials with a caller-allocated
partial buffer, on two calls in a row (the second with a smaller volume, so stale partials of the
first call would show), and a volume beyond the array's limit is refused before any launch."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _rand(torch, rng, dims):
    return torch.from_numpy((rng.standard_normal(dims) + 1j * rng.standard_normal(dims)).astype(np.complex64)).cuda()


def test_noalloc_primal_matches_two_call_path():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from toporecon_hip import _lib
    lib = _lib.load()
    s = _lib.stream_handle(0)
    rng = np.random.default_rng(17)
    for dims in ((3, 4, 10, 12, 14), (2, 2, 6, 8, 10)):
        B, E, Z, Y, X = dims
        print("no-alloc primal", dims, flush=True)
        u, ubar, grad, pm, px, py, pz = (_rand(torch, rng, dims) for _ in range(7))
        u2, ubar2 = u.clone(), ubar.clone()
        n = int(lib.tr_partial_count(B, Z, Y, X))
        partial = torch.empty((n, 2), dtype=torch.float64, device="cuda")
        ref = torch.empty(2, dtype=torch.float64, device="cuda")
        _lib.call("tr_tvme_primal", B, E, Z, Y, X, u.data_ptr(), ubar.data_ptr(), grad.data_ptr(), pm.data_ptr(),
                  px.data_ptr(), py.data_ptr(), pz.data_ptr(), 1 / 6, partial.data_ptr(), s)
        _lib.call("tr_reduce_partials", partial.data_ptr(), n, ref.data_ptr(), s)
        got = torch.full((2,), float("nan"), dtype=torch.float64, device="cuda")
        _lib.call("tr_tvme_primal_noalloc", B, E, Z, Y, X, u2.data_ptr(), ubar2.data_ptr(), grad.data_ptr(),
                  pm.data_ptr(), px.data_ptr(), py.data_ptr(), pz.data_ptr(), 1 / 6, got.data_ptr(), s)
        torch.cuda.synchronize()
        assert torch.equal(torch.view_as_real(u2), torch.view_as_real(u))
        assert torch.equal(torch.view_as_real(ubar2), torch.view_as_real(ubar))
        assert torch.equal(got, ref), (got.tolist(), ref.tolist())
        assert float(ref[1]) > 0.0


def test_noalloc_primal_refuses_volumes_beyond_its_array():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from toporecon_hip import _lib
    rc = _lib.load().tr_tvme_primal_noalloc(9, 1, 128, 256, 256, None, None, None, None, None, None, None,
                                            1 / 6, None, _lib.stream_handle(0))
    assert rc == -1
    assert "no-allocation limit" in _lib.last_error()
