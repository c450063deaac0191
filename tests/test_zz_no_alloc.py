"""The allocation-free likelihood entry point (eobhip_fd_likelihood_static)
against eobhip_fd_likelihood with a caller buffer, through the C ABI: two calls
in a row, the second on a smaller batch, bit for bit; sizes past the static
buffer are rejected."""
import numpy as np
import pytest
import torch

from eobgpu_amd._lib import EobHipError, call, lib
from eobgpu_amd.utils import current_stream

pytestmark = pytest.mark.gpu


def _spectra(rng, n, F, dev):
    h = rng.normal(size=(n, F)) + 1j * rng.normal(size=(n, F))
    d = rng.normal(size=F) + 1j * rng.normal(size=F)
    return torch.tensor(h, device=dev), torch.tensor(d, device=dev)


def test_static_likelihood_matches_buffered(dev):
    rng = np.random.default_rng(17)
    for n, F in ((96, (1 << 15) + 1), (5, 4097)):
        ht, dt = _spectra(rng, n, F, dev)
        df = 1.0 / 64
        part = torch.empty(int(lib().eobhip_fd_likelihood_partial_doubles(n, F)), dtype=torch.float64, device=dev)
        ref = torch.empty(n, dtype=torch.float64, device=dev)
        call("eobhip_fd_likelihood", ht.data_ptr(), dt.data_ptr(), part.data_ptr(), n, F, df, 200.0, ref.data_ptr(),
             current_stream())
        out = torch.full((n,), float("nan"), dtype=torch.float64, device=dev)
        call("eobhip_fd_likelihood_static", ht.data_ptr(), dt.data_ptr(), n, F, df, 200.0, out.data_ptr(), current_stream())
        torch.cuda.synchronize()
        np.testing.assert_array_equal(out.cpu().numpy(), ref.cpu().numpy())
    cap = int(lib().eobhip_fd_likelihood_static_max_partials())
    F = (1 << 19) + 1
    n_too_many = cap // 257 + 1
    with pytest.raises(EobHipError, match="static buffer"):
        call("eobhip_fd_likelihood_static", ht.data_ptr(), dt.data_ptr(), n_too_many, F, 1.0, 200.0, out.data_ptr(),
             current_stream())
