"""The allocation-free MATINV (sevp_matinv_dev_noalloc) against the existing
MATINV entry point (sevp_matinv), bitwise, on successive calls that reuse
its fixed device scratch: a blocked-path order, a smaller blocked one, then
an order the LDS kernel takes."""
import numpy as np
import pytest

from musoc_solver_amd import sevp as S
from tests.conftest import fields_equal

pytestmark = pytest.mark.gpu


def test_matinv_dev_noalloc_matches_matinv(device):
    import torch
    rng = np.random.default_rng(77)
    for n in (1440, 300, 48):
        print(f"[noalloc] n={n}", flush=True)
        b = np.asfortranarray(rng.standard_normal((n, n)) + 0.5 * n * np.eye(n))
        ref = b.copy(order="F")
        S.matinv(ref)
        # a column-major n x n matrix is the row-major transpose
        bd = torch.from_numpy(np.ascontiguousarray(b.T)).to(device)
        zp = torch.full((1,), -1, dtype=torch.int32, device=device)
        S.matinv_dev_noalloc(bd, n, zp)
        torch.cuda.synchronize()
        got = np.asfortranarray(bd.cpu().numpy().T)
        assert int(zp.item()) == 0
        assert fields_equal(got, ref), n
    with pytest.raises(S.SevpError, match="4096"):
        S.matinv_dev_noalloc(bd, 4097)
