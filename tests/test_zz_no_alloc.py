"""The no-allocation FP64 entry point (pcrp_solve_f64_noalloc) against
pcrp_solve_f64 with a caller workspace, through the C ABI: two calls in a row
(the second smaller, reusing the fixed scratch the first left behind), both
bit-identical to the workspace entry point and to the oracle; an LDS-size
call (no scratch); a request beyond the scratch is refused with
PCRP_ERR_SCRATCH."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

C, I = 0, 1


def _p(t):
    return ctypes.c_void_p(t.data_ptr())


def _case(gpu, oracle_lib, lib, kind, seed, n, m, layout):
    import torch
    print("[case] noalloc", kind, seed, n, m, layout, flush=True)
    arrs = oracle_lib.gen_batch(kind, seed, n, m, layout)
    dev = [torch.from_numpy(a.copy()).cuda() for a in arrs]
    y_ws = dev[5].clone()
    nb = int(lib.pcrp_workspace_size(n, m, 0))
    ws = torch.empty(max(nb, 1), dtype=torch.uint8, device="cuda")
    st = torch.cuda.current_stream()
    s = ctypes.c_void_p(st.cuda_stream)
    rc = lib.pcrp_solve_f64(*[_p(t) for t in dev[:5]], _p(y_ws), n, m, layout, _p(ws), nb, s)
    assert rc == 0, rc
    y_na = dev[5].clone()
    rc = lib.pcrp_solve_f64_noalloc(*[_p(t) for t in dev[:5]], _p(y_na), n, m, layout, s)
    assert rc == 0, rc
    torch.cuda.synchronize()
    x_ws, x_na = y_ws.cpu().numpy(), y_na.cpu().numpy()
    assert np.array_equal(x_na, x_ws), (n, m, layout)
    assert np.array_equal(x_na, oracle_lib.pcr_batch(*arrs, n, m, layout)), (n, m, layout)
    return nb


def test_noalloc_matches_workspace_entry_point(gpu, oracle_lib):
    lib = gpu.load()
    cap = 64 << 20
    nb1 = _case(gpu, oracle_lib, lib, 1, 11, 16384, 40, I)      # split path, 63 MB of scratch
    assert 0 < nb1 <= cap
    nb2 = _case(gpu, oracle_lib, lib, 0, 12, 3000, 7, C)        # smaller, over the old contents
    assert 0 < nb2 < nb1
    assert _case(gpu, oracle_lib, lib, 0, 13, 1000, 33, I) == 0   # LDS size: no scratch
    # beyond the scratch: refused before any launch
    fake = [ctypes.c_void_p(0x1000 * (k + 1)) for k in range(6)]
    assert int(lib.pcrp_workspace_size(1 << 20, 32, 0)) > cap
    assert lib.pcrp_solve_f64_noalloc(*fake, 1 << 20, 32, I, None) == 8
    assert lib.pcrp_solve_f64_noalloc(*fake, 0, 32, I, None) == 1
    assert lib.pcrp_solve_f64_noalloc(*fake[:5], None, 100, 32, I, None) == 2
    assert lib.pcrp_solve_f64_noalloc(*fake, 5000, 32, 9, None) == 3
