"""hdg_max_abs4_noalloc (the Newton norms with the library's static scratch)
against hdg_max_abs4 with a caller-provided scratch, through the C ABI: two
calls in a row, the second on smaller vectors with smaller values, so that a
partial maximum left over from the first call would show."""
import ctypes as ct

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _call(torch, L, vectors, scratch=None, lengths=None):
    arr = (ct.c_void_p * 4)(*[t.data_ptr() for t in vectors])
    lens = (ct.c_longlong * 4)(*(lengths or [t.numel() for t in vectors]))
    out = torch.full((4,), -1.0, dtype=torch.float64, device="cuda")
    stream = ct.c_void_p(torch.cuda.current_stream().cuda_stream)
    if scratch is None:
        rc = L.hdg_max_abs4_noalloc(arr, lens, ct.c_void_p(out.data_ptr()), stream)
    else:
        rc = L.hdg_max_abs4(arr, lens, ct.c_void_p(scratch.data_ptr()),
                            ct.c_void_p(out.data_ptr()), stream)
    assert rc == 0, L.hdg_last_error().decode()
    torch.cuda.synchronize()
    return out.cpu().numpy()


def test_noalloc_norms_match_scratch_entry_point():
    import torch
    from ddhdg_amd import hip_lib
    assert torch.cuda.is_available()
    print("case max_abs4_noalloc vs max_abs4, two calls", flush=True)
    L = hip_lib.lib()
    rng = np.random.default_rng(3)
    scratch = torch.empty(1024, dtype=torch.float64, device="cuda")
    for call, (lens, scale) in enumerate((((1_000_003, 300_000, 700_001, 5), 10.0),
                                          ((100, 7, 64, 1), 1e-3))):
        host = [scale * rng.standard_normal(n) for n in lens]
        vecs = [torch.as_tensor(h, device="cuda") for h in host]
        got = _call(torch, L, vecs)
        ref = _call(torch, L, vecs, scratch)
        want = np.array([np.abs(h).max() for h in host])
        print(f"  call {call}: noalloc {got} scratch {ref}", flush=True)
        assert np.array_equal(got, ref) and np.array_equal(got, want)


def test_noalloc_norms_edge_cases():
    import torch
    from ddhdg_amd import hip_lib
    L = hip_lib.lib()
    print("case max_abs4_noalloc empty vector and invalid length", flush=True)
    # a vector of length 0 (its pointer valid, as both entry points require)
    vecs = [torch.tensor([-2.0, 1.0], dtype=torch.float64, device="cuda"),
            torch.tensor([9.0], dtype=torch.float64, device="cuda"),
            torch.tensor([0.5], dtype=torch.float64, device="cuda"),
            torch.tensor([-7.0], dtype=torch.float64, device="cuda")]
    got = _call(torch, L, vecs, lengths=[2, 0, 1, 1])
    scratch = torch.empty(1024, dtype=torch.float64, device="cuda")
    ref = _call(torch, L, vecs, scratch, lengths=[2, 0, 1, 1])
    assert np.array_equal(got, [2.0, 0.0, 0.5, 7.0]) and np.array_equal(got, ref)
    # a NaN anywhere makes its norm NaN (a NaN update must not read as converged)
    vecs[2] = torch.tensor([0.5, float("nan"), 1.0], dtype=torch.float64, device="cuda")
    got = _call(torch, L, vecs, lengths=[2, 0, 3, 1])
    assert np.isnan(got[2]) and np.array_equal(got[[0, 1, 3]], [2.0, 0.0, 7.0])
    arr = (ct.c_void_p * 4)(*[t.data_ptr() for t in vecs])
    lens = (ct.c_longlong * 4)(2, -1, 1, 1)
    out = torch.empty(4, dtype=torch.float64, device="cuda")
    rc = L.hdg_max_abs4_noalloc(arr, lens, ct.c_void_p(out.data_ptr()),
                                ct.c_void_p(torch.cuda.current_stream().cuda_stream))
    assert rc < 0 and "invalid vector" in L.hdg_last_error().decode()
