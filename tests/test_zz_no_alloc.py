"""tc_thomas_batched_static (the allocation-free batched Thomas sweep) through the C ABI against
tc_thomas_batched with a caller workspace: bit-identical results on two calls in a row, the
second smaller than the first, and a clear error above the static scratch's capacity."""
import ctypes

import pytest
import torch

pytestmark = pytest.mark.gpu

TC_ERR_SCRATCH = -1006
MAX_F64 = 1 << 22


@pytest.fixture(scope="module")
def lib():
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    from torchannel_amd import _lib as L
    return L.load()


def systems(n, batch, dtype, cplx, shared, seed):
    g = torch.Generator().manual_seed(seed)
    shape = (n,) if shared else (batch, n)
    a = torch.rand(shape, generator=g, dtype=torch.float64) - 0.5
    c = torch.rand(shape, generator=g, dtype=torch.float64) - 0.5
    b = 2.0 + torch.rand(shape, generator=g, dtype=torch.float64)          # diagonally dominant
    d = torch.rand((batch, n, 2) if cplx else (batch, n), generator=g, dtype=torch.float64) - 0.5
    return [t.to(dtype).cuda() for t in (a, b, c, d)]


def solve(lib, name, n, batch, dtype, cplx, shared, a, b, c, d):
    x = torch.empty_like(d)
    code = 0 if dtype == torch.float64 else 1
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    args = [n, batch, code, int(cplx), ctypes.c_void_p(a.data_ptr()), ctypes.c_void_p(b.data_ptr()),
            ctypes.c_void_p(c.data_ptr()), 0 if shared else n, ctypes.c_void_p(d.data_ptr()),
            ctypes.c_void_p(x.data_ptr())]
    if name == "tc_thomas_batched":
        work = torch.empty(batch * n, dtype=dtype, device="cuda")
        args.append(ctypes.c_void_p(work.data_ptr()))
    rc = getattr(lib, name)(*args, stream)
    return rc, x


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_static_scratch_matches_workspace(lib, dtype):
    # the 192x192x128 Poisson batch (192 x 97 complex systems of 128 rows), then a smaller real one
    cases = [(128, 192 * 97, True, False, 1), (48, 1000, False, True, 2)]
    if dtype == torch.float32:
        cases.insert(1, (256, 20000, False, False, 3))     # above the float64 limit, within float32's
    for n, batch, cplx, shared, seed in cases:
        print(f"[no-alloc] n={n} batch={batch} complex={cplx} shared={shared} {dtype}", flush=True)
        a, b, c, d = systems(n, batch, dtype, cplx, shared, seed)
        rc0, want = solve(lib, "tc_thomas_batched", n, batch, dtype, cplx, shared, a, b, c, d)
        rc1, got = solve(lib, "tc_thomas_batched_static", n, batch, dtype, cplx, shared, a, b, c, d)
        torch.cuda.synchronize()
        assert rc0 == 0 and rc1 == 0
        assert torch.equal(got, want)
        # the sweep solves the systems: residual of the first and last rows
        x = torch.view_as_complex(got.double()) if cplx else got.double()
        dd = torch.view_as_complex(d.double()) if cplx else d.double()
        aa, bb, cc = (t.double().expand(batch, n) for t in (a, b, c))
        r0 = bb[:, 0] * x[:, 0] + cc[:, 0] * x[:, 1] - dd[:, 0]
        r1 = aa[:, -1] * x[:, -2] + bb[:, -1] * x[:, -1] - dd[:, -1]
        tol = 1e-12 if dtype == torch.float64 else 1e-5
        assert float(r0.abs().max()) < tol and float(r1.abs().max()) < tol


def test_static_scratch_rejects_oversize(lib):
    n, batch = 256, MAX_F64 // 256 + 1                    # one system past the float64 capacity
    a, b, c, d = (torch.zeros(1, dtype=torch.float64, device="cuda") for _ in range(4))
    x = torch.zeros(1, dtype=torch.float64, device="cuda")
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    p = [ctypes.c_void_p(t.data_ptr()) for t in (a, b, c)]
    rc = lib.tc_thomas_batched_static(n, batch, 0, 0, *p, 0, ctypes.c_void_p(d.data_ptr()),
                                      ctypes.c_void_p(x.data_ptr()), stream)
    assert rc == TC_ERR_SCRATCH
    lib.tc_strerror.restype = ctypes.c_char_p
    assert b"static scratch" in lib.tc_strerror(rc)
