"""qd_fp16_update (the FP16 trailing update of the mixed factorisation as an
operator) against a plain PyTorch FP32 reference of the same operation:
C -= colscale[j/64] * A16 B16^T, FP16 operands exact in FP32, FP32 sums.
The K = 1024 cases cover ragged M and N edges (N not a multiple of 4), row
pitches wider than N and launches of many tiles."""
import ctypes as C

import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("M,N,K,pad", [(300, 200, 64, 0), (513, 129, 96, 7), (256, 128, 32, 0),
                                       (1000, 700, 1024, 64), (2000, 1500, 1024, 0), (1000, 701, 1024, 3),
                                       (4100, 2050, 1024, 62)])
def test_fp16_update_matches_torch(M, N, K, pad):
    import torch
    from bem3d_elastostatic_amd import _lib
    lib = _lib.load()
    h = C.c_void_p()
    assert lib.qd_create(C.byref(h), 0) == 0
    print(f"[fp16_update] M={M} N={N} K={K} ldc={N + pad}", flush=True)
    g = torch.Generator().manual_seed(M + N + K)
    A = torch.randn(M, K, generator=g).half()
    B = torch.randn(N, K, generator=g).half()
    sc = torch.rand((N + 63) // 64, generator=g) + 0.5
    C0 = torch.randn(M, N + pad, generator=g)
    ref = C0.clone()
    ref[:, :N] -= (A.float() @ B.float().T) * sc.repeat_interleave(64)[:N][None, :]
    dev = torch.device("cuda:0")
    dA, dB, dS, dC = A.to(dev), B.to(dev), sc.to(dev), C0.to(dev)
    rc = lib.qd_fp16_update(h, M, N, K, C.c_void_p(dA.data_ptr()), C.c_void_p(dB.data_ptr()),
                            C.c_void_p(dS.data_ptr()), C.c_void_p(dC.data_ptr()), N + pad)
    assert rc == 0, lib.qd_last_error(h)
    assert lib.qd_synchronize(h) == 0
    out = dC.cpu()
    err = float((out - ref).abs().max() / ref.abs().max())
    assert err < 2e-6, err
    # the padding columns are untouched
    assert torch.equal(out[:, N:], C0[:, N:])
    lib.qd_destroy(h)
