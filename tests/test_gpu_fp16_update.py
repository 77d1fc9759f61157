"""qd_fp16_update (the FP16 trailing update of the mixed factorisation as an
operator) against a plain PyTorch FP32 reference of the same operation:
C -= colscale[j/64] * A16 B16^T, FP16 operands exact in FP32, FP32 sums,
held to the elementwise error bound of FP32 summation against an FP64 reference.
The K = 1024 cases cover ragged M and N edges (N not a multiple of 4), row
pitches wider than N and launches of many tiles; the last three are launches of
at least 1,024 tiles of 256 x 256, which take the wide-tile kernel (edge tiles,
a row pitch that is not a multiple of 4, K = 2048), with the reference product
formed on the GPU."""
import ctypes as C

import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("M,N,K,pad", [(300, 200, 64, 0), (513, 129, 96, 7), (256, 128, 32, 0),
                                       (1000, 700, 1024, 64), (2000, 1500, 1024, 0), (1000, 701, 1024, 3),
                                       (4100, 2050, 1024, 62), (9000, 8300, 1024, 64), (8200, 8450, 2048, 3),
                                       (16384, 4096, 1024, 0)])
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
    dev = torch.device("cuda:0")
    dA, dB, dS, dC = A.to(dev), B.to(dev), sc.to(dev), C0.to(dev)
    rc = lib.qd_fp16_update(h, M, N, K, C.c_void_p(dA.data_ptr()), C.c_void_p(dB.data_ptr()),
                            C.c_void_p(dS.data_ptr()), C.c_void_p(dC.data_ptr()), N + pad)
    assert rc == 0, lib.qd_last_error(h)
    assert lib.qd_synchronize(h) == 0
    # FP64 reference (the FP16 products are exact in it) and the elementwise
    # bound of an FP32 evaluation in any summation order: (K + 2) u (s |A||B|^T
    # + |C|), u = 2^-24.  Large cases are formed on the GPU.
    on = dev if M * N * K > 10**10 else torch.device("cpu")
    A64, B64, s64 = A.to(on).double(), B.to(on).double(), sc.to(on).double().repeat_interleave(64)[:N][None, :]
    C64 = C0.to(on).double()[:, :N]
    ref = C64 - (A64 @ B64.T) * s64
    bound = (K + 2) * 2.0**-24 * ((A64.abs() @ B64.abs().T) * s64 + C64.abs())
    out = dC.to(on).double()
    excess = float(((out[:, :N] - ref).abs() - bound).max())
    err = float((out[:, :N] - ref).abs().max() / ref.abs().max())
    print(f"[fp16_update] max rel error vs FP64 {err:.3e}", flush=True)
    assert excess <= 0.0, (excess, err)
    out = out.cpu()
    # the padding columns are untouched
    assert torch.equal(out[:, N:].float(), C0[:, N:])
    lib.qd_destroy(h)
