"""GPU: batched spectral transforms (kad_coef_3d / kad_coef_i_3d / kad_coef_1d /
kad_coef_i_1d) against the oracle's restatement of Base_spectral::coef / coef_i.

Inputs follow the line-transform workload: coefficients U(-1,1) 2^-k, values
their inverse transform.  Tolerance: 1e-13 absolute on O(1) data (FP64 roundoff
of O(N) sums; the oracle uses FFTs, the GPU dense line matrices).
"""
import numpy as np
import pytest
import torch

from oracle import spectral as S

pytestmark = pytest.mark.gpu


def decaying(rng, shape):
    k = np.arange(shape[-1])
    return rng.uniform(-1, 1, shape) * 2.0 ** (-k)


@pytest.mark.parametrize("base,n", [(S.CHEB, 9), (S.CHEB, 13), (S.CHEB, 17), (S.CHEB, 25), (S.CHEB, 33),
                                    (S.COS_EVEN, 13), (S.SIN_ODD, 13), (S.COS_ODD, 9), (S.SIN_EVEN, 17),
                                    (S.COSSIN, 4), (S.COSSIN, 8), (S.COSSIN, 12), (S.COSSIN, 16),
                                    (S.COSSIN, 24), (S.COSSIN, 32)])
def test_lines(ctx, base, n):
    print(f"lines base={base} n={n}", flush=True)
    rng = np.random.default_rng(base * 100 + n)
    L = 1 << 14
    nc = n + 2 if base == S.COSSIN else n
    c = decaying(rng, (L, nc))
    if base == S.COSSIN:
        c[:, 1] = 0
        c[:, -1] = 0
    elif base in (S.COS_ODD, S.SIN_ODD):
        c[:, -1] = 0
    elif base == S.SIN_EVEN:
        c[:, 0] = 0
        c[:, -1] = 0
    sample = rng.choice(L, 256, replace=False)
    vals_ref = np.stack([S.coef_i_1d(base, c[i]) for i in sample])
    dc = torch.tensor(c, device="cuda")
    v = ctx.coef_i_1d(base, dc, n)
    np.testing.assert_allclose(v.cpu().numpy()[sample], vals_ref, atol=1e-13)
    c2 = ctx.coef_1d(base, v)
    cref = np.stack([S.coef_1d(base, vals_ref[q]) for q in range(len(sample))])
    np.testing.assert_allclose(c2.cpu().numpy()[sample], cref, atol=1e-13)
    np.testing.assert_allclose(c2.cpu().numpy(), c, atol=1e-13)


@pytest.mark.parametrize("shape", [(9, 7, 4), (17, 13, 12), (13, 9, 8), (33, 13, 16)])
@pytest.mark.parametrize("anti", [False, True])
def test_domain_3d(ctx, shape, anti):
    nr, nt, npt = shape
    print(f"3d {shape} anti={anti}", flush=True)
    rng = np.random.default_rng(nr * nt + npt)
    br, bt, bp = S.std_base(nt, npt, anti)
    nb = 6
    confs = rng.uniform(-1, 1, (nb, npt, nt, nr))
    # project to the range of coef_i (valid data for the basis)
    cfs = np.stack([S.coef_3d(confs[b], br, bt, bp) for b in range(nb)])
    vals = np.stack([S.coef_i_3d(cfs[b], br, bt, bp) for b in range(nb)])
    dv = torch.tensor(vals, device="cuda")
    got = ctx.coef_3d(dv, br, bt, bp).cpu().numpy()
    np.testing.assert_allclose(got, cfs, atol=1e-12)
    back = ctx.coef_i_3d(torch.tensor(cfs, device="cuda"), br, bt, bp).cpu().numpy()
    np.testing.assert_allclose(back, vals, atol=1e-12)
