"""TwoNO1 (misc.f:166) on the GPU against the oracle and the closed form."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _system(n, rank, seed):
    from gammcor_amd import synth
    chol = synth.cholesky_factors(n, rank, 0.1, seed)
    return chol, synth.ao_integrals_packed(chol), synth.orbitals(n, seed + 1).T.copy()


def _closed_form(chol, u):
    """(ab|cd) = sum_P (U L_P U^T)_ab (U L_P U^T)_cd from the factors."""
    n = u.shape[0]
    il, jl = np.tril_indices(n)
    lp = np.zeros((chol.shape[0], n, n))
    lp[:, il, jl] = chol
    lp[:, jl, il] = chol
    b = np.einsum("am,Pmn,bn->Pab", u, lp, u, optimize=True)[:, il, jl]   # (P, T)
    v = b.T @ b
    i, j = np.tril_indices(v.shape[0])
    return v[i, j]


@pytest.mark.parametrize("n,rank", [(5, 7), (12, 20), (48, 60), (67, 40)])
def test_twono1_device(cuda, n, rank):
    import torch
    from gammcor_amd import _lib
    import accas_oracle as O
    chol, ao, u = _system(n, rank, n)
    ref = O.two_no1(ao, u)
    closed = _closed_form(chol, u)
    assert np.max(np.abs(ref - closed)) < 1e-12
    d = torch.from_numpy(ao).to(cuda)
    du = torch.from_numpy(np.asfortranarray(u).ravel(order="F")).to(cuda)
    _lib.call("gmc_twono1_dev", d.data_ptr(), du.data_ptr(), n, len(ao), None)
    out = d.cpu().numpy()
    assert np.max(np.abs(out - ref)) < 1e-12


def test_twono1_host(cuda):
    from gammcor_amd import _lib
    import accas_oracle as O
    chol, ao, u = _system(20, 30, 3)
    ref = O.two_no1(ao, u)
    buf = ao.copy()
    uf = np.asfortranarray(u)
    _lib.call("gmc_twono1", buf.ctypes.data, uf.ctypes.data, 20, len(buf))
    assert np.max(np.abs(buf - ref)) < 1e-12
