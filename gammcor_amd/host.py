"""Host-side mirror of the reference's AC-CAS interface over the C ABI.

Names follow the reference subroutines so a caller reads the same flow:
ACCAS pair selection (accas.f:30-99), TwoNO1 (misc.f:166), ACECORR
(acfd.f:2), GauLeg (misc.f:828).  Arrays that the GPU consumes are torch
tensors on ``cuda:0``; torch only provides device memory and streams.
"""
from __future__ import annotations

import json
import math
import os

import numpy as np

from . import _lib

GMC_FP64, GMC_FP32, GMC_MIXED, GMC_MIXED2 = 0, 1, 2, 3


def GauLeg(x1, x2, n):
    """Gauss-Legendre nodes and weights on [x1, x2] (misc.f:828-865): Newton
    iteration on P_n from the Chebyshev-like initial guess."""
    x = np.zeros(n)
    w = np.zeros(n)
    xm, xl = 0.5 * (x2 + x1), 0.5 * (x2 - x1)
    for i in range(1, (n + 1) // 2 + 1):
        z = math.cos(math.pi * (i - 0.25) / (n + 0.5))
        for _ in range(100):
            p1, p2 = 1.0, 0.0
            for j in range(1, n + 1):
                p1, p2 = ((2 * j - 1) * z * p1 - (j - 1) * p2) / j, p1
            pp = n * (z * p1 - p2) / (z * z - 1.0)
            z1, z = z, z - p1 / pp
            if abs(z - z1) <= 3.0e-14:
                break
        x[i - 1], x[n - i] = xm - xl * z, xm + xl * z
        w[i - 1] = w[n - i] = 2.0 * xl / ((1.0 - z * z) * pp * pp)
    return x, w


def ACCAS_pairs(occ, nele, thr_sel_act=1e-8, thr_qvirt=1e-7, iflcore=1):
    """IndN(2, NDimX) of the ERPA pair space (accas.f:30-99), 1-based, p > q."""
    occ = np.asarray(occ, dtype=np.float64)
    nb = len(occ)
    aux = np.where(np.arange(nb) < nele, 0, 2)
    aux[(occ < 1.0) & (occ != 0.0)] = 1
    out = []
    for i in range(nb):
        for j in range(i):
            s = aux[i] + aux[j]
            if s in (0, 4):
                continue
            if aux[i] == 1 and aux[j] == 1 and abs(occ[i] - occ[j]) / occ[i] < thr_sel_act:
                continue
            if iflcore == 1 or (occ[i] != 1.0 and occ[j] != 1.0):
                if abs(occ[i] + occ[j] - 2.0) > 1e-10 and abs(occ[i] + occ[j]) > thr_qvirt:
                    out.append((i + 1, j + 1))
    return np.array(out, dtype=np.int32).reshape(-1, 2)


def geminals(nbasis, ninact, nact):
    """IGem labels: 1 inactive, 2 active, 3 virtual (initia.f:1202-1216)."""
    g = np.full(nbasis, 3, dtype=np.int32)
    if ninact == 0:
        g[:nact] = 1
        g[nact:] = 2
    else:
        g[:ninact] = 1
        g[ninact:ninact + nact] = 2
    return g


def pack_rdm2act(gamma):
    """RDM2Act(NAddrRDM(i,j,k,l)) = Gamma(i,j,k,l) (interpa.f:4529-4541)."""
    na = gamma.shape[0]
    out = np.zeros(na * na * (na * na + 1) // 2)
    for i in range(na):
        for j in range(na):
            ij = i * na + j + 1
            for k in range(na):
                for l in range(na):
                    kl = k * na + l + 1
                    if ij >= kl:
                        out[ij * (ij - 1) // 2 + kl - 1] = gamma[i, j, k, l]
    return out


def TwoNO1(tno, ure, nbasis):
    """In-place AO->NO transform of packed integrals (misc.f:166).  ``tno`` is
    a host float64 array or a cuda float64 tensor, ``ure`` likewise
    (URe(NO,AO), column-major, i.e. a Fortran-ordered numpy array)."""
    ninte2 = len(tno)
    if isinstance(tno, np.ndarray):
        u = np.asfortranarray(ure, dtype=np.float64)
        _lib.call("gmc_twono1", tno.ctypes.data, u.ctypes.data, nbasis, ninte2)
        return tno
    _lib.call("gmc_twono1_dev", tno.data_ptr(), ure.data_ptr(), nbasis, ninte2, None)
    return tno


class DeviceSystem:
    """Device-resident inputs of one AC-CAS solve."""

    def __init__(self, twono, occ, xone, gamma_act, nbasis, ninact, nact, nele, device="cuda:0", ure=None):
        import torch
        self.nbasis, self.ninact, self.nact = nbasis, ninact, nact
        self.occ_host = np.asarray(occ, dtype=np.float64)
        self.indn_host = ACCAS_pairs(self.occ_host, nele)
        self.ndimx = len(self.indn_host)
        dev = torch.device(device)
        to = lambda a, dt: torch.as_tensor(np.ascontiguousarray(a), dtype=dt).to(dev)
        self.twono = twono if hasattr(twono, "data_ptr") else to(twono, torch.float64)
        self.occ = to(self.occ_host, torch.float64)
        self.igem = to(geminals(nbasis, ninact, nact), torch.int32)
        self.xone = to(xone, torch.float64)
        self.rdm2act = to(pack_rdm2act(gamma_act) if nact > 0 else np.zeros(1), torch.float64)
        self.indn = to(self.indn_host.ravel(), torch.int32)
        self.ure = None if ure is None else to(np.asfortranarray(ure).ravel(order="F"), torch.float64)


def ACECORR(dsys, alphas, weights, prec=GMC_FP64, return_eig=False, stream=None, alpha_batch=0):
    """ECorr = sum_i w_i W(alpha_i) over all alpha points in one batched solve
    (acfd.f:161-201).  Returns a dict with ecorr, walpha, ecasscf, nneg and,
    if asked, the ERPA eigenvalues per alpha (device tensor)."""
    import torch
    al = np.ascontiguousarray(alphas, dtype=np.float64)
    wt = np.ascontiguousarray(weights, dtype=np.float64)
    na = len(al)
    ec = np.zeros(1)
    wa = np.zeros(na)
    es = np.zeros(1)
    nneg = np.zeros(na, dtype=np.int32)
    eig = torch.zeros(na, dsys.ndimx, dtype=torch.float64, device=dsys.occ.device) if return_eig else None
    ure = dsys.ure.data_ptr() if dsys.ure is not None else None
    _lib.call("gmc_acecorr_dev", prec, alpha_batch, dsys.nbasis, dsys.ninact, dsys.nact, dsys.twono.data_ptr(),
              dsys.occ.data_ptr(), dsys.igem.data_ptr(), dsys.xone.data_ptr(), ure, dsys.rdm2act.data_ptr(),
              dsys.ndimx, dsys.indn.data_ptr(), na, al.ctypes.data, wt.ctypes.data, ec.ctypes.data, wa.ctypes.data,
              es.ctypes.data, eig.data_ptr() if eig is not None else None, nneg.ctypes.data, stream)
    out = dict(ecorr=float(ec[0]), walpha=wa, ecasscf=float(es[0]), nneg=nneg)
    if return_eig:
        out["eig"] = eig
    return out


def AC0CAS(dsys, stream=None):
    """The linearised AC correction (AC0CAS, acfd.f:1902-2614), FP64: the
    alpha=0 problem in AC0CAS's blocks and the first-order term with
    A' = A(1) - A(0).  Returns a dict with ecorr and ecasscf (w/o ENuc)."""
    ec, es = np.zeros(1), np.zeros(1)
    ure = dsys.ure.data_ptr() if dsys.ure is not None else None
    _lib.call("gmc_ac0cas_dev", dsys.nbasis, dsys.ninact, dsys.nact, dsys.twono.data_ptr(), dsys.occ.data_ptr(),
              dsys.igem.data_ptr(), dsys.xone.data_ptr(), ure, dsys.rdm2act.data_ptr(), dsys.ndimx,
              dsys.indn.data_ptr(), ec.ctypes.data, es.ctypes.data, stream)
    return dict(ecorr=float(ec[0]), ecasscf=float(es[0]))


def AC0CAS_FOFO(ao, uaomo, dsys, stream=None):
    """The FOFO form of the linearised AC correction (AC0CAS_FOFO,
    abfofo.f90:1485-1750), FP64, from packed AO integrals ``ao`` and orbitals
    ``uaomo`` as ACECORR_FOFO takes them.  Returns ecorr and ecasscf."""
    ec, es = np.zeros(1), np.zeros(1)
    _lib.call("gmc_ac0cas_fofo_dev", dsys.nbasis, dsys.ninact, dsys.nact, ao.data_ptr(), uaomo.data_ptr(),
              dsys.occ.data_ptr(), dsys.igem.data_ptr(), dsys.xone.data_ptr(), dsys.rdm2act.data_ptr(), dsys.ndimx,
              dsys.indn.data_ptr(), ec.ctypes.data, es.ctypes.data, stream)
    return dict(ecorr=float(ec[0]), ecasscf=float(es[0]))


def ACECORR_FOFO(ao, uaomo, dsys, alphas, weights, prec=GMC_FP64, alpha_batch=0, return_eig=False, stream=None):
    """The FOFO variant (ACEInteg_FOFO, abfofo.f90:3664): packed AO integrals
    ``ao`` and orbitals ``uaomo`` (URe(NO,AO), column-major, both cuda
    tensors); only (FF|OO) and (FO|FO) are built.  ``dsys`` supplies occ,
    IGem, XOne (NO basis), the active 2-RDM and IndN."""
    import torch
    al = np.ascontiguousarray(alphas, dtype=np.float64)
    wt = np.ascontiguousarray(weights, dtype=np.float64)
    na = len(al)
    ec, es = np.zeros(1), np.zeros(1)
    wa = np.zeros(na)
    nneg = np.zeros(na, dtype=np.int32)
    eig = torch.zeros(na, dsys.ndimx, dtype=torch.float64, device=dsys.occ.device) if return_eig else None
    _lib.call("gmc_acecorr_fofo_dev", prec, alpha_batch, dsys.nbasis, dsys.ninact, dsys.nact, ao.data_ptr(),
              uaomo.data_ptr(), dsys.occ.data_ptr(), dsys.igem.data_ptr(), dsys.xone.data_ptr(),
              dsys.rdm2act.data_ptr(), dsys.ndimx, dsys.indn.data_ptr(), na, al.ctypes.data, wt.ctypes.data,
              ec.ctypes.data, wa.ctypes.data, es.ctypes.data, eig.data_ptr() if eig is not None else None,
              nneg.ctypes.data, stream)
    out = dict(ecorr=float(ec[0]), walpha=wa, ecasscf=float(es[0]), nneg=nneg)
    if return_eig:
        out["eig"] = eig
    return out


# ---------------------------------------------------------------- host drop-ins
# numpy (host) arrays in and out, the reference's argument meaning; COMMON
# inputs (NInAcCAS, NAcCAS, IGem) and the rdm2.dat contents are explicit.

def _f(a, dt=np.float64):
    return np.asfortranarray(a, dtype=dt)


def _i(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def _ptr(a):
    return None if a is None else a.ctypes.data


def AB_CAS(twono, ure, occ, xone, indn, nbasis, ninact, nact, igem, rdm2act, acalpha):
    """A+B, A-B at ACAlpha (interpa.f:3659): (ABPLUS, ABMIN, ETot, CICoef),
    NDimX x NDimX Fortran-ordered, not symmetrised."""
    # every converted array is bound to a name: ctypes keeps only the address
    indn, ig = _i(indn), _i(igem)
    oc, xo, rd = _f(occ), _f(xone), _f(rdm2act)
    u = None if ure is None else _f(ure)
    nd = len(indn)
    ap = np.zeros((nd, nd), order="F")
    am = np.zeros((nd, nd), order="F")
    et = np.zeros(1)
    ci = np.zeros(nbasis)
    _lib.call("gmc_ab_cas", ap.ctypes.data, am.ctypes.data, et.ctypes.data, _ptr(u), oc.ctypes.data, xo.ctypes.data,
              twono.ctypes.data, len(twono), indn.ctypes.data, nd, nbasis, ninact, nact, ig.ctypes.data,
              rd.ctypes.data, float(acalpha), ci.ctypes.data)
    return ap, am, float(et[0]), ci


def ERPASYMM1(abplus, abmin):
    """ERPASYMM1 (interpa.f:1652): returns (EigVecR, Eig, APLSQRT, ABMIN
    symmetrised, number of negative eigenvalues of A+B)."""
    ap = _f(abplus).copy(order="F")
    am = _f(abmin).copy(order="F")
    nd = ap.shape[0]
    y = np.zeros((nd, nd), order="F")
    e = np.zeros(nd)
    nneg = np.zeros(1, dtype=np.int32)
    _lib.call("gmc_erpasymm1", y.ctypes.data, e.ctypes.data, ap.ctypes.data, am.ctypes.data, nd, nneg.ctypes.data)
    return y, e, ap, am, int(nneg[0])


def ACEneERPA(eigvecr, eig, twono, occ, igem, cicoef, indn, nbasis):
    """W(alpha) from ERPASYMM1's Y and Eig (acfd.f:4968)."""
    indn, ig = _i(indn), _i(igem)
    y, e, oc, ci = _f(eigvecr), _f(eig), _f(occ), _f(cicoef)
    nd = len(indn)
    ec = np.zeros(1)
    _lib.call("gmc_aceneerpa", ec.ctypes.data, y.ctypes.data, e.ctypes.data, twono.ctypes.data, len(twono),
              oc.ctypes.data, ig.ctypes.data, ci.ctypes.data, indn.ctypes.data, nbasis, nd)
    return float(ec[0])


def ACEInteg(twono, ure, occ, xone, indn, nbasis, ninact, nact, igem, rdm2act, acalpha, prec=GMC_FP64,
             side_outputs=False):
    """One alpha point (acfd.f:4832): returns dict(ecorr=W(alpha), egone=ECASSCF
    and, with side_outputs, ABPLUS=(A+B)^(1/2), ABMIN, EigVecR, Eig)."""
    indn, ig = _i(indn), _i(igem)
    oc, xo, rd = _f(occ), _f(xone), _f(rdm2act)
    u = None if ure is None else _f(ure)
    nd = len(indn)
    ec = np.zeros(1)
    eg = np.zeros(1)
    outs = [np.zeros((nd, nd), order="F") for _ in range(3)] + [np.zeros(nd)] if side_outputs else [None] * 4
    _lib.call("gmc_aceinteg", prec, ec.ctypes.data, eg.ctypes.data, twono.ctypes.data, len(twono), _ptr(u),
              oc.ctypes.data, xo.ctypes.data, ig.ctypes.data, rd.ctypes.data, indn.ctypes.data, nd, nbasis, ninact,
              nact, float(acalpha), *[_ptr(o) for o in outs])
    res = dict(ecorr=float(ec[0]), egone=float(eg[0]))
    if side_outputs:
        res.update(abplus=outs[0], abmin=outs[1], eigvecr=outs[2], eig=outs[3])
    return res


def ACEInteg_FOFO(jints, kints, ure, occ, xone, indn, nbasis, ninact, nact, igem, rdm2act, acalpha,
                  prec=GMC_FP64, side_outputs=False):
    """One alpha point of the FOFO variant (abfofo.f90:3664): jints and kints
    are the (FF|OO) and (FO|FO) classes in the record order of the FFOO and
    FOFO files.  Returns the same dict as ACEInteg."""
    indn, ig = _i(indn), _i(igem)
    oc, xo, rd = _f(occ), _f(xone), _f(rdm2act)
    jj, kk = _f(jints), _f(kints)
    u = None if ure is None else _f(ure)
    nd = len(indn)
    nO = ninact + nact
    if jj.size != nO * nO * nbasis * nbasis or kk.size != nbasis * nO * nbasis * nO:
        raise ValueError("ACEInteg_FOFO: jints/kints sizes do not match NBasis and the occupied count")
    ec = np.zeros(1)
    eg = np.zeros(1)
    outs = [np.zeros((nd, nd), order="F") for _ in range(3)] + [np.zeros(nd)] if side_outputs else [None] * 4
    _lib.call("gmc_aceinteg_fofo", prec, ec.ctypes.data, eg.ctypes.data, jj.ctypes.data, kk.ctypes.data, _ptr(u),
              oc.ctypes.data, xo.ctypes.data, ig.ctypes.data, rd.ctypes.data, indn.ctypes.data, nd, nbasis, ninact,
              nact, float(acalpha), *[_ptr(o) for o in outs])
    res = dict(ecorr=float(ec[0]), egone=float(eg[0]))
    if side_outputs:
        res.update(abplus=outs[0], abmin=outs[1], eigvecr=outs[2], eig=outs[3])
    return res


def ACECORR_host(twono, ure, occ, xone, indn, nbasis, ninact, nact, igem, rdm2act, alphas, weights,
                 prec=GMC_FP64, alpha_batch=0):
    """ACECORR's alpha loop with host inputs (acfd.f:161-201)."""
    indn, ig = _i(indn), _i(igem)
    oc, xo, rd = _f(occ), _f(xone), _f(rdm2act)
    u = None if ure is None else _f(ure)
    al = np.ascontiguousarray(alphas, dtype=np.float64)
    wt = np.ascontiguousarray(weights, dtype=np.float64)
    na = len(al)
    ec, es = np.zeros(1), np.zeros(1)
    wa = np.zeros(na)
    nneg = np.zeros(na, dtype=np.int32)
    _lib.call("gmc_acecorr", prec, alpha_batch, nbasis, ninact, nact, twono.ctypes.data, len(twono), oc.ctypes.data,
              ig.ctypes.data, xo.ctypes.data, _ptr(u), rd.ctypes.data, len(indn), indn.ctypes.data, na,
              al.ctypes.data, wt.ctypes.data, ec.ctypes.data, wa.ctypes.data, es.ctypes.data, nneg.ctypes.data)
    return dict(ecorr=float(ec[0]), walpha=wa, ecasscf=float(es[0]), nneg=nneg)


# ---------------------------------------------------------------- batching form
TUNING = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuning", "batch_gfx950.json")


def batch_cost():
    """The batch cost model {c0, c1, c2, c3} fitted by tune_batch.py (None
    when no tuning file is present: the library's built-in fit)."""
    if not os.path.exists(TUNING):
        return None
    with open(TUNING) as f:
        return _cost4(json.load(f)["cost"])


def _cost4(c):
    c = np.ascontiguousarray(c, dtype=np.float64)
    if c.shape != (4,):
        raise ValueError(f"batch cost model: 4 coefficients (c0 + c1 npad + c2 B npad^2 + c3 B npad^3), got {c.shape}")
    return c


def batch_plan(ndimx, nalpha, cost=None):
    """(group per system, padded order per system) of the batching form."""
    nd = np.ascontiguousarray(ndimx, dtype=np.int32)
    grp = np.zeros(len(nd), dtype=np.int32)
    npad = np.zeros(len(nd), dtype=np.int32)
    c = batch_cost() if cost is None else _cost4(cost)
    ng = _lib.lib().gmc_batch_plan(len(nd), nd.ctypes.data, nalpha, None if c is None else c.ctypes.data,
                                   grp.ctypes.data, npad.ctypes.data)
    _lib.check(0 if ng > 0 else ng)
    return grp, npad


def ACECORR_BATCH(systems, alphas, weights, prec=GMC_FP64, graph=False, cost=None, stream=None):
    """The batching form: ACECORR for many systems (DeviceSystem list) with
    all their alpha points in shared, padded batches.  Returns dict with
    ecorr and ecasscf arrays (one per system) and walpha (systems x alpha
    points)."""
    import ctypes
    n = len(systems)
    P = ctypes.c_void_p * n
    ints = lambda f: np.ascontiguousarray([f(d) for d in systems], dtype=np.int32)
    nb, ni, na, nd = ints(lambda d: d.nbasis), ints(lambda d: d.ninact), ints(lambda d: d.nact), ints(lambda d: d.ndimx)
    tw = P(*[d.twono.data_ptr() for d in systems])
    oc = P(*[d.occ.data_ptr() for d in systems])
    ig = P(*[d.igem.data_ptr() for d in systems])
    xo = P(*[d.xone.data_ptr() for d in systems])
    rd = P(*[d.rdm2act.data_ptr() for d in systems])
    ix = P(*[d.indn.data_ptr() for d in systems])
    al = np.ascontiguousarray(alphas, dtype=np.float64)
    wt = np.ascontiguousarray(weights, dtype=np.float64)
    c = batch_cost() if cost is None else _cost4(cost)
    ec, es, wa = np.zeros(n), np.zeros(n), np.zeros((n, len(al)))
    _lib.call("gmc_acecorr_batch_dev", prec, n, nb.ctypes.data, ni.ctypes.data, na.ctypes.data, nd.ctypes.data,
              ctypes.addressof(tw), ctypes.addressof(oc), ctypes.addressof(ig), ctypes.addressof(xo),
              ctypes.addressof(rd), ctypes.addressof(ix), len(al), al.ctypes.data, wt.ctypes.data,
              None if c is None else c.ctypes.data, 1 if graph else 0, ec.ctypes.data, es.ctypes.data,
              wa.ctypes.data, stream)
    return dict(ecorr=ec, ecasscf=es, walpha=wa)


# ---------------------------------------------------------------- SAPT
def SAPT_RESPONSE(dsys, stream=None):
    """A monomer's ERPA response for SAPT (sapt_main.f90:2593-2714): AB_CAS at
    alpha = 1, ERPASYMMXY, EigVecR = (C_p - C_q)(EigY - EigX).  Returns a
    dict of device tensors evec (ndimx x ndimx, column-major: evec[k] is
    column k), eig and cicoef."""
    import torch
    dev = dsys.occ.device
    n = dsys.ndimx
    evec = torch.zeros(n * n, dtype=torch.float64, device=dev)
    eig = torch.zeros(n, dtype=torch.float64, device=dev)
    cic = torch.zeros(dsys.nbasis, dtype=torch.float64, device=dev)
    ure = dsys.ure.data_ptr() if dsys.ure is not None else None
    _lib.call("gmc_sapt_response_dev", dsys.nbasis, dsys.ninact, dsys.nact, dsys.twono.data_ptr(),
              dsys.occ.data_ptr(), dsys.igem.data_ptr(), dsys.xone.data_ptr(), ure, dsys.rdm2act.data_ptr(), n,
              dsys.indn.data_ptr(), evec.data_ptr(), eig.data_ptr(), cic.data_ptr(), stream)
    return dict(evec=evec, eig=eig, cicoef=cic, indn=dsys.indn, ndimx=n, nbasis=dsys.nbasis)


def E2DISP(ra, rb, vab, stream=None):
    """Coupled E2disp (sapt_pol.f90:1684-2034) of two monomer responses and
    the intermolecular pair integrals vab (device, ndimx_A x ndimx_B,
    column-major, (pq|rs)_AB over the two pair lists).  Hartree."""
    e = np.zeros(1)
    _lib.call("gmc_e2disp_dev", ra["nbasis"], ra["ndimx"], ra["indn"].data_ptr(), ra["cicoef"].data_ptr(),
              ra["evec"].data_ptr(), ra["eig"].data_ptr(), rb["ndimx"], rb["indn"].data_ptr(),
              rb["cicoef"].data_ptr(), rb["evec"].data_ptr(), rb["eig"].data_ptr(), vab.data_ptr(), e.ctypes.data,
              stream)
    return float(e[0])


# ---------------------------------------------------------------- input files
def _buf(n):
    return np.zeros(n, dtype=np.float64)


def read_aotwoint_molpro(path, nbasis):
    """AO integrals of a Molpro AOTWOINT.mol (readtwoint, sorter.f90:133-313)
    as the NAddr3-packed array TwoNO1 takes."""
    t = nbasis * (nbasis + 1) // 2
    out = _buf(t * (t + 1) // 2)
    _lib.call("gmc_read_aotwoint_molpro", str(path).encode(), nbasis, out.ctypes.data)
    return out


def read_rdm2_molpro(path, nact, nost=1, nosym=0):
    """RDM2Act (NAddrRDM packing) from a '2RDM' record (types.f90:622-731)."""
    out = _buf(nact * nact * (nact * nact + 1) // 2)
    _lib.call("gmc_read_rdm2_molpro", str(path).encode(), nact, nost, nosym, out.ctypes.data)
    return out


def read_rdm1_molpro(path, nact, nost=1, nosym=0):
    """Halved active 1-RDM, column-packed upper triangle (types.f90:793-866)."""
    out = _buf(nact * (nact + 1) // 2)
    _lib.call("gmc_read_rdm1_molpro", str(path).encode(), nact, nost, nosym, out.ctypes.data)
    return out


def read_mo_molpro(path, nbasis, label="CASORB"):
    """Orbital coefficients (column-major nbasis x nbasis) under `label` of
    MOLPRO.MOPUN (types.f90:1078-1130)."""
    out = _buf(nbasis * nbasis)
    _lib.call("gmc_read_mo_molpro", str(path).encode(), label.encode(), nbasis, out.ctypes.data)
    return out.reshape(nbasis, nbasis, order="F")


def read_oneint_molpro(path, nbasis, label="ONEHAMIL"):
    """Packed lower triangle of the one-electron integrals under `label` of
    AOONEINT.mol (types.f90:1221-1283, expanded over all irreps)."""
    out = _buf(nbasis * (nbasis + 1) // 2)
    _lib.call("gmc_read_oneint_molpro", str(path).encode(), label.encode(), nbasis, out.ctypes.data)
    return out


def read_dpqrs(path, nbasis, header=False):
    """DPQRS.bin (initia.f:440-466): the NAddr3-packed integrals."""
    t = nbasis * (nbasis + 1) // 2
    out = _buf(t * (t + 1) // 2)
    _lib.call("gmc_read_dpqrs", str(path).encode(), nbasis, 1 if header else 0, out.ctypes.data)
    return out
