"""Seeded synthetic AC-CAS inputs.

The reference reads integrals, orbitals and RDMs from Molpro/Dalton dumps
(initia.f:995 ``LdInteg``).  For benchmarking and parity work we build the same
objects from seeds instead:

* AO integrals (mu nu|la si) = sum_P L_P[mu,nu] L_P[la,si] with symmetric L_P,
  stored 8-fold packed exactly as ``NAddr3`` addresses them (misc.f:505-525).
* Orthonormal MO coefficients from the QR factor of a seeded Gaussian matrix.
* An exact 1-/2-RDM of a seeded random CI vector over all determinants of a
  CAS(ne, no) space (Sz = 0), rotated to natural orbitals with descending
  occupations as ``LdInteg`` does (initia.f:1165-1171).
* The active 2-RDM is written as ``rdm2.dat`` in the reference's F19.12 text
  format (initia.f:1552-1566) and read back, so every consumer sees the same
  rounded values.

Conventions (all 0-based here):
  pair(i, j) = max*(max+1)/2 + min, packed4(ij, kl) = pair(ij, kl)
  Gamma[p,q,r,s] = 1/2 sum_{st} <a+_{p s} a+_{q t} a_{s t} a_{r s}>  so that
  E2 = sum Gamma[p,q,r,s] (pr|qs)   (interpa.f:3755-3771, FRDM2 5244-5272)
  occupations n_p = gamma_pp / 2 in [0, 1].
"""
from __future__ import annotations

import dataclasses
import itertools
import math
import os
from typing import Optional

import numpy as np

# Named workloads.  sigma/rank/ramp choices that the workload descriptions leave
# open are fixed here and recorded with every fixture.
CONFIGS = {
    "n48_cas44": dict(nbasis=48, ninact=6, nact=4, nelact=4, rank=144, sigma=0.1, seed=0, nalpha=8),
    "n160_cas88": dict(nbasis=160, ninact=10, nact=8, nelact=8, rank=480, sigma=0.05, seed=1, nalpha=8),
    "n240_cas88": dict(nbasis=240, ninact=12, nact=8, nelact=8, rank=720, sigma=0.05, seed=2, nalpha=16),
    "n400_cas1010": dict(nbasis=400, ninact=20, nact=10, nelact=10, rank=1200, sigma=0.05, seed=4, nalpha=16),
}

H_SIGMA = 0.1          # spread of the symmetric Gaussian part of h
# Target orbital energies of the Fock-like operator h + G(n): inactive in
# [-GAP-W_INACT, -GAP], active in [-W_ACT, W_ACT], virtual in [GAP, GAP+W_VIRT].
GAP, W_INACT, W_ACT, W_VIRT = 10.0, 2.0, 1.0, 10.0
CI_REF_WEIGHT = 0.8    # |c_ref|^2 of the leading determinant in the random CI vector


def pair_index(i, j):
    """0-based packed index of the unordered pair (i, j) (misc.f:516-519)."""
    a = np.maximum(i, j)
    b = np.minimum(i, j)
    return a * (a + 1) // 2 + b


def naddr3(p, q, r, s):
    """0-based position of (pq|rs) in the 8-fold packed array (misc.f:505-525)."""
    return pair_index(pair_index(p, q), pair_index(r, s))


@dataclasses.dataclass
class System:
    nbasis: int
    ninact: int
    nact: int
    nelact: int
    occ: np.ndarray          # (N,) NO occupations, inactive 1.0, virtual 0.0
    xone: np.ndarray         # (N(N+1)/2,) one-electron matrix in the NO basis, packed
    uaomo: np.ndarray        # (N, N) UAOMO[no, ao]
    rdm2_text: str           # rdm2.dat contents
    gamma_act: np.ndarray    # (nA,nA,nA,nA) active Gamma as read back from rdm2.dat
    twoel_ao: Optional[np.ndarray] = None   # (NInte2,) packed AO integrals
    chol: Optional[np.ndarray] = None        # (rank, T) packed Cholesky factors
    meta: dict = dataclasses.field(default_factory=dict)

    @property
    def nele(self):
        return self.ninact + self.nelact // 2

    @property
    def ninte1(self):
        return self.nbasis * (self.nbasis + 1) // 2

    @property
    def ninte2(self):
        t = self.ninte1
        return t * (t + 1) // 2

    @property
    def igem(self):
        """Geminal labels 1 inactive / 2 active / 3 virtual (initia.f:1202-1216)."""
        g = np.full(self.nbasis, 3, dtype=np.int32)
        g[: self.ninact] = 1
        g[self.ninact: self.ninact + self.nact] = 2
        if self.ninact == 0:
            g[: self.nact] = 1
            g[self.nact:] = 2
        return g


# ---------------------------------------------------------------- CI / RDMs

def _strings(norb, nel):
    return [sum(1 << k for k in c) for c in itertools.combinations(range(norb), nel)]


def _excitation_ops(norb, strings):
    """Dense matrices of a+_p a_q (one spin) on the given string space."""
    index = {s: i for i, s in enumerate(strings)}
    ns = len(strings)
    ops = np.zeros((norb, norb, ns, ns))
    for col, s in enumerate(strings):
        for q in range(norb):
            if not (s >> q) & 1:
                continue
            sign_q = (-1) ** bin(s & ((1 << q) - 1)).count("1")
            s1 = s ^ (1 << q)
            for p in range(norb):
                if (s1 >> p) & 1:
                    continue
                sign_p = (-1) ** bin(s1 & ((1 << p) - 1)).count("1")
                s2 = s1 | (1 << p)
                ops[p, q, index[s2], col] += sign_q * sign_p
    return ops


def cas_rdms(nact, nelact, seed, ref_weight=None):
    """Spin-summed 1-RDM gamma[p,q] = <E_pq> and 2-RDM D[p,q,r,s] =
    <E_pr E_qs> - delta_qr <E_ps> of a normalised random CI vector: a seeded
    Gaussian vector over all determinants scaled to norm sqrt(1 - w) plus
    sqrt(w) on the determinant with the lowest nelact/2 orbitals doubly
    occupied (w = ref_weight), then renormalised."""
    if ref_weight is None:
        ref_weight = CI_REF_WEIGHT
    if nact == 0:
        return np.zeros((0, 0)), np.zeros((0, 0, 0, 0))
    na = nelact // 2
    strings = _strings(nact, na)
    ns = len(strings)
    rng = np.random.default_rng(seed)
    psi = rng.standard_normal((ns, ns))
    psi *= math.sqrt(1.0 - ref_weight) / np.linalg.norm(psi)
    psi[0, 0] += math.sqrt(ref_weight)
    psi /= np.linalg.norm(psi)
    ops = _excitation_ops(nact, strings)
    # E_pq Psi = Ea_pq Psi + Psi Eb_pq^T (same string space for both spins)
    phi = np.einsum("pqij,jk->pqik", ops, psi) + np.einsum("ij,pqkj->pqik", psi, ops)
    gamma = np.einsum("ik,pqik->pq", psi, phi)
    flat = phi.reshape(nact * nact, -1)
    g = (flat @ flat.T).reshape(nact, nact, nact, nact)       # g[r,p,q,s] = <E_rp Psi|E_qs Psi>
    d = np.einsum("rpqs->pqrs", g) - np.einsum("qr,ps->pqrs", np.eye(nact), gamma)
    return gamma, d


def rdm2_dat_text(gamma2):
    """rdm2.dat as written by initia.f:1552-1566: for I,J,K,L over active with
    (I,J) >= (K,L) in row-major pair order, the line 'K I L J  2*Gamma(I,J,K,L)'
    in (4I4,F19.12)."""
    na = gamma2.shape[0]
    lines = []
    for i in range(na):
        for j in range(na):
            ij = i * na + j
            for k in range(na):
                for l in range(na):
                    kl = k * na + l
                    if ij >= kl:
                        lines.append("%4d%4d%4d%4d%19.12f" % (k + 1, i + 1, l + 1, j + 1, 2.0 * gamma2[i, j, k, l]))
    return "\n".join(lines) + "\n"


def read_rdm2_dat(text, nact):
    """Parse rdm2.dat the way AB_CAS does (interpa.f:3734-3750): each line
    'I J K L X' sets Gamma(J,L,I,K) = X/2, stored symmetric in (JL)<->(IK)."""
    g = np.zeros((nact, nact, nact, nact))
    for line in text.splitlines():
        if not line.strip():
            continue
        parts = line.split()
        i, j, k, l = (int(v) - 1 for v in parts[:4])
        x = 0.5 * float(parts[4])
        g[j, l, i, k] = x
        g[i, k, j, l] = x
    return g


# ---------------------------------------------------------------- integrals

def cholesky_factors(nbasis, rank, sigma, seed):
    """Packed symmetric factors: row P, column pair(mu,nu) ~ N(0, sigma^2),
    drawn in the order P, then packed pair index."""
    t = nbasis * (nbasis + 1) // 2
    rng = np.random.default_rng(seed)
    return rng.standard_normal((rank, t)) * sigma


def cholesky_factors_sym(nt, rank, sigma, seed):
    """Packed factors of integrals with point-group symmetry: the AOs come
    in irrep blocks of sizes nt (irrep product = XOR of the 0-based labels,
    len(nt) a power of two); factor P belongs to irrep P mod len(nt) and is
    N(0, sigma^2) on the pairs (mu, nu) whose product is that irrep, zero
    elsewhere, so (mu nu|la si) vanishes unless the product of all four
    irreps is totally symmetric."""
    nirr = len(nt)
    assert nirr in (1, 2, 4, 8)
    lab = np.repeat(np.arange(nirr), nt)
    n = len(lab)
    il, jl = np.tril_indices(n)
    gam = lab[il] ^ lab[jl]                     # pair order i >= j, row-major == packed order
    rng = np.random.default_rng(seed)
    f = rng.standard_normal((rank, len(il))) * sigma
    f *= (gam[None, :] == (np.arange(rank) % nirr)[:, None])
    return f


def ao_integrals_packed(chol):
    """(ij|kl) = sum_P L[P,ij] L[P,kl] for ij >= kl, packed by NAddr3."""
    t = chol.shape[1]
    out = np.empty(t * (t + 1) // 2)
    blk = 2048
    for r0 in range(0, t, blk):
        r1 = min(t, r0 + blk)
        v = chol[:, r0:r1].T @ chol[:, :r1]        # rows ij in [r0,r1), cols kl < r1
        for ij in range(r0, r1):
            base = ij * (ij + 1) // 2
            out[base: base + ij + 1] = v[ij - r0, : ij + 1]
    return out


def orbitals(nbasis, seed):
    rng = np.random.default_rng(seed)
    q, r = np.linalg.qr(rng.standard_normal((nbasis, nbasis)))
    return q * np.sign(np.diag(r))[None, :]


def orbital_energies(ninact, nact, nvirt):
    return np.concatenate([np.linspace(-GAP - W_INACT, -GAP, ninact), np.linspace(-W_ACT, W_ACT, nact),
                           np.linspace(GAP, GAP + W_VIRT, nvirt)])


def mean_field(chol, uaomo, occ):
    """G_pq = sum_t n_t [2 (pq|tt) - (pt|qt)] in the NO basis, from the packed
    Cholesky factors (B_P = U L_P U^T)."""
    nb = uaomo.shape[0]
    il, jl = np.tril_indices(nb)
    g = np.zeros((nb, nb))
    for p0 in range(0, chol.shape[0], 64):
        lp = np.zeros((min(64, chol.shape[0] - p0), nb, nb))
        lp[:, il, jl] = chol[p0:p0 + 64]
        lp[:, jl, il] = chol[p0:p0 + 64]
        b = np.matmul(np.matmul(uaomo[None], lp), uaomo.T[None])
        dn = np.einsum("Ptt,t->P", b, occ)
        g += 2.0 * np.einsum("Ppq,P->pq", b, dn)
        g -= np.einsum("Ppt,Pqt->pq", b * occ[None, None, :], b)
    return g


def one_electron(nbasis, ninact, nact, seed, gmf):
    """h = diag(eps) - G(n) + symmetric N(0, H_SIGMA^2) noise (NO basis)."""
    rng = np.random.default_rng(seed)
    h = np.zeros((nbasis, nbasis))
    iu = np.triu_indices(nbasis)
    h[iu] = rng.standard_normal(len(iu[0])) * H_SIGMA
    h = h + np.triu(h, 1).T
    h += np.diag(orbital_energies(ninact, nact, nbasis - ninact - nact)) - gmf
    return h


def pack_sym(m):
    n = m.shape[0]
    i, j = np.tril_indices(n)
    return m[i, j]            # row-major lower triangle == pair order i>=j


def unpack_sym(v, n):
    m = np.zeros((n, n))
    i, j = np.tril_indices(n)
    m[i, j] = v
    m[j, i] = v
    return m


def make_system(nbasis, ninact, nact, nelact, rank, sigma, seed, with_integrals=True, ci_ref_weight=None,
                energy_scale=1.0, occ_act_set=(), degenerate_active=None, **_):
    """One seeded system.  ci_ref_weight: |c|^2 put on the leading determinant
    of the CI vector (None: CI_REF_WEIGHT; 0: a plain normalised Gaussian
    vector over all determinants).

    Edge cases of the reference's branches (the RDMs stay those of the CI
    vector; reference and library read the same inputs):
      energy_scale: h and the two-electron integrals times this factor (the
        Cholesky factors times its square root), so every ERPA omega scales
        by it - used to put omegas on either side of the 1e-3 filter;
      occ_act_set: (k, value) pairs overriding active occupation k;
      degenerate_active: (k1, k2) gives active orbital k2 the occupation of
        k1, a pair the pair selection discards (accas.f:67-71)."""
    if ci_ref_weight is None:
        ci_ref_weight = CI_REF_WEIGHT
    s_l, s_c, s_ci, s_h = seed * 10 + 0, seed * 10 + 1, seed * 10 + 2, seed * 10 + 3
    gamma, d = cas_rdms(nact, nelact, s_ci, ref_weight=ci_ref_weight)
    w, v = np.linalg.eigh(gamma) if nact else (np.zeros(0), np.zeros((0, 0)))
    order = np.argsort(-w, kind="stable")
    w, v = w[order], v[:, order]
    if nact:
        v = v * np.where(v[np.argmax(np.abs(v), axis=0), np.arange(nact)] < 0, -1.0, 1.0)[None, :]
    occ_act = np.abs(w) / 2.0
    d_no = np.einsum("pqrs,pa,qb,rc,sd->abcd", d, v, v, v, v, optimize=True)
    gamma2 = 0.5 * d_no
    text = rdm2_dat_text(gamma2)
    gamma_read = read_rdm2_dat(text, nact)

    c = orbitals(nbasis, s_c)
    c_no = c.copy()
    c_no[:, ninact:ninact + nact] = c[:, ninact:ninact + nact] @ v
    occ = np.zeros(nbasis)
    occ[:ninact] = 1.0
    occ[ninact:ninact + nact] = occ_act
    chol = cholesky_factors(nbasis, rank, sigma, s_l)
    h = one_electron(nbasis, ninact, nact, s_h, mean_field(chol, c_no.T, occ))
    for k, v in occ_act_set:
        occ[ninact + int(k)] = float(v)
    if degenerate_active is not None:
        occ[ninact + int(degenerate_active[1])] = occ[ninact + int(degenerate_active[0])]
    if energy_scale != 1.0:
        h = h * energy_scale
        chol = chol * math.sqrt(energy_scale)
    sysm = System(nbasis=nbasis, ninact=ninact, nact=nact, nelact=nelact, occ=occ,
                  xone=pack_sym(h), uaomo=np.ascontiguousarray(c_no.T), rdm2_text=text,
                  gamma_act=gamma_read)
    sysm.meta = dict(nbasis=nbasis, ninact=ninact, nact=nact, nelact=nelact, rank=rank, sigma=sigma,
                     seed=seed, seeds=dict(chol=s_l, orbitals=s_c, ci=s_ci, hcore=s_h),
                     h_sigma=H_SIGMA, gap=GAP, widths=[W_INACT, W_ACT, W_VIRT],
                     ci_ref_weight=ci_ref_weight, enuc=0.0)
    if energy_scale != 1.0:
        sysm.meta["energy_scale"] = energy_scale
    if occ_act_set:
        sysm.meta["occ_act_set"] = [[int(k), float(v)] for k, v in occ_act_set]
    if degenerate_active is not None:
        sysm.meta["degenerate_active"] = [int(x) for x in degenerate_active]
    sysm.chol = chol
    if with_integrals:
        sysm.twoel_ao = ao_integrals_packed(chol)
    return sysm


def fofo_records(twono, nbasis, nocc):
    """The (FF|OO) and (FO|FO) integral classes in the record order of the
    reference's FFOO and FOFO files (tran4_gen, tran.f90:313-423), taken from
    packed NO integrals: J[(k + nO l) N^2 + i + N j] = (kl|ij) with k, l
    occupied; K[(f + N o) N nO + f' + N o'] = (fo|f'o') with o, o' occupied."""
    n, no = nbasis, nocc
    i = np.arange(n)
    o = np.arange(no)
    l_, k_, j_, i_ = np.meshgrid(o, o, i, i, indexing="ij")
    jrec = np.asarray(twono)[naddr3(k_, l_, i_, j_)]
    o_, f_, o2, f2 = np.meshgrid(o, i, o, i, indexing="ij")
    krec = np.asarray(twono)[naddr3(f_, o_, f2, o2)]
    return np.ascontiguousarray(jrec.reshape(-1)), np.ascontiguousarray(krec.reshape(-1))


# The batching workload: 64 independent systems of uneven size, NBasis uniform
# in [40, 140] from seed 3, NBasis/12 inactive orbitals, CAS(4e,4o) and
# CAS(6e,6o) alternating, rank 3 NBasis integrals.  The description leaves
# sigma and the per-system seeds open: sigma 0.1 as for the NBasis=48 system,
# system k seeded 1000 + k.
BATCH = dict(count=64, nbasis_range=(40, 140), seed=3, sigma=0.1, nalpha=16, seed0=1000)


def batch_configs(count=BATCH["count"], seed=BATCH["seed"]):
    """Per-system keyword arguments of make_system for the batching workload."""
    rng = np.random.default_rng(seed)
    lo, hi = BATCH["nbasis_range"]
    out = []
    for k in range(count):
        n = int(rng.integers(lo, hi + 1))
        na = 4 if k % 2 == 0 else 6
        out.append(dict(nbasis=n, ninact=n // 12, nact=na, nelact=na, rank=3 * n, sigma=BATCH["sigma"],
                        seed=BATCH["seed0"] + k))
    return out


def ao_integrals_device(chol, device):
    """The packed AO integrals (mn|ls) = sum_P L_Pmn L_Pls of ao_integrals_packed,
    formed on the device in row blocks of the pair matrix (torch)."""
    import torch
    t = chol.shape[1]
    c = torch.from_numpy(np.ascontiguousarray(chol)).to(device)
    ao = torch.empty(t * (t + 1) // 2, dtype=torch.float64, device=device)
    blk = 1024
    cols = torch.arange(t, device=device)
    for r0 in range(0, t, blk):
        r1 = min(t, r0 + blk)
        block = c[:, r0:r1].T @ c[:, :r1]
        rows = torch.arange(r0, r1, device=device)[:, None]
        mask = cols[None, :r1] <= rows
        ao[r0 * (r0 + 1) // 2: r1 * (r1 + 1) // 2] = block[mask]
    return ao


def system_from_meta(meta, with_integrals=True):
    """The system a fixture records (its `system` entry)."""
    keys = ("nbasis", "ninact", "nact", "nelact", "rank", "sigma", "seed", "ci_ref_weight", "energy_scale",
            "occ_act_set", "degenerate_active")
    return make_system(with_integrals=with_integrals, **{k: meta[k] for k in keys if k in meta})


def config_system(name, with_integrals=True):
    cfg = dict(CONFIGS[name])
    return make_system(with_integrals=with_integrals, **cfg)


def gauss_legendre_01(n):
    """Gauss-Legendre nodes/weights on [0,1] from numpy (used by hosts that do
    not call the library's own GauLeg)."""
    x, w = np.polynomial.legendre.leggauss(n)
    return 0.5 * (x + 1.0), 0.5 * w


def write_inputs(sysm, directory, alphas, weights, tag=""):
    """Write a system in the binary layout read by the reference driver
    (oracle/ref_driver.f90): little-endian int32 header then float64 arrays.
    tag "_A" / "_B" names a monomer's files (inputs_A.bin, rdm2_A.dat)."""
    os.makedirs(directory, exist_ok=True)
    hdr = np.array([sysm.nbasis, sysm.ninact, sysm.nact, sysm.nele, len(alphas)], dtype=np.int32)
    with open(os.path.join(directory, f"inputs{tag}.bin"), "wb") as f:
        f.write(hdr.tobytes())
        f.write(np.ascontiguousarray(sysm.occ, dtype=np.float64).tobytes())
        f.write(np.ascontiguousarray(sysm.xone, dtype=np.float64).tobytes())
        # UAOMO(NO,AO) column-major for Fortran == row-major transpose
        f.write(np.asfortranarray(sysm.uaomo).tobytes(order="F"))
        f.write(np.asarray(alphas, dtype=np.float64).tobytes())
        f.write(np.asarray(weights, dtype=np.float64).tobytes())
    with open(os.path.join(directory, "twoel_ao.bin"), "wb") as f:
        f.write(np.ascontiguousarray(sysm.twoel_ao, dtype=np.float64).tobytes())
    with open(os.path.join(directory, f"rdm2{tag}.dat"), "w") as f:
        f.write(sysm.rdm2_text)


def dimer(nbasis, ninact, nact, nelact, rank, sigma, seed_a, seed_b):
    """Two monomers in one basis for SAPT: A and B each from its own seed
    (orbitals, CI vector, h), both with the AO integrals of A's Cholesky
    factors.  Returns (A, B, vab) with the intermolecular NO integrals
    vab[p, q, r, s] = (pq|rs)_AB = sum_P (U_A L_P U_A^T)_pq (U_B L_P U_B^T)_rs,
    p, q in A's orbitals and r, s in B's."""
    a = make_system(nbasis, ninact, nact, nelact, rank, sigma, seed_a)
    b = make_system(nbasis, ninact, nact, nelact, rank, sigma, seed_b, with_integrals=False)
    b.twoel_ao, b.chol = a.twoel_ao, a.chol
    b.meta["shares_ao_of_seed"] = seed_a
    il, jl = np.tril_indices(nbasis)
    lp = np.zeros((a.chol.shape[0], nbasis, nbasis))
    lp[:, il, jl] = a.chol
    lp[:, jl, il] = a.chol
    ba = np.matmul(np.matmul(a.uaomo[None], lp), a.uaomo.T[None])
    bb = np.matmul(np.matmul(b.uaomo[None], lp), b.uaomo.T[None])
    vab = np.einsum("Ppq,Prs->pqrs", ba, bb, optimize=True)
    return a, b, vab


def twomoab_records(vab, a, b):
    """The TWOMOAB direct-access records e2disp reads (sapt_pol.f90:1652-1664):
    record iq + (ip - num0_A - 1) dimO_A (1-based ip > num0_A, iq <= dimO_A)
    holds, at is + (ir - num0_B - 1) dimO_B, (ip iq|ir is)_AB; num0 = inactive
    orbitals, dimO = inactive + active."""
    n = a.nbasis
    n0a, doa = a.ninact, a.ninact + a.nact
    n0b, dob = b.ninact, b.ninact + b.nact
    nova, novb = doa * (n - n0a), dob * (n - n0b)
    out = np.zeros((nova, novb))
    for ip in range(n0a + 1, n + 1):
        for iq in range(1, doa + 1):
            rec = iq + (ip - n0a - 1) * doa
            blk = vab[ip - 1, iq - 1, n0b:, :dob]          # [ir - n0b - 1, is - 1]
            out[rec - 1, :] = blk.reshape(-1)                # is fastest
    return out
