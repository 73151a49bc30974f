"""Writers of the reference's input files, so seeded synthetic systems can be
handed to the reference's own readers and to this library's readers
(gmc_read_* in include/gammcor_hip.h) alike.

Files (Fortran sequential unformatted records: a 4-byte length before and
after each record, little-endian), as the reference reads them:

* AOTWOINT.mol - Molpro AO two-electron integrals, symmetry-blocked
  (readtwoint, aosource 2: sorter.f90:133-313);
* 2RDM / 1RDM labelled records of the active density matrices
  (read_2rdm_molpro types.f90:622-731, read_1rdm_molpro 793-866);
* MOLPRO.MOPUN with the orbitals under a label such as 'CASORB  '
  (read_mo_molpro types.f90:1078-1130);
* AOONEINT.mol with the one-electron integrals under a label such as
  'ONEHAMIL' (readoneint_molpro types.f90:1221-1283);
* DPQRS.bin - the packed integrals as a raw stream of doubles in NAddr3
  order, optionally after three 4-byte integers (ReadDMRG, initia.f:440-466).

Index conventions here are 0-based; the records hold the reference's 1-based
indices where they hold indices.
"""
from __future__ import annotations

import struct

import numpy as np


class FortranWriter:
    """Sequential unformatted records with 4-byte length markers."""

    def __init__(self, path):
        self.f = open(path, "wb")

    def record(self, *items):
        payload = b"".join(_bytes(x) for x in items)
        n = struct.pack("<i", len(payload))
        self.f.write(n + payload + n)

    def close(self):
        self.f.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def _bytes(x):
    if isinstance(x, bytes):
        return x
    if isinstance(x, str):
        return x.encode("ascii")
    a = np.asarray(x)
    if a.dtype.kind in "iu":
        return np.ascontiguousarray(a, dtype="<i4").tobytes()
    return np.ascontiguousarray(a, dtype="<f8").tobytes()


def label8(text):
    t = text.ljust(8)[:8]
    return t


def _pair(i, j):
    """0-based packed index of (i, j), i >= j."""
    return i * (i + 1) // 2 + j


def _naddr3(p, q, r, s):
    a, b = (p, q) if p >= q else (q, p)
    c, d = (r, s) if r >= s else (s, r)
    x, y = _pair(a, b), _pair(c, d)
    return _pair(x, y) if x >= y else _pair(y, x)


def symmetry_offsets(nt):
    off = [0] * 8
    for i in range(1, 8):
        off[i] = off[i - 1] + (nt[i - 1] if i - 1 < len(nt) else 0)
    return off


def write_aotwoint_molpro(path, twoel, nt):
    """AO integrals (packed NAddr3 array over all AOs, AOs ordered by irrep)
    in the symmetry-blocked layout readtwoint reads (sorter.f90:133-313).
    nt: AOs per irrep (irreps 1..nsk, product of irreps = XOR of their
    0-based labels); integrals between irreps whose product is not totally
    symmetric must be zero and are not written."""
    nt = list(nt)
    nsk = len(nt)
    assert 1 <= nsk <= 8
    nt8 = nt + [0] * (8 - nsk)
    off = symmetry_offsets(nt8)
    tw = np.asarray(twoel)

    def v(p, q, r, s):   # 1-based global indices
        return tw[_naddr3(p - 1, q - 1, r - 1, s - 1)]

    with FortranWriter(path) as w:
        w.record(np.array([nsk]))
        w.record(np.array(nt8))
        # all four indices in one irrep: for every rs (r <= s) the (pq|rs) with pq <= rs
        for sp in range(nsk):
            if nt8[sp] == 0:
                continue
            w.record(np.array([sp + 1]))          # section header (skipped by the reader)
            o = off[sp]
            for s in range(1, nt8[sp] + 1):
                for r in range(1, s + 1):
                    vals = []
                    for q in range(1, s + 1):
                        pend = q if q < s else r
                        for p in range(1, pend + 1):
                            vals.append(v(o + p, o + q, o + r, o + s))
                    w.record(np.array([s, r]))
                    w.record(np.array(vals, dtype=np.float64))
            w.record(np.array([0, 0]))
        # two irreps: p, q in irrep P < R, r, s in irrep R
        for sr in range(1, nsk):
            for sp in range(sr):
                if nt8[sp] == 0 or nt8[sr] == 0:
                    continue
                w.record(np.array([sp + 1, sr + 1]))
                op, orr = off[sp], off[sr]
                for s in range(1, nt8[sr] + 1):
                    for r in range(1, s + 1):
                        vals = []
                        for q in range(1, nt8[sp] + 1):
                            for p in range(1, q + 1):
                                P, Q, Rr, S = op + p, op + q, orr + r, orr + s
                                vals += [v(P, Q, Rr, S), v(P, Rr, Q, S), v(Q, Rr, P, S)]
                        w.record(np.array([s, r]))
                        w.record(np.array(vals, dtype=np.float64))
                w.record(np.array([0, 0]))
        # four different irreps: sym_s > sym_r > sym_q > sym_p with p^q = r^s
        for ss in range(3, nsk):
            for sr in range(2, ss):
                srs = sr ^ ss
                for sq in range(1, sr):
                    sp = sq ^ srs
                    if sp >= sq:
                        continue
                    if 0 in (nt8[sp], nt8[sq], nt8[sr], nt8[ss]):
                        continue
                    w.record(np.array([sp + 1, sq + 1, sr + 1, ss + 1]))
                    op, oq, orr, os_ = off[sp], off[sq], off[sr], off[ss]
                    for s in range(1, nt8[ss] + 1):
                        for r in range(1, nt8[sr] + 1):
                            vals = []
                            for q in range(1, nt8[sq] + 1):
                                for p in range(1, nt8[sp] + 1):
                                    P, Q, Rr, S = op + p, oq + q, orr + r, os_ + s
                                    vals += [v(P, Q, Rr, S), v(P, Rr, Q, S), v(Q, Rr, P, S)]
                            w.record(np.array([s, r]))
                            w.record(np.array(vals, dtype=np.float64))
                    w.record(np.array([0, 0]))


def write_rdm_molpro(path, gamma_act, occ_act, state=1, sym=1):
    """The labelled '1RDM' and '2RDM' records of one state.

    gamma_act: active Gamma(p,q,r,s) in the AB_CAS convention
    (E2 = sum Gamma(p,q,r,s) (pr|qs)); the 2RDM record holds
    D(i,j,k,l) = 2 Gamma(i,k,j,l) for i, j and k <= i (l <= j when k == i),
    the order read_2rdm_molpro unpacks (types.f90:707-727).
    occ_act: natural occupations in [0, 1]; the 1RDM record holds the
    spin-summed matrix 2 diag(occ) (read_1rdm_molpro halves it, 853-860)."""
    g = np.asarray(gamma_act)
    na = g.shape[0]
    work = []
    for i in range(na):
        for j in range(na):
            for k in range(i + 1):
                lend = j + 1 if k == i else na
                for l in range(lend):
                    work.append(2.0 * g[i, k, j, l])
    d1 = np.zeros((na, na))
    d1[np.arange(na), np.arange(na)] = 2.0 * np.asarray(occ_act)
    with FortranWriter(path) as w:
        w.record(label8("1RDM"))
        w.record(np.array([na, na * na, 1]))
        w.record(np.array([sym, 1]))
        w.record(np.array([state]))
        w.record(d1.ravel(order="F"))
        w.record(label8("2RDM"))
        w.record(np.array([1, 1]))            # ic1d, number of state symmetries
        w.record(np.array([sym, 1]))
        w.record(np.array([state]))
        w.record(np.array(work, dtype=np.float64))


def write_mo_molpro(path, cmo, nbas, label="CASORB  "):
    """Orbital coefficients under `label` (read_mo_molpro,
    types.f90:1078-1130): per irrep its nbas x nbas block, column-major,
    placed at the irrep's offset."""
    cmo = np.asarray(cmo)
    nsym = len(nbas)
    offs = [sum(nbas[:i]) for i in range(nsym)]
    tmp = []
    for irep in range(nsym):
        o = offs[irep]
        blk = cmo[o:o + nbas[irep], o:o + nbas[irep]]
        tmp.append(blk.ravel(order="F"))
    with FortranWriter(path) as w:
        w.record(label8("OTHER"))
        w.record(np.zeros(3))
        w.record(label8(label))
        w.record(np.array([nsym] + list(nbas) + offs))
        w.record(np.concatenate(tmp))


def write_oneint_molpro(path, mats, nbas):
    """One-electron integrals (AOONEINT.mol, readoneint_molpro
    types.f90:1221-1283): a first record, the symmetry record (nsym,
    nbas, offsets), a third record, then for each label of `mats` (label ->
    full symmetric matrix) the label and the per-irrep packed triangles
    (column-wise, i <= j)."""
    nsym = len(nbas)
    offs = [sum(nbas[:i]) for i in range(nsym)]
    with FortranWriter(path) as w:
        w.record(label8("AOONEINT"))
        w.record(np.array([nsym] + list(nbas) + offs))
        w.record(label8("BASINFO"))
        for lab, m in mats.items():
            m = np.asarray(m)
            tmp = []
            for irep in range(nsym):
                o = offs[irep]
                for j in range(nbas[irep]):
                    for i in range(j + 1):
                        tmp.append(m[o + i, o + j])
            w.record(label8(lab))
            w.record(np.array(tmp, dtype=np.float64))


def write_dpqrs(path, twoel, header=None):
    """DPQRS.bin (ReadDMRG, initia.f:440-466): the packed integrals as a
    raw stream in NAddr3 order, after three 4-byte integers when `header`
    is given (the LiborNew layout)."""
    with open(path, "wb") as f:
        if header is not None:
            f.write(np.asarray(header, dtype="<i4").tobytes())
        f.write(np.ascontiguousarray(twoel, dtype="<f8").tobytes())


def read_aotwosort(path, nbasis):
    """The reference's AOTWOSORT direct-access file (readtwoint SORT2,
    sorter.f90:375-387: record rs holds (pq|rs) for every pq) as the packed
    NAddr3 array."""
    t = nbasis * (nbasis + 1) // 2
    recs = np.fromfile(path, dtype="<f8").reshape(t, t)     # [rs, pq]
    out = np.empty(t * (t + 1) // 2)
    for ij in range(t):
        out[ij * (ij + 1) // 2: ij * (ij + 1) // 2 + ij + 1] = recs[ij, : ij + 1]
    return out
