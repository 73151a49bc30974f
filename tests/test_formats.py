"""Input files of the reference's front end: the writers of
gammcor_amd.formats and the library's readers (gmc_read_*, host code, no
GPU), pinned by the reference's own readers.

- Round trips: symmetry-blocked AO integrals (1, 2, 4 and 8 irreps), the
  labelled RDM records, orbitals, one-electron integrals and DPQRS.bin come
  back bit for bit.
- The compiled reference (oracle/_ref, driver mode `formats`) reads the same
  files with readtwoint (sorter.f90:31), read_2rdm_molpro, read_1rdm_molpro,
  read_mo_molpro and readoneint_molpro (types.f90): its arrays equal the
  library's and the seeded inputs'.
- DPQRS.bin: the stream order is the loop of initia.f:446-460, restated
  here, and equals the NAddr3 order.
"""
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF_DRIVER = os.path.join(ROOT, "oracle", "_ref", "ref_driver")


def _sym_integrals(nt, rank=None, seed=0):
    from gammcor_amd import synth
    n = sum(nt)
    chol = synth.cholesky_factors_sym(nt, rank or 3 * n, 0.1, seed)
    return synth.ao_integrals_packed(chol)


@pytest.mark.parametrize("nt", [[7], [4, 3], [3, 2, 4, 2], [2, 1, 2, 1, 1, 2, 1, 2], [3, 0, 2, 2]])
def test_aotwoint_round_trip(tmp_path, nt):
    from gammcor_amd import formats, host, synth
    ao = _sym_integrals(nt, seed=len(nt))
    n = sum(nt)
    # symmetry zeros are present (the file leaves them out)
    if len([x for x in nt if x]) > 1:
        lab = np.repeat(np.arange(len(nt)), nt)
        i, j, k, l_ = 0, int(np.argmax(lab == lab[-1])), 0, 0
        if lab[i] ^ lab[j] ^ lab[k] ^ lab[l_]:
            assert ao[synth.naddr3(i, j, k, l_)] == 0.0
    f = tmp_path / "AOTWOINT.mol"
    formats.write_aotwoint_molpro(f, ao, nt)
    got = host.read_aotwoint_molpro(f, n)
    assert np.array_equal(got, ao)


def test_dpqrs_stream_order(tmp_path):
    from gammcor_amd import formats, host
    n = 6
    ao = _sym_integrals([n], seed=3)
    for header in (None, (n, n, n)):
        f = tmp_path / "DPQRS.bin"
        formats.write_dpqrs(f, ao, header)
        assert np.array_equal(host.read_dpqrs(f, n, header is not None), ao)
    # initia.f:446-460: I, J <= I, K, L <= K, IJ >= KL; the value read next
    # goes to NAddr3(I, J, K, L)
    from gammcor_amd import synth
    order = []
    ij = 0
    for i in range(1, n + 1):
        for j in range(1, i + 1):
            ij += 1
            kl = 0
            for k in range(1, n + 1):
                for l_ in range(1, k + 1):
                    kl += 1
                    if ij >= kl:
                        order.append(synth.naddr3(i - 1, j - 1, k - 1, l_ - 1))
    assert order == list(range(len(ao)))


def _system():
    from gammcor_amd import synth
    return synth.make_system(12, 2, 4, 4, 36, 0.1, 0)


def _write_all(d, s, nt):
    from gammcor_amd import formats, synth
    n = s.nbasis
    ao = _sym_integrals(nt, seed=11)
    formats.write_aotwoint_molpro(d / "AOTWOINT.mol", ao, nt)
    occ_act = s.occ[s.ninact:s.ninact + s.nact]
    formats.write_rdm_molpro(d / "RDM.mol", s.gamma_act, occ_act)
    cmo = s.uaomo.T.copy()                    # C(AO, MO)
    formats.write_mo_molpro(d / "MOLPRO.MOPUN", cmo, [n])
    h_no = synth.unpack_sym(s.xone, n)
    h_ao = cmo @ h_no @ cmo.T
    h_ao = 0.5 * (h_ao + h_ao.T)              # exactly symmetric
    formats.write_oneint_molpro(d / "AOONEINT.mol", {"OVERLAP": np.eye(n), "ONEHAMIL": h_ao}, [n])
    return ao, occ_act, cmo, h_ao


def test_readers_round_trip(tmp_path):
    from gammcor_amd import host, synth
    s = _system()
    nt = [5, 3, 2, 2]
    ao, occ_act, cmo, h_ao = _write_all(tmp_path, s, nt)
    n, na = s.nbasis, s.nact
    assert np.array_equal(host.read_aotwoint_molpro(tmp_path / "AOTWOINT.mol", n), ao)
    assert np.array_equal(host.read_rdm2_molpro(tmp_path / "RDM.mol", na), host.pack_rdm2act(s.gamma_act))
    d1 = host.read_rdm1_molpro(tmp_path / "RDM.mol", na)
    iu = np.triu_indices(na)
    ref1 = np.diag(occ_act)[iu[0], iu[1]]
    # column-packed upper triangle: order by column, then row
    order = np.lexsort((iu[0], iu[1]))
    assert np.array_equal(d1, ref1[order])
    assert np.array_equal(host.read_mo_molpro(tmp_path / "MOLPRO.MOPUN", n), cmo)
    assert np.array_equal(host.read_oneint_molpro(tmp_path / "AOONEINT.mol", n), synth.pack_sym(h_ao))


def test_missing_label_is_an_error(tmp_path):
    from gammcor_amd import host, _lib
    s = _system()
    _write_all(tmp_path, s, [12])
    with pytest.raises(_lib.GmcError, match="not found"):
        host.read_mo_molpro(tmp_path / "MOLPRO.MOPUN", s.nbasis, label="NATORB")
    with pytest.raises(_lib.GmcError, match="NBasis"):
        host.read_aotwoint_molpro(tmp_path / "AOTWOINT.mol", s.nbasis + 1)


@pytest.mark.skipif(not os.path.exists(REF_DRIVER), reason="oracle/_ref/ref_driver not built (make -C oracle ref)")
def test_reference_readers_read_the_same_arrays(tmp_path):
    from gammcor_amd import formats, host, synth
    s = _system()
    nt = [5, 3, 2, 2]
    ao, occ_act, cmo, h_ao = _write_all(tmp_path, s, nt)
    n, na = s.nbasis, s.nact
    hdr = np.array([n, s.ninact, na, s.nele, 1], dtype=np.int32)
    (tmp_path / "inputs.bin").write_bytes(hdr.tobytes())
    r = subprocess.run(["bash", "-c", f"ulimit -s unlimited; {REF_DRIVER} formats"], cwd=tmp_path,
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0 and "FORMATS_OK    0" in r.stdout, r.stdout[-2000:] + r.stderr[-2000:]
    # readtwoint's sorted records: every (pq|rs), both orientations
    assert np.array_equal(formats.read_aotwosort(tmp_path / "AOTWOSORT", n), ao)
    rdm2 = np.fromfile(tmp_path / "ref_rdm2.bin")
    assert np.array_equal(rdm2, host.read_rdm2_molpro(tmp_path / "RDM.mol", na))
    assert np.array_equal(rdm2, host.pack_rdm2act(s.gamma_act))
    rdm1 = np.fromfile(tmp_path / "ref_rdm1.bin")
    assert np.array_equal(rdm1[: na * (na + 1) // 2], host.read_rdm1_molpro(tmp_path / "RDM.mol", na))
    assert np.all(rdm1[na * (na + 1) // 2:] == 0.0)
    assert np.array_equal(np.fromfile(tmp_path / "ref_cmo.bin").reshape(n, n, order="F"), cmo)
    assert np.array_equal(np.fromfile(tmp_path / "ref_oneint.bin"),
                          host.read_oneint_molpro(tmp_path / "AOONEINT.mol", n))
    assert np.array_equal(np.fromfile(tmp_path / "ref_oneint.bin"), synth.pack_sym(h_ao))
