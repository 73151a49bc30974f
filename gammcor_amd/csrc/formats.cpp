// Readers of the reference's input files (host code, no GPU): they return
// the arrays the AC-CAS entry points take, from the files the reference's
// front end reads.  Fortran sequential unformatted records carry a 4-byte
// length before and after their payload.
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/gammcor_hip.h"

namespace gmc {
void set_error(const char* fmt, ...);
}

namespace {

using gmc::set_error;

struct RecFile {
    FILE* f = nullptr;
    explicit RecFile(const char* path) { f = std::fopen(path, "rb"); }
    ~RecFile()
    {
        if (f) std::fclose(f);
    }
    // next record's payload; false at end of file or on a malformed record
    bool next(std::vector<char>& rec)
    {
        int32_t n = 0, m = 0;
        if (std::fread(&n, 4, 1, f) != 1) return false;
        if (n < 0) return false;
        rec.resize((size_t)n);
        if (n && std::fread(rec.data(), 1, (size_t)n, f) != (size_t)n) return false;
        if (std::fread(&m, 4, 1, f) != 1 || m != n) return false;
        return true;
    }
    bool skip()
    {
        std::vector<char> r;
        return next(r);
    }
};

template <class T>
bool get(const std::vector<char>& rec, size_t idx, T* out)
{
    if ((idx + 1) * sizeof(T) > rec.size()) return false;
    std::memcpy(out, rec.data() + idx * sizeof(T), sizeof(T));
    return true;
}

inline int64_t pair1(int64_t i, int64_t j)   // 1-based packed index of (i, j), any order
{
    const int64_t a = i > j ? i : j, b = i > j ? j : i;
    return a * (a - 1) / 2 + b;
}

inline int64_t naddr3(int64_t p, int64_t q, int64_t r, int64_t s)   // 0-based position (misc.f:505-525)
{
    const int64_t pq = pair1(p, q), rs = pair1(r, s);
    return pair1(pq, rs) - 1;
}

bool label_is(const std::vector<char>& rec, const char* text)
{
    char lab[8];
    std::memset(lab, ' ', 8);
    std::memcpy(lab, text, std::min<size_t>(8, std::strlen(text)));
    return rec.size() >= 8 && std::memcmp(rec.data(), lab, 8) == 0;
}

}  // namespace

extern "C" {

// readtwoint, aosource 2 (Molpro), sorter.f90:133-313: the three sections
// of AOTWOINT.mol - one irrep, two irreps, four irreps - read into the
// NAddr3-packed array instead of the reference's sorted AOTWOSORT records.
int gmc_read_aotwoint_molpro(const char* path, int nbasis, double* twono)
{
    RecFile rf(path);
    if (!rf.f) {
        set_error("AOTWOINT: cannot open %s", path);
        return -1;
    }
    const int64_t t = (int64_t)nbasis * (nbasis + 1) / 2, ninte2 = t * (t + 1) / 2;
    std::memset(twono, 0, (size_t)ninte2 * sizeof(double));
    std::vector<char> rec;
    int32_t nsk = 0, nt[8] = {0};
    if (!rf.next(rec) || !get(rec, 0, &nsk) || nsk < 1 || nsk > 8) {
        set_error("AOTWOINT: bad symmetry count in %s", path);
        return -2;
    }
    if (!rf.next(rec) || rec.size() < 8 * sizeof(int32_t)) {
        set_error("AOTWOINT: bad irrep sizes in %s", path);
        return -2;
    }
    std::memcpy(nt, rec.data(), 8 * sizeof(int32_t));
    int ntoff[8] = {0};
    int tot = 0;
    for (int i = 0; i < nsk; i++) tot += nt[i];
    for (int i = 1; i < 8; i++) ntoff[i] = ntoff[i - 1] + nt[i - 1];
    if (tot != nbasis) {
        set_error("AOTWOINT: irrep sizes sum to %d, not NBasis = %d", tot, nbasis);
        return -3;
    }
    auto put = [&](int64_t p, int64_t q, int64_t r, int64_t s, double v) { twono[naddr3(p, q, r, s)] = v; };
    // reads the (idx_s, idx_r) + values records of one section until (<=0, <=0)
    auto section = [&](auto&& body) -> int {
        for (;;) {
            int32_t is = 0, ir = 0;
            if (!rf.next(rec) || !get(rec, 0, &is) || !get(rec, 1, &ir)) return -4;
            if (is <= 0 && ir <= 0) return 0;
            std::vector<char> vals;
            if (!rf.next(vals)) return -4;
            const int rc = body(is, ir, vals);
            if (rc) return rc;
        }
    };
    int rc = 0;
    for (int sp = 0; sp < nsk && !rc; sp++) {
        if (nt[sp] == 0) continue;
        if (!rf.skip()) return set_error("AOTWOINT: truncated file %s", path), -4;
        const int o = ntoff[sp];
        rc = section([&](int is, int ir, const std::vector<char>& vals) -> int {
            const size_t nints = (size_t)ir + (size_t)is * (is - 1) / 2;
            if (vals.size() != nints * sizeof(double)) return -5;
            const int S = is + o, R = ir + o;
            size_t idx = 0;
            for (int q = o + 1; q <= S; q++) {
                const int pend = q < S ? q : R;
                for (int p = o + 1; p <= pend; p++) {
                    double v;
                    std::memcpy(&v, vals.data() + 8 * idx++, 8);
                    put(p, q, R, S, v);
                }
            }
            return 0;
        });
    }
    for (int sr = 1; sr < nsk && !rc; sr++)
        for (int sp = 0; sp < sr && !rc; sp++) {
            if (nt[sp] == 0 || nt[sr] == 0) continue;
            if (!rf.skip()) return set_error("AOTWOINT: truncated file %s", path), -4;
            rc = section([&](int is, int ir, const std::vector<char>& vals) -> int {
                const size_t nints = 3 * (size_t)nt[sp] * (nt[sp] + 1) / 2;
                if (vals.size() != nints * sizeof(double)) return -5;
                const int S = is + ntoff[sr], R = ir + ntoff[sr];
                size_t idx = 0;
                for (int q = ntoff[sp] + 1; q <= ntoff[sp] + nt[sp]; q++)
                    for (int p = ntoff[sp] + 1; p <= q; p++) {
                        double v[3];
                        std::memcpy(v, vals.data() + 8 * idx, 24);
                        idx += 3;
                        put(p, q, R, S, v[0]);
                        put(p, R, q, S, v[1]);
                        put(q, R, p, S, v[2]);
                    }
                return 0;
            });
        }
    for (int ss = 3; ss < nsk && !rc; ss++)
        for (int sr = 2; sr < ss && !rc; sr++) {
            const int srs = sr ^ ss;
            for (int sq = 1; sq < sr && !rc; sq++) {
                const int sp = sq ^ srs;
                if (sp >= sq) continue;
                if (!nt[sp] || !nt[sq] || !nt[sr] || !nt[ss]) continue;
                if (!rf.skip()) return set_error("AOTWOINT: truncated file %s", path), -4;
                rc = section([&](int is, int ir, const std::vector<char>& vals) -> int {
                    const size_t nints = 3 * (size_t)nt[sp] * nt[sq];
                    if (vals.size() != nints * sizeof(double)) return -5;
                    const int S = is + ntoff[ss], R = ir + ntoff[sr];
                    size_t idx = 0;
                    for (int q = ntoff[sq] + 1; q <= ntoff[sq] + nt[sq]; q++)
                        for (int p = ntoff[sp] + 1; p <= ntoff[sp] + nt[sp]; p++) {
                            double v[3];
                            std::memcpy(v, vals.data() + 8 * idx, 24);
                            idx += 3;
                            put(p, q, R, S, v[0]);
                            put(p, R, q, S, v[1]);
                            put(q, R, p, S, v[2]);
                        }
                    return 0;
                });
            }
        }
    if (rc == -4) {
        set_error("AOTWOINT: truncated or malformed record in %s", path);
        return -4;
    }
    if (rc == -5) {
        set_error("AOTWOINT: a value record of %s has the wrong length", path);
        return -5;
    }
    return 0;
}

// read_2rdm_molpro, types.f90:622-731: the active 2-RDM of state
// (nost, nosym) from the '2RDM' record into RDM2Act (NAddrRDM packing,
// nact^2 (nact^2+1)/2 values), halved as the reference stores it.  nosym = 0
// takes the file's first state, as the reference does without a state input.
int gmc_read_rdm2_molpro(const char* path, int nact, int nost, int nosym, double* rdm2act)
{
    RecFile rf(path);
    if (!rf.f) return set_error("2RDM: cannot open %s", path), -1;
    const int64_t n2 = (int64_t)nact * nact, trsq = n2 * (n2 + 1) / 2;
    std::vector<char> rec, work;
    bool found = false;
    while (!found && rf.next(rec)) {
        if (!label_is(rec, "2RDM")) continue;
        int32_t ic1d, nstsym;
        if (!rf.next(rec) || !get(rec, 0, &ic1d) || !get(rec, 1, &nstsym)) break;
        const int nsets = nosym > 0 ? nstsym : 1;
        for (int k = 0; k < nsets && !found; k++) {
            int32_t isym, nstate;
            if (!rf.next(rec) || !get(rec, 0, &isym) || !get(rec, 1, &nstate)) return -4;
            const int nst = nosym > 0 ? nstate : 1;
            for (int i = 0; i < nst; i++) {
                int32_t ist;
                if (!rf.next(rec) || !get(rec, 0, &ist) || !rf.next(work)) return -4;
                if (nosym <= 0 || (ist == nost && isym == nosym)) {
                    found = true;
                    break;
                }
            }
        }
        break;
    }
    if (!found) return set_error("2RDM: state %d.%d not in %s", nost, nosym, path), -6;
    if ((int64_t)work.size() < trsq * 8) return set_error("2RDM: record of %s too short", path), -5;
    const double* w = reinterpret_cast<const double*>(work.data());
    int64_t ijkl = 0;
    for (int i = 1; i <= nact; i++)
        for (int j = 1; j <= nact; j++)
            for (int k = 1; k <= i; k++) {
                const int lend = k == i ? j : nact;
                for (int l = 1; l <= lend; l++) {
                    const double v = 0.5 * w[ijkl++];
                    rdm2act[pair1((int64_t)(i - 1) * nact + k, (int64_t)(j - 1) * nact + l) - 1] = v;
                    rdm2act[pair1((int64_t)(k - 1) * nact + i, (int64_t)(l - 1) * nact + j) - 1] = v;
                }
            }
    return 0;
}

// read_1rdm_molpro, types.f90:793-866: the active 1-RDM of state
// (nost, nosym), halved, as the upper triangle packed by columns
// (nact (nact+1)/2 values: the first entries of the reference's onerdm).
int gmc_read_rdm1_molpro(const char* path, int nact, int nost, int nosym, double* onerdm)
{
    RecFile rf(path);
    if (!rf.f) return set_error("1RDM: cannot open %s", path), -1;
    std::vector<char> rec, work;
    bool found = false;
    int32_t na = 0, na2 = 0;
    while (!found && rf.next(rec)) {
        if (!label_is(rec, "1RDM")) continue;
        int32_t nstsym;
        if (!rf.next(rec) || !get(rec, 0, &na) || !get(rec, 1, &na2) || !get(rec, 2, &nstsym)) break;
        const int nsets = nosym > 0 ? nstsym : 1;
        for (int k = 0; k < nsets && !found; k++) {
            int32_t isym, nstate;
            if (!rf.next(rec) || !get(rec, 0, &isym) || !get(rec, 1, &nstate)) return -4;
            const int nst = nosym > 0 ? nstate : 1;
            for (int i = 0; i < nst; i++) {
                int32_t ist;
                if (!rf.next(rec) || !get(rec, 0, &ist) || !rf.next(work)) return -4;
                if (nosym <= 0 || (ist == nost && isym == nosym)) {
                    found = true;
                    break;
                }
            }
        }
        break;
    }
    if (!found) return set_error("1RDM: state %d.%d not in %s", nost, nosym, path), -6;
    if (na != nact || (int64_t)work.size() < (int64_t)na * na * 8)
        return set_error("1RDM: %s holds %d active orbitals, expected %d", path, na, nact), -5;
    const double* w = reinterpret_cast<const double*>(work.data());
    int64_t idx = 0;
    for (int j = 1; j <= nact; j++)
        for (int i = 1; i <= j; i++) onerdm[idx++] = 0.5 * w[(i - 1) + (int64_t)(j - 1) * nact];
    return 0;
}

// read_mo_molpro, types.f90:1078-1130: the orbitals under `label` (such as
// "CASORB") of a MOLPRO.MOPUN file, irrep blocks placed at their offsets,
// cmo column-major nbasis x nbasis.
int gmc_read_mo_molpro(const char* path, const char* label, int nbasis, double* cmo)
{
    RecFile rf(path);
    if (!rf.f) return set_error("MO: cannot open %s", path), -1;
    std::vector<char> rec, tmp;
    while (rf.next(rec)) {
        if (!label_is(rec, label)) continue;
        int32_t nsym;
        if (!rf.next(rec) || !get(rec, 0, &nsym) || nsym < 1 || nsym > 8) return -4;
        int32_t nbas[8], offs[8];
        for (int i = 0; i < nsym; i++)
            if (!get(rec, 1 + i, &nbas[i]) || !get(rec, 1 + nsym + i, &offs[i])) return -4;
        if (!rf.next(tmp)) return -4;
        std::memset(cmo, 0, (size_t)nbasis * nbasis * sizeof(double));
        size_t idx = 0;
        const size_t nval = tmp.size() / 8;
        for (int r = 0; r < nsym; r++)
            for (int j = 0; j < nbas[r]; j++)
                for (int i = 0; i < nbas[r]; i++) {
                    if (idx >= nval || offs[r] + nbas[r] > nbasis) return set_error("MO: bad block in %s", path), -5;
                    double v;
                    std::memcpy(&v, tmp.data() + 8 * idx++, 8);
                    cmo[(offs[r] + i) + (size_t)(offs[r] + j) * nbasis] = v;
                }
        return 0;
    }
    return set_error("MO: label %s not found in %s", label, path), -6;
}

// readoneint_molpro, types.f90:1221-1283 (expand = true): the one-electron
// integrals under `label` (such as "ONEHAMIL") of an AOONEINT.mol file, as
// the full packed lower triangle over all orbitals (N(N+1)/2, pair order
// max(max-1)/2 + min).
int gmc_read_oneint_molpro(const char* path, const char* label, int nbasis, double* mone)
{
    RecFile rf(path);
    if (!rf.f) return set_error("ONEINT: cannot open %s", path), -1;
    std::vector<char> rec, tmp;
    int32_t nsym = 0, nbas[8] = {0}, offs[8] = {0};
    if (!rf.skip() || !rf.next(rec) || !get(rec, 0, &nsym) || nsym < 1 || nsym > 8 || !rf.skip())
        return set_error("ONEINT: bad header in %s", path), -4;
    for (int i = 0; i < nsym; i++)
        if (!get(rec, 1 + i, &nbas[i]) || !get(rec, 1 + nsym + i, &offs[i])) return -4;
    while (rf.next(rec)) {
        if (!label_is(rec, label)) continue;
        if (!rf.next(tmp)) return -4;
        const int64_t ntr = (int64_t)nbasis * (nbasis + 1) / 2;
        std::memset(mone, 0, (size_t)ntr * sizeof(double));
        size_t idx = 0;
        const size_t nval = tmp.size() / 8;
        for (int r = 0; r < nsym; r++)
            for (int j = 1; j <= nbas[r]; j++)
                for (int i = 1; i <= j; i++) {
                    if (idx >= nval || offs[r] + j > nbasis) return set_error("ONEINT: bad block in %s", path), -5;
                    double v;
                    std::memcpy(&v, tmp.data() + 8 * idx++, 8);
                    mone[pair1(offs[r] + j, offs[r] + i) - 1] = v;
                }
        return 0;
    }
    return set_error("ONEINT: label %s not found in %s", label, path), -6;
}

// ReadDMRG's DPQRS.bin, initia.f:440-466: NInte2 doubles in NAddr3 order as
// a raw stream, after three 4-byte integers when header != 0 (LiborNew).
int gmc_read_dpqrs(const char* path, int nbasis, int header, double* twono)
{
    FILE* f = std::fopen(path, "rb");
    if (!f) return set_error("DPQRS: cannot open %s", path), -1;
    const int64_t t = (int64_t)nbasis * (nbasis + 1) / 2, ninte2 = t * (t + 1) / 2;
    int ok = 1;
    if (header) {
        int32_t h[3];
        ok = std::fread(h, 4, 3, f) == 3;
    }
    const size_t got = ok ? std::fread(twono, sizeof(double), (size_t)ninte2, f) : 0;
    std::fclose(f);
    if ((int64_t)got != ninte2) return set_error("DPQRS: %s holds %zu of %lld integrals", path, got, (long long)ninte2), -4;
    return 0;
}

}  // extern "C"
