/*
 * ORACLE — test infrastructure only.  CPU restatement of the reference's
 * AB_CAS (interpa.f:3659-4104) with FRDM2 (interpa.f:5244-5272), NAddrRDM
 * (interpa.f:4529-4541) and NAddr3 (misc.f:505-525).  Used by tests/ and by
 * bench.py's cpu_baseline leg as the checker; never by the product path.
 *
 * Indices are 0-based here.  Matrices are column-major like the Fortran.
 * The 2-RDM is passed as a dense nA^4 array gact[((p*nA+q)*nA+r)*nA+s] =
 * Gamma(p,q,r,s), symmetric under (pq)<->(rs), as read from rdm2.dat.
 * IGFact(NAddr3(i,j,k,l)) of the reference is the predicate "i,j,k,l all in
 * one geminal", evaluated here directly from igem[] (same value for every
 * index permutation, so identical to the packed table).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

static inline int64_t pidx(int64_t i, int64_t j)
{
    int64_t a = i > j ? i : j, b = i > j ? j : i;
    return a * (a + 1) / 2 + b;
}

static inline int64_t naddr3(int p, int q, int r, int s) { return pidx(pidx(p, q), pidx(r, s)); }

typedef struct {
    const double *occ, *gact;
    const int *ind2; /* active index or -1 */
    int na;
} rdm_t;

/* FRDM2 (interpa.f:5244-5272) */
static double frdm2(const rdm_t *R, int p, int q, int r, int s)
{
    const double *occ = R->occ;
    double v = 0.0;
    if (p == r && q == s && (occ[p] == 1.0 || occ[q] == 1.0)) v += 2.0 * occ[p] * occ[q];
    if (p == s && q == r && (occ[p] == 1.0 || occ[q] == 1.0)) v -= occ[p] * occ[q];
    if (R->ind2[p] >= 0 && R->ind2[q] >= 0 && R->ind2[r] >= 0 && R->ind2[s] >= 0) {
        int na = R->na;
        v = R->gact[(((int64_t)R->ind2[p] * na + R->ind2[q]) * na + R->ind2[r]) * na + R->ind2[s]];
    }
    return v;
}

static inline double fac(const int *igem, double alpha, int i, int j, int k, int l)
{
    return (igem[i] == igem[j] && igem[j] == igem[k] && igem[k] == igem[l]) ? 1.0 : alpha;
}

/*
 * ure: URe(NBasis,NBasis) column-major; xone: packed one-electron matrix;
 * twono: packed NO integrals; indn: IndN(2,NDimX) column-major, 1-based
 * values as in the reference.  Outputs ABPLUS/ABMIN (ndimx x ndimx,
 * column-major), ETot (ECASSCF w/o ENuc) and the CICoef vector.
 */
int oracle_ab_cas(int nb, int ninact, int nact, const double *occ, const double *xone,
                  const double *ure, const double *twono, const double *gact, const int *igem,
                  int ndimx, const int *indn, double alpha, double *abplus, double *abmin,
                  double *etot, double *cicoef)
{
    const int nocc = ninact + nact;
    const int64_t ninte1 = (int64_t)nb * (nb + 1) / 2;
    double *hno = calloc(ninte1, sizeof(double));
    double *auxi = calloc(ninte1, sizeof(double));
    double *auxio = calloc(ninte1, sizeof(double));
    double *wmat = calloc((size_t)nb * nb, sizeof(double));
    double *c = calloc(nb, sizeof(double));
    int *ind1 = calloc(nact > 0 ? nact : 1, sizeof(int));
    int *ind2 = malloc(nb * sizeof(int));
    if (!hno || !auxi || !auxio || !wmat || !c || !ind1 || !ind2) return -1;
    memset(abplus, 0, sizeof(double) * ndimx * (size_t)ndimx);
    memset(abmin, 0, sizeof(double) * ndimx * (size_t)ndimx);

    /* HNO = U h U^T (interpa.f:3702-3716) */
#pragma omp parallel for schedule(dynamic)
    for (int i = 0; i < nb; i++)
        for (int j = 0; j <= i; j++) {
            double s = 0.0;
            for (int a = 0; a < nb; a++)
                for (int b = 0; b < nb; b++)
                    s += ure[i + (size_t)a * nb] * ure[j + (size_t)b * nb] * xone[pidx(a, b)];
            hno[pidx(i, j)] = s;
        }

    for (int i = 0; i < nb; i++) ind2[i] = -1;
    for (int i = 0; i < nact; i++) {
        ind1[i] = ninact + i;
        ind2[ninact + i] = i;
    }
    rdm_t R = {occ, gact, ind2, nact};

    /* ECASSCF (interpa.f:3755-3771) */
    double e = 0.0;
    for (int i = 0; i < nb; i++) e += 2.0 * occ[i] * hno[pidx(i, i)];
    for (int p = 0; p < nocc; p++)
        for (int q = 0; q < nocc; q++)
            for (int r = 0; r < nocc; r++)
                for (int s = 0; s < nocc; s++) e += frdm2(&R, p, q, r, s) * twono[naddr3(p, r, q, s)];
    *etot = e;

    /* C_p = +-sqrt(n_p) (interpa.f:3777-3781) */
    for (int i = 0; i < nb; i++) {
        c[i] = sqrt(occ[i]);
        if (occ[i] < 0.5) c[i] = -c[i];
        cicoef[i] = c[i];
    }

    /* alpha one-electron part (interpa.f:3785-3810) */
    for (int i = 0; i < nb; i++)
        for (int j = 0; j <= i; j++) {
            int64_t ij = pidx(i, j);
            if (igem[i] != igem[j]) {
                hno[ij] = alpha * hno[ij];
            } else {
                double aux = 0.0;
                for (int t = 0; t < nb; t++)
                    if (igem[t] != igem[i])
                        aux += occ[t] * (2.0 * twono[naddr3(t, t, i, j)] - twono[naddr3(t, i, t, j)]);
                hno[ij] += (1.0 - alpha) * aux;
            }
        }

    /* AuxI, AuxIO (interpa.f:3843-3858) */
    for (int p = 0; p < nb; p++)
        for (int q = 0; q <= p; q++) {
            double a1 = 0.0, a2 = 0.0;
            for (int t = 0; t < nocc; t++) {
                double f = fac(igem, alpha, t, t, p, q);
                double v = occ[t] * f * (2.0 * twono[naddr3(p, q, t, t)] - twono[naddr3(p, t, q, t)]);
                a1 += v;
                if (t < ninact) a2 += v;
            }
            auxi[pidx(p, q)] = a1;
            auxio[pidx(p, q)] = a2;
        }

    /* WMAT (interpa.f:3862-3890) */
#pragma omp parallel for schedule(dynamic)
    for (int p = 0; p < nb; p++)
        for (int r = 0; r < nocc; r++) {
            double s = 0.0;
            for (int t = 0; t < nocc; t++)
                for (int w = 0; w < nocc; w++)
                    for (int u = 0; u < nocc; u++) {
                        double f = fac(igem, alpha, t, w, p, u);
                        s += twono[naddr3(t, w, p, u)] * frdm2(&R, w, u, t, r) * f +
                             twono[naddr3(t, u, p, w)] * frdm2(&R, w, u, r, t) * f;
                    }
            wmat[p + (size_t)r * nb] = s;
        }

    /* main loop (interpa.f:3894-4075) */
#pragma omp parallel for schedule(dynamic)
    for (int irow = 0; irow < ndimx; irow++) {
        const int ir = indn[2 * irow] - 1, is = indn[2 * irow + 1] - 1;
        for (int icol = 0; icol < ndimx; icol++) {
            const int ipp = indn[2 * icol] - 1, iqq = indn[2 * icol + 1] - 1;
            for (int o = 0; o < 2; o++) {
                /* the two orientations (p,q) = (qq,pp) then (pp,qq) */
                const int ip = o == 0 ? iqq : ipp, iq = o == 0 ? ipp : iqq;
                const int64_t iqs = pidx(is, iq), ipr = pidx(ir, ip);
                double a = 0.0;
                if (ip == ir) a += (occ[ip] - occ[is]) * hno[iqs];
                if (is == iq) a += (occ[iq] - occ[ir]) * hno[ipr];
                const double fpqrs = fac(igem, alpha, ip, iq, ir, is);
                const double auxpqrs =
                    fpqrs * (2.0 * twono[naddr3(ip, iq, ir, is)] - twono[naddr3(ip, ir, iq, is)]);
                /* T1+T2 */
                if (occ[ip] * occ[ir] != 0.0) {
                    if (occ[ip] == 1.0 || occ[ir] == 1.0) {
                        a += occ[ip] * occ[ir] * auxpqrs;
                        if (ip == ir) a += occ[ip] * auxi[iqs];
                    } else {
                        for (int tt = 0; tt < nact; tt++)
                            for (int uu = 0; uu < nact; uu++) {
                                int it = ind1[tt], iu = ind1[uu];
                                double f = fac(igem, alpha, is, iq, it, iu);
                                a += twono[naddr3(is, iq, it, iu)] * frdm2(&R, ip, iu, ir, it) * f +
                                     twono[naddr3(is, iu, it, iq)] * frdm2(&R, ip, iu, it, ir) * f;
                            }
                        if (ip == ir) a += occ[ip] * auxio[iqs];
                    }
                }
                /* T3+T4 */
                if (occ[iq] * occ[is] != 0.0) {
                    if (occ[iq] == 1.0 || occ[is] == 1.0) {
                        a += occ[iq] * occ[is] * auxpqrs;
                        if (iq == is) a += occ[iq] * auxi[ipr];
                    } else {
                        for (int tt = 0; tt < nact; tt++)
                            for (int uu = 0; uu < nact; uu++) {
                                int it = ind1[tt], iu = ind1[uu];
                                double f = fac(igem, alpha, iu, it, ip, ir);
                                a += twono[naddr3(iu, it, ip, ir)] * frdm2(&R, is, it, iq, iu) * f +
                                     twono[naddr3(iu, ir, ip, it)] * frdm2(&R, is, it, iu, iq) * f;
                            }
                        if (iq == is) a += occ[iq] * auxio[ipr];
                    }
                }
                /* T5 */
                if (occ[ir] * occ[iq] != 0.0) {
                    if (occ[iq] == 1.0 || occ[ir] == 1.0) {
                        a -= occ[iq] * occ[ir] * auxpqrs;
                    } else {
                        for (int tt = 0; tt < nact; tt++)
                            for (int uu = 0; uu < nact; uu++) {
                                int it = ind1[tt], iu = ind1[uu];
                                double f = fac(igem, alpha, ip, it, is, iu);
                                a -= twono[naddr3(ip, it, is, iu)] * frdm2(&R, it, iu, iq, ir) * f;
                            }
                    }
                }
                /* T6 */
                if (occ[ip] * occ[is] != 0.0) {
                    if (occ[ip] == 1.0 || occ[is] == 1.0) {
                        a -= occ[ip] * occ[is] * auxpqrs;
                    } else {
                        for (int tt = 0; tt < nact; tt++)
                            for (int uu = 0; uu < nact; uu++) {
                                int it = ind1[tt], iu = ind1[uu];
                                double f = fac(igem, alpha, it, iq, iu, ir);
                                a -= twono[naddr3(it, iq, iu, ir)] * frdm2(&R, is, ip, iu, it) * f;
                            }
                    }
                }
                if (is == iq) a -= 0.5 * wmat[ip + (size_t)ir * nb];
                if (ip == ir) a -= 0.5 * wmat[iq + (size_t)is * nb];
                const size_t at = irow + (size_t)icol * ndimx;
                if (ir > is && ip > iq) {
                    abplus[at] += a;
                    abmin[at] += a;
                }
                if (ir > is && iq > ip) {
                    abplus[at] += a;
                    abmin[at] -= a;
                }
            }
        }
    }

    /* C scaling (interpa.f:4079-4097) */
    for (int irow = 0; irow < ndimx; irow++) {
        const int ir = indn[2 * irow] - 1, is = indn[2 * irow + 1] - 1;
        for (int icol = 0; icol < ndimx; icol++) {
            const int ip = indn[2 * icol] - 1, iq = indn[2 * icol + 1] - 1;
            const size_t at = irow + (size_t)icol * ndimx;
            if ((c[ip] + c[iq]) * (c[ir] + c[is]) != 0.0) abplus[at] = abplus[at] / (c[ip] + c[iq]) / (c[ir] + c[is]);
            if ((c[ip] - c[iq]) * (c[ir] - c[is]) != 0.0) abmin[at] = abmin[at] / (c[ip] - c[iq]) / (c[ir] - c[is]);
        }
    }
    free(hno); free(auxi); free(auxio); free(wmat); free(c); free(ind1); free(ind2);
    return 0;
}
