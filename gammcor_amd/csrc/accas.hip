// AC-CAS on gfx950: A+-B assembly (AB_CAS, interpa.f:3659-4104), the
// symmetric ERPA solve (ERPASYMM1, interpa.f:1652-1781) batched over alpha
// points, and the alpha-integrand (ACEneERPA, acfd.f:4968-5043).
//
// A+-B depends on alpha only through factors alpha (inter-geminal terms) and
// 1-alpha (intra-geminal one-electron shift), so every element is evaluated
// once for alpha = 0 and alpha = 1 and A(alpha) = (1-alpha) A(0) + alpha A(1)
// is formed on the fly by the solve.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <vector>

#include <hip/hip_fp16.h>

#include "accas.hpp"
#include "chol.hpp"
#include "oa_policy.hpp"
#include "gemm.hpp"
#include "prof.hpp"

namespace gmc {

struct DevSys {
    int N, nI, nA, nO, nd;
    const double* twono;     // packed NO integrals, or nullptr for the J/K classes
    const double* occ;
    const int* igem;
    const double* gam;
    const int* indn;
    const double* J;         // (k l|i j), k, l occupied (fofo.hpp)
    const double* K;         // (f o|f' o'), o, o' occupied
};

// (pq|rs) from the packed integrals, or from the J/K classes: two occupied
// indices in one pair -> J, one occupied index in each pair -> K.  Other
// integrals are never needed with a non-zero factor; they read as 0.
__device__ inline double tw(const DevSys& s, int a, int b, int c, int d)
{
    if (s.twono) return s.twono[pidx(pidx(a, b), pidx(c, d))];
    const int nO = s.nO, N = s.N;
    if (a < nO && b < nO) return s.J[(size_t)(a + nO * b) * N * N + c + (size_t)N * d];
    if (c < nO && d < nO) return s.J[(size_t)(c + nO * d) * N * N + a + (size_t)N * b];
    int f1, o1, f2, o2;
    if (b < nO) { o1 = b; f1 = a; } else if (a < nO) { o1 = a; f1 = b; } else return 0.0;
    if (d < nO) { o2 = d; f2 = c; } else if (c < nO) { o2 = c; f2 = d; } else return 0.0;
    return s.K[(size_t)(f1 + (size_t)N * o1) * N * nO + f2 + (size_t)N * o2];
}

__device__ inline bool same4(const int* g, int i, int j, int k, int l)
{
    return g[i] == g[j] && g[j] == g[k] && g[k] == g[l];
}

// FRDM2 (interpa.f:5244-5272)
__device__ inline double frdm2(const DevSys& s, int p, int q, int r, int s_)
{
    const double* occ = s.occ;
    double v = 0.0;
    if (p == r && q == s_ && (occ[p] == 1.0 || occ[q] == 1.0)) v += 2.0 * occ[p] * occ[q];
    if (p == s_ && q == r && (occ[p] == 1.0 || occ[q] == 1.0)) v -= occ[p] * occ[q];
    const int a = p - s.nI, b = q - s.nI, c = r - s.nI, d = s_ - s.nI;
    if (a >= 0 && a < s.nA && b >= 0 && b < s.nA && c >= 0 && c < s.nA && d >= 0 && d < s.nA)
        v = s.gam[((a * s.nA + b) * s.nA + c) * s.nA + d];
    return v;
}

// ------------------------------------------------------------ small kernels
__global__ void unpack_sym_kernel(const double* packed, double* M, int N)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y;
    if (i < N) M[i + (size_t)j * N] = packed[pidx(i, j)];
}

__global__ void cicoef_kernel(const double* occ, double* c, int N)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < N) {
        double v = sqrt(occ[i]);
        c[i] = occ[i] < 0.5 ? -v : v;
    }
}

// HNO alpha parts and AuxI/AuxIO at alpha = 0 and 1, one thread per pair p>=q.
// h packed arrays: index pidx(p,q).
__global__ void ab_aux_kernel(DevSys s, const double* hno, double* h0, double* h1, double* ai0, double* ai1,
                              double* aio0, double* aio1)
{
    const int64_t T = (int64_t)s.N * (s.N + 1) / 2;
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= T) return;
    int p = (int)((sqrt(8.0 * (double)idx + 1.0) - 1.0) * 0.5);
    while ((int64_t)(p + 1) * (p + 2) / 2 <= idx) p++;
    while ((int64_t)p * (p + 1) / 2 > idx) p--;
    const int q = (int)(idx - (int64_t)p * (p + 1) / 2);
    const double h = hno[p + (size_t)q * s.N];
    const int* g = s.igem;
    // alpha one-electron part (interpa.f:3785-3810), element (I,J) = (p,q)
    if (g[p] != g[q]) {
        h0[idx] = 0.0;
        h1[idx] = h;
    } else {
        double aux = 0.0;
        for (int t = 0; t < s.N; t++)
            if (g[t] != g[p]) aux += s.occ[t] * (2.0 * tw(s, t, t, p, q) - tw(s, t, p, t, q));
        h0[idx] = h + aux;
        h1[idx] = h;
    }
    // AuxI, AuxIO (interpa.f:3843-3858)
    double a0 = 0, a1 = 0, o0 = 0, o1 = 0;
    for (int t = 0; t < s.nO; t++) {
        const double v = s.occ[t] * (2.0 * tw(s, p, q, t, t) - tw(s, p, t, q, t));
        const bool same = g[t] == g[p] && g[p] == g[q];
        a1 += v;
        if (same) a0 += v;
        if (t < s.nI) {
            o1 += v;
            if (same) o0 += v;
        }
    }
    ai0[idx] = a0;
    ai1[idx] = a1;
    aio0[idx] = o0;
    aio1[idx] = o1;
}

// WMAT (interpa.f:3862-3890): one wave per (p, r), r < nO.  ld = N.
__global__ __launch_bounds__(256) void ab_wmat_kernel(DevSys s, double* w0, double* w1)
{
    const int lane = threadIdx.x & 63;
    const int64_t wid = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    if (wid >= (int64_t)s.N * s.nO) return;
    const int p = (int)(wid % s.N), r = (int)(wid / s.N);
    const int nO = s.nO;
    const int tot = nO * nO * nO;
    double a0 = 0, a1 = 0;
    for (int x = lane; x < tot; x += 64) {
        const int t = x / (nO * nO), w = (x / nO) % nO, u = x % nO;
        const double v = tw(s, t, w, p, u) * frdm2(s, w, u, t, r) + tw(s, t, u, p, w) * frdm2(s, w, u, r, t);
        a1 += v;
        if (same4(s.igem, t, w, p, u)) a0 += v;
    }
    for (int o = 32; o > 0; o >>= 1) {
        a0 += __shfl_xor(a0, o, 64);
        a1 += __shfl_xor(a1, o, 64);
    }
    if (lane == 0) {
        w0[p + (size_t)r * s.N] = a0;
        w1[p + (size_t)r * s.N] = a1;
    }
}

// ECASSCF (interpa.f:3755-3771) with HNO before the alpha modification:
// ECAS_BLOCKS workgroups each sum a strided share of the one- and
// two-electron terms into part[blockIdx.x]; ab_ecas_sum_kernel adds the
// partials in block order (the same bits on every run).
constexpr int ECAS_BLOCKS = 128;
__global__ __launch_bounds__(256) void ab_ecas_kernel(DevSys s, const double* hno, double* part)
{
    __shared__ double red[4];
    const int64_t stride = (int64_t)ECAS_BLOCKS * 256, t0 = (int64_t)blockIdx.x * 256 + threadIdx.x;
    double e = 0.0;
    for (int64_t i = t0; i < s.N; i += stride) e += 2.0 * s.occ[i] * hno[i + (size_t)i * s.N];
    const int nO = s.nO;
    const int64_t tot = (int64_t)nO * nO * nO * nO;
    for (int64_t x = t0; x < tot; x += stride) {
        const int p = (int)(x / ((int64_t)nO * nO * nO)), q = (int)((x / (nO * nO)) % nO), r = (int)((x / nO) % nO),
                  ss = (int)(x % nO);
        e += frdm2(s, p, q, r, ss) * tw(s, p, r, q, ss);
    }
    for (int o = 32; o > 0; o >>= 1) e += __shfl_xor(e, o, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = e;
    __syncthreads();
    if (threadIdx.x == 0) part[blockIdx.x] = red[0] + red[1] + red[2] + red[3];
}
__global__ void ab_ecas_sum_kernel(const double* part, double* out)
{
    if (threadIdx.x != 0) return;
    double e = 0.0;
    for (int b = 0; b < ECAS_BLOCKS; b++) e += part[b];
    *out = e;
}

struct AbAux {
    const double *h0, *h1, *ai0, *ai1, *aio0, *aio1, *w0, *w1, *c;
    // The active-space sums of T1..T6 contracted ahead (ab_gather_kernel,
    // ab_gam_kernel and four MFMA products), N^2 x nA^2 each, column-major:
    //   x12[(s + N q) + N^2 (p' + nA r')] = sum_tu (sq|tu) G(pu,rt) + (su|tq) G(pu,tr)
    //   x34[(p + N r) + N^2 (s' + nA q')] = sum_tu (ut|pr) G(st,qu) + (ur|pt) G(st,uq)
    //   x5 [(p + N s) + N^2 (q' + nA r')] = sum_tu (pt|su) G(tu,qr)
    //   x6 [(q + N r) + N^2 (s' + nA p')] = sum_tu (tq|ur) G(sp,ut)
    // (x' = x - nI for an active index; G the active 2-RDM); null without
    // active orbitals
    const double *x12, *x34, *x5, *x6;
};

// The integral blocks the contractions read, one thread per ((a, b), (t, u)),
// a, b over all orbitals, t, u active; each N^2 x nA^2 block column-major
// (row a + N b, column t' + nA u'), with the argument order of the loop
// terms they replace:
//   iab = [tw(a,b,t,u) | tw(a,u,t,b)], icd = [tw(u,t,a,b) | tw(u,b,a,t)],
//   ie = tw(a,t,b,u), if_ = tw(t,a,u,b).
__global__ void ab_gather_kernel(DevSys s, double* iab, double* icd, double* ie, double* if_)
{
    const int N = s.N, nA = s.nA, nI = s.nI, na2 = nA * nA;
    const int64_t NN = (int64_t)N * N;
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= NN * na2) return;
    const int64_t row = idx % NN;
    const int col = (int)(idx / NN);
    const int a = (int)(row % N), b = (int)(row / N);
    const int t = nI + col % nA, u = nI + col / nA;
    iab[row + NN * col] = tw(s, a, b, t, u);
    iab[row + NN * (na2 + col)] = tw(s, a, u, t, b);
    icd[row + NN * col] = tw(s, u, t, a, b);
    icd[row + NN * (na2 + col)] = tw(s, u, b, a, t);
    ie[row + NN * col] = tw(s, a, t, b, u);
    if_[row + NN * col] = tw(s, t, a, u, b);
}

// The 2-RDM arranged for the four products: row c = t' + nA u' (+ nA^2 for
// the second term), column d = x' + nA y'.
//   gab[c][d = p' + nA r'] = G(p,u,r,t) | G(p,u,t,r)
//   gcd[c][d = s' + nA q'] = G(s,t,q,u) | G(s,t,u,q)
//   ge [c][d = q' + nA r'] = G(t,u,q,r)
//   gf [c][d = s' + nA p'] = G(s,p,u,t)
__global__ void ab_gam_kernel(const double* gam, int nA, double* gab, double* gcd, double* ge, double* gf)
{
    const int na2 = nA * nA;
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= na2 * na2) return;
    const int c = idx % na2, d = idx / na2;
    const int t = c % nA, u = c / nA, x = d % nA, y = d / nA;
    auto G = [&](int i, int j, int k, int l) { return gam[((i * nA + j) * nA + k) * nA + l]; };
    gab[c + 2 * na2 * d] = G(x, u, y, t);
    gab[na2 + c + 2 * na2 * d] = G(x, u, t, y);
    gcd[c + 2 * na2 * d] = G(x, t, y, u);
    gcd[na2 + c + 2 * na2 * d] = G(x, t, u, y);
    ge[c + na2 * d] = G(t, u, x, y);
    gf[c + na2 * d] = G(x, y, u, t);
}

// A_{rs,pq} of interpa.f:3905-4054 at alpha = 0 (e0) and alpha = 1 (e1).
__device__ inline void arspq(const DevSys& s, const AbAux& x, int ip, int iq, int ir, int is, double& e0,
                             double& e1)
{
    const double* occ = s.occ;
    const int* g = s.igem;
    const int N = s.N;
    const int64_t iqs = pidx(is, iq), ipr = pidx(ir, ip);
    double a0 = 0.0, a1 = 0.0;
    if (ip == ir) {
        a0 += (occ[ip] - occ[is]) * x.h0[iqs];
        a1 += (occ[ip] - occ[is]) * x.h1[iqs];
    }
    if (is == iq) {
        a0 += (occ[iq] - occ[ir]) * x.h0[ipr];
        a1 += (occ[iq] - occ[ir]) * x.h1[ipr];
    }
    const double f0pqrs = same4(g, ip, iq, ir, is) ? 1.0 : 0.0;
    const double xpqrs = 2.0 * tw(s, ip, iq, ir, is) - tw(s, ip, ir, iq, is);
    const int nI = s.nI, nA = s.nA;
    const int64_t NN = (int64_t)N * N;
    // an active orbital: its 2-RDM entries are the active 2-RDM's, and the
    // term's sum over t, u is a table entry (an inactive or virtual orbital
    // that reaches a loop, e.g. with an occupation just below 1, sums it here)
    auto act = [&](int o) { return o >= nI && o < nI + nA; };
    // T1+T2
    if (occ[ip] * occ[ir] != 0.0) {
        if (occ[ip] == 1.0 || occ[ir] == 1.0) {
            const double v = occ[ip] * occ[ir] * xpqrs;
            a0 += f0pqrs * v;
            a1 += v;
            if (ip == ir) {
                a0 += occ[ip] * x.ai0[iqs];
                a1 += occ[ip] * x.ai1[iqs];
            }
        } else {
            double sum = 0.0;
            if (x.x12 && act(ip) && act(ir)) {
                sum = x.x12[(is + (int64_t)N * iq) + NN * ((ip - nI) + nA * (ir - nI))];
            } else {
                for (int tt = 0; tt < nA; tt++)
                    for (int uu = 0; uu < nA; uu++) {
                        const int it = nI + tt, iu = nI + uu;
                        sum += tw(s, is, iq, it, iu) * frdm2(s, ip, iu, ir, it) +
                               tw(s, is, iu, it, iq) * frdm2(s, ip, iu, it, ir);
                    }
            }
            if (same4(g, is, iq, nI, nI)) a0 += sum;
            a1 += sum;
            if (ip == ir) {
                a0 += occ[ip] * x.aio0[iqs];
                a1 += occ[ip] * x.aio1[iqs];
            }
        }
    }
    // T3+T4
    if (occ[iq] * occ[is] != 0.0) {
        if (occ[iq] == 1.0 || occ[is] == 1.0) {
            const double v = occ[iq] * occ[is] * xpqrs;
            a0 += f0pqrs * v;
            a1 += v;
            if (iq == is) {
                a0 += occ[iq] * x.ai0[ipr];
                a1 += occ[iq] * x.ai1[ipr];
            }
        } else {
            double sum = 0.0;
            if (x.x34 && act(iq) && act(is)) {
                sum = x.x34[(ip + (int64_t)N * ir) + NN * ((is - nI) + nA * (iq - nI))];
            } else {
                for (int tt = 0; tt < nA; tt++)
                    for (int uu = 0; uu < nA; uu++) {
                        const int it = nI + tt, iu = nI + uu;
                        sum += tw(s, iu, it, ip, ir) * frdm2(s, is, it, iq, iu) +
                               tw(s, iu, ir, ip, it) * frdm2(s, is, it, iu, iq);
                    }
            }
            if (same4(g, nI, nI, ip, ir)) a0 += sum;
            a1 += sum;
            if (iq == is) {
                a0 += occ[iq] * x.aio0[ipr];
                a1 += occ[iq] * x.aio1[ipr];
            }
        }
    }
    // T5
    if (occ[ir] * occ[iq] != 0.0) {
        if (occ[iq] == 1.0 || occ[ir] == 1.0) {
            const double v = occ[iq] * occ[ir] * xpqrs;
            a0 -= f0pqrs * v;
            a1 -= v;
        } else {
            double sum = 0.0;
            if (x.x5 && act(iq) && act(ir)) {
                sum = x.x5[(ip + (int64_t)N * is) + NN * ((iq - nI) + nA * (ir - nI))];
            } else {
                for (int tt = 0; tt < nA; tt++)
                    for (int uu = 0; uu < nA; uu++) {
                        const int it = nI + tt, iu = nI + uu;
                        sum += tw(s, ip, it, is, iu) * frdm2(s, it, iu, iq, ir);
                    }
            }
            if (same4(g, ip, nI, is, nI)) a0 -= sum;
            a1 -= sum;
        }
    }
    // T6
    if (occ[ip] * occ[is] != 0.0) {
        if (occ[ip] == 1.0 || occ[is] == 1.0) {
            const double v = occ[ip] * occ[is] * xpqrs;
            a0 -= f0pqrs * v;
            a1 -= v;
        } else {
            double sum = 0.0;
            if (x.x6 && act(ip) && act(is)) {
                sum = x.x6[(iq + (int64_t)N * ir) + NN * ((is - nI) + nA * (ip - nI))];
            } else {
                for (int tt = 0; tt < nA; tt++)
                    for (int uu = 0; uu < nA; uu++) {
                        const int it = nI + tt, iu = nI + uu;
                        sum += tw(s, it, iq, iu, ir) * frdm2(s, is, ip, iu, it);
                    }
            }
            if (same4(g, nI, iq, nI, ir)) a0 -= sum;
            a1 -= sum;
        }
    }
    // WMAT(p, r) is stored for r < nO only; the reference's N x N array is
    // zero beyond that (interpa.f:3866-3870)
    if (is == iq && ir < s.nO) {
        a0 -= 0.5 * x.w0[ip + (size_t)ir * N];
        a1 -= 0.5 * x.w1[ip + (size_t)ir * N];
    }
    if (ip == ir && is < s.nO) {
        a0 -= 0.5 * x.w0[iq + (size_t)is * N];
        a1 -= 0.5 * x.w1[iq + (size_t)is * N];
    }
    e0 = a0;
    e1 = a1;
}

// Main A+-B loop (interpa.f:3894-4075) and C scaling (4079-4097).
__global__ __launch_bounds__(256) void ab_main_kernel(DevSys s, AbAux x, double* p0, double* m0, double* p1,
                                                      double* m1)
{
    const int irow = blockIdx.x * blockDim.x + threadIdx.x;
    const int icol = blockIdx.y;
    const int nd = s.nd;
    if (irow >= nd) return;
    const int ir = s.indn[2 * irow] - 1, is = s.indn[2 * irow + 1] - 1;
    const int ipp = s.indn[2 * icol] - 1, iqq = s.indn[2 * icol + 1] - 1;
    double P0 = 0, P1 = 0, M0 = 0, M1 = 0;
    for (int o = 0; o < 2; o++) {
        const int ip = o == 0 ? iqq : ipp, iq = o == 0 ? ipp : iqq;
        double e0, e1;
        arspq(s, x, ip, iq, ir, is, e0, e1);
        P0 += e0;
        P1 += e1;
        if (ir > is && ip > iq) {
            M0 += e0;
            M1 += e1;
        } else if (ir > is && iq > ip) {
            M0 -= e0;
            M1 -= e1;
        }
    }
    const double* c = x.c;
    const double cp = (c[ipp] + c[iqq]) * (c[ir] + c[is]);
    const double cm = (c[ipp] - c[iqq]) * (c[ir] - c[is]);
    if (cp != 0.0) {
        P0 = P0 / (c[ipp] + c[iqq]) / (c[ir] + c[is]);
        P1 = P1 / (c[ipp] + c[iqq]) / (c[ir] + c[is]);
    }
    if (cm != 0.0) {
        M0 = M0 / (c[ipp] - c[iqq]) / (c[ir] - c[is]);
        M1 = M1 / (c[ipp] - c[iqq]) / (c[ir] - c[is]);
    }
    const size_t at = irow + (size_t)icol * nd;
    p0[at] = P0;
    m0[at] = M0;
    p1[at] = P1;
    m1[at] = M1;
}

static DevSys devsys(const AcSystem& a)
{
    return DevSys{a.N, a.nI, a.nA, a.nI + a.nA, a.ndimx, a.twono, a.occ, a.igem, a.gam, a.indn, a.J, a.K};
}

size_t ab_scratch_bytes(const AcSystem& s)
{
    const size_t N = s.N, T = N * (N + 1) / 2, nO = s.nI + s.nA;
    const size_t na2 = (size_t)s.nA * s.nA, xs = N * N * na2 * 8;
    return carve_bytes({N * N * 8, N * N * 8, N * N * 8, T * 8, T * 8, T * 8, T * 8, T * 8, T * 8, N * nO * 8,
                        N * nO * 8, 8, ECAS_BLOCKS * 8, 2 * xs, 2 * xs, xs, xs, xs, xs, xs, xs,
                        6 * na2 * na2 * 8});
}

int assemble_ab01(hipStream_t st, const AcSystem& a, void* scratch, AbPair& out)
{
    GMC_CHECK(a.N > 0 && a.ndimx > 0, "AB_CAS: empty system");
    const int N = a.N;
    const size_t T = (size_t)N * (N + 1) / 2, nO = a.nI + a.nA;
    Arena ar{static_cast<char*>(scratch), ab_scratch_bytes(a)};
    double* H = ar.take<double>((size_t)N * N);
    double* UH = ar.take<double>((size_t)N * N);
    double* hno = ar.take<double>((size_t)N * N);
    double* h0 = ar.take<double>(T);
    double* h1 = ar.take<double>(T);
    double* ai0 = ar.take<double>(T);
    double* ai1 = ar.take<double>(T);
    double* aio0 = ar.take<double>(T);
    double* aio1 = ar.take<double>(T);
    double* w0 = ar.take<double>((size_t)N * nO);
    double* w1 = ar.take<double>((size_t)N * nO);
    double* d_e = ar.take<double>(1);
    double* d_epart = ar.take<double>(ECAS_BLOCKS);
    const int nA = a.nA, na2 = nA * nA;
    const size_t xs = (size_t)N * N * na2;
    double* iab = ar.take<double>(2 * xs);
    double* icd = ar.take<double>(2 * xs);
    double* ie = ar.take<double>(xs);
    double* if_ = ar.take<double>(xs);
    double* x12 = ar.take<double>(xs);
    double* x34 = ar.take<double>(xs);
    double* x5 = ar.take<double>(xs);
    double* x6 = ar.take<double>(xs);
    double* gab = ar.take<double>(6 * (size_t)na2 * na2);
    double* gcd = gab + 2 * (size_t)na2 * na2;
    double* ge = gcd + 2 * (size_t)na2 * na2;
    double* gf = ge + (size_t)na2 * na2;
    GMC_CHECK(ar.fits(), "AB_CAS: scratch layout overflow");
    DevSys s = devsys(a);

    // HNO = U h U^T (interpa.f:3702-3716)
    hipLaunchKernelGGL(unpack_sym_kernel, dim3((N + 255) / 256, N), dim3(256), 0, st, a.xone, a.ure ? H : hno, N);
    GMC_HIP(hipGetLastError());
    if (a.ure) {
        GemmShape g1{N, N, N, false, false, N, N, N, 0, 0, 0, 1, 0};
        GMC_TRY(gemm<double>(st, g1, 1.0, a.ure, H, 0.0, UH));
        GemmShape g2{N, N, N, false, true, N, N, N, 0, 0, 0, 1, 0};
        GMC_TRY(gemm<double>(st, g2, 1.0, UH, a.ure, 0.0, hno));
    }
    hipLaunchKernelGGL(ab_ecas_kernel, dim3(ECAS_BLOCKS), dim3(256), 0, st, s, hno, d_epart);
    hipLaunchKernelGGL(ab_ecas_sum_kernel, dim3(1), dim3(64), 0, st, d_epart, d_e);
    hipLaunchKernelGGL(cicoef_kernel, dim3((N + 255) / 256), dim3(256), 0, st, a.occ, out.cicoef, N);
    hipLaunchKernelGGL(ab_aux_kernel, dim3((unsigned)((T + 255) / 256)), dim3(256), 0, st, s, hno, h0, h1, ai0, ai1,
                       aio0, aio1);
    if (nO > 0)
        hipLaunchKernelGGL(ab_wmat_kernel, dim3((unsigned)(((int64_t)N * nO * 64 + 255) / 256)), dim3(256), 0, st, s,
                           w0, w1);
    GMC_HIP(hipGetLastError());
    AbAux x{h0, h1, ai0, ai1, aio0, aio1, w0, w1, out.cicoef, nullptr, nullptr, nullptr, nullptr};
    if (na2 > 0) {
        // the active-space sums as products over (t, u): the integral blocks
        // and the 2-RDM arrangements, then X = I G on the FP64 MFMA tile
        // (algorithmic bytes: the ten N^2 x nA^2 tables written once and the
        // six integral blocks read once by the products)
        const int64_t NN = (int64_t)N * N;
        ProfScope ps(st, "ab_cas_tables", 16.0 * 8.0 * (double)xs);
        GMC_CHECK(NN <= INT32_MAX && NN * 2 * na2 < ((int64_t)1 << 40), "AB_CAS: N=%d, nA=%d too large", N, nA);
        hipLaunchKernelGGL(ab_gather_kernel, dim3((unsigned)((NN * na2 + 255) / 256)), dim3(256), 0, st, s, iab, icd,
                           ie, if_);
        hipLaunchKernelGGL(ab_gam_kernel, dim3((na2 * na2 + 255) / 256), dim3(256), 0, st, a.gam, nA, gab, gcd, ge,
                           gf);
        GMC_HIP(hipGetLastError());
        const int m = (int)NN;
        GemmShape g2{m, na2, 2 * na2, false, false, NN, 2 * na2, NN, 0, 0, 0, 1, 0};
        GMC_TRY(gemm<double>(st, g2, 1.0, iab, gab, 0.0, x12));
        GMC_TRY(gemm<double>(st, g2, 1.0, icd, gcd, 0.0, x34));
        GemmShape g1{m, na2, na2, false, false, NN, na2, NN, 0, 0, 0, 1, 0};
        GMC_TRY(gemm<double>(st, g1, 1.0, ie, ge, 0.0, x5));
        GMC_TRY(gemm<double>(st, g1, 1.0, if_, gf, 0.0, x6));
        x.x12 = x12;
        x.x34 = x34;
        x.x5 = x5;
        x.x6 = x6;
    }
    GMC_CHECK(a.ndimx <= 65535, "AB_CAS: NDimX=%d exceeds the grid y limit", a.ndimx);
    {
        // algorithmic bytes: the four ndimx^2 output matrices written once
        ProfScope ps(st, "ab_cas_main", 4.0 * 8.0 * (double)a.ndimx * a.ndimx);
        hipLaunchKernelGGL(ab_main_kernel, dim3((a.ndimx + 255) / 256, a.ndimx), dim3(256), 0, st, s, x, out.p0,
                           out.m0, out.p1, out.m1);
    }
    GMC_HIP(hipGetLastError());
    GMC_HIP(hipMemcpyAsync(&out.ecasscf, d_e, sizeof(double), hipMemcpyDeviceToHost, st));
    GMC_HIP(hipStreamSynchronize(st));
    return 0;
}

// ------------------------------------------------------------ integrand weights
__global__ void gw_kernel(DevSys s, const double* c, double* Gw)
{
    const int I = blockIdx.x * blockDim.x + threadIdx.x, J = blockIdx.y;
    if (I >= s.nd) return;
    const int p = s.indn[2 * I] - 1, r = s.indn[2 * I + 1] - 1;
    const int q = s.indn[2 * J] - 1, ss = s.indn[2 * J + 1] - 1;
    const int* g = s.igem;
    double v = 0.0;
    if (!(g[p] == g[r] && g[q] == g[ss] && g[p] == g[q]))
        v = 2.0 * (c[ss] + c[q]) * (c[p] + c[r]) * tw(s, p, r, q, ss);
    Gw[I + (size_t)J * s.nd] = v;
}

__global__ __launch_bounds__(256) void wdiag_kernel(DevSys s, double* out)
{
    __shared__ double red[4];
    double e = 0.0;
    const int* g = s.igem;
    for (int I = threadIdx.x; I < s.nd; I += 256) {
        const int p = s.indn[2 * I] - 1, r = s.indn[2 * I + 1] - 1;
        if (!(g[p] == g[r])) {
            const double np = s.occ[p], nr = s.occ[r];
            e -= (np * (1.0 - nr) + nr * (1.0 - np)) * tw(s, p, r, p, r);
        }
    }
    for (int o = 32; o > 0; o >>= 1) e += __shfl_xor(e, o, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = e;
    __syncthreads();
    if (threadIdx.x == 0) *out = red[0] + red[1] + red[2] + red[3];
}

int integrand_weights(hipStream_t st, const AcSystem& a, const double* cicoef, double* Gw, double* d_wdiag)
{
    DevSys s = devsys(a);
    hipLaunchKernelGGL(gw_kernel, dim3((a.ndimx + 255) / 256, a.ndimx), dim3(256), 0, st, s, cicoef, Gw);
    hipLaunchKernelGGL(wdiag_kernel, dim3(1), dim3(256), 0, st, s, d_wdiag);
    GMC_HIP(hipGetLastError());
    return 0;
}

// ------------------------------------------------------------ ERPA
// A+(alpha_b) and A-(alpha_b) of every alpha point: the symmetric parts of
// A(0) and A(1) combined affinely (A+-B is affine in alpha; ERPASYMM1
// symmetrises, interpa.f:1680-1687).  One workgroup per 64 x 64 tile on or
// below the diagonal: each base tile and its transposed partner are read
// once and symmetrised in LDS, then every alpha point's tile and its mirror
// are written along columns.
template <class T>
__global__ __launch_bounds__(256) void form_kernel(const double* p0, const double* p1, const double* m0,
                                                   const double* m1, const double* alphas, int nalpha, T* Ap, T* Am,
                                                   int n)
{
    __shared__ double s0[64][65], s1[64][65], tmp[64][65];
    int ti, tj;
    lower_tile(blockIdx.x, ti, tj);
    const int i0 = ti * 64, j0 = tj * 64, tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
    // dst[y][x] = sym(X)(i0 + x, j0 + y)
    auto load_sym = [&](const double* X, double (*dst)[65]) {
        for (int y = ty; y < 64; y += 4) {
            const int i = i0 + tx, j = j0 + y;
            dst[y][tx] = (i < n && j < n) ? X[i + (size_t)j * n] : 0.0;
            const int i2 = j0 + tx, j2 = i0 + y;
            tmp[y][tx] = (i2 < n && j2 < n) ? X[i2 + (size_t)j2 * n] : 0.0;   // X(j0 + x, i0 + y)
        }
        __syncthreads();
        for (int y = ty; y < 64; y += 4) dst[y][tx] = 0.5 * (dst[y][tx] + tmp[tx][y]);
        __syncthreads();
    };
    for (int which = 0; which < 2; which++) {
        load_sym(which ? m0 : p0, s0);
        load_sym(which ? m1 : p1, s1);
        T* out = which ? Am : Ap;
        for (int b = 0; b < nalpha; b++) {
            const double al = alphas[b];
            T* o = out + (size_t)b * n * n;
            for (int y = ty; y < 64; y += 4) {
                const int i = i0 + tx, j = j0 + y;
                if (i < n && j < n) o[i + (size_t)j * n] = (T)((1.0 - al) * s0[y][tx] + al * s1[y][tx]);
                const int i2 = j0 + tx, j2 = i0 + y;
                if (ti != tj && i2 < n && j2 < n) o[i2 + (size_t)j2 * n] = (T)((1.0 - al) * s0[tx][y] + al * s1[tx][y]);
            }
        }
        __syncthreads();
    }
}

// The batching form: A+(alpha) and A-(alpha) of system k of a batch,
// padded to the batch order npad.  Within the system's order n the matrices
// are those of form_kernel; beyond it they are zero.  The padded modes stay
// decoupled through every step: their reflector tails are zero (tau = 0),
// their divide & conquer weights z are exactly zero so they deflate before
// any close-pair rotation, and with eigenvalue 0 they leave the deflation
// tolerance (8 eps max(|d|, |z|) of a merge) as it is without them - a pad
// above the spectrum would raise it; S, M and Y are zero on them, and the
// integrand reads only the system's own block.  One workgroup per 64 x 64
// tile on or below the diagonal and system (blockIdx.y), writing the tile of
// every alpha point.
template <class T>
__global__ __launch_bounds__(256) void form_pad_kernel(const BatchSys* sys, const double* alphas, int nalpha,
                                                       T* Ap, T* Am, int npad)
{
    __shared__ double s0[64][65], s1[64][65], tmp[64][65];
    int ti, tj;
    lower_tile(blockIdx.x, ti, tj);
    const BatchSys sy = sys[blockIdx.y];
    const int n = sy.n;
    const int i0 = ti * 64, j0 = tj * 64, tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
    auto load_sym = [&](const double* X, double (*dst)[65]) {
        for (int y = ty; y < 64; y += 4) {
            const int i = i0 + tx, j = j0 + y;
            dst[y][tx] = (i < n && j < n) ? X[i + (size_t)j * n] : 0.0;
            const int i2 = j0 + tx, j2 = i0 + y;
            tmp[y][tx] = (i2 < n && j2 < n) ? X[i2 + (size_t)j2 * n] : 0.0;
        }
        __syncthreads();
        for (int y = ty; y < 64; y += 4) dst[y][tx] = 0.5 * (dst[y][tx] + tmp[tx][y]);
        __syncthreads();
    };
    for (int which = 0; which < 2; which++) {
        load_sym(which ? sy.m0 : sy.p0, s0);
        load_sym(which ? sy.m1 : sy.p1, s1);
        T* out = which ? Am : Ap;
        for (int a = 0; a < nalpha; a++) {
            const double al = alphas[a];
            T* o = out + ((size_t)blockIdx.y * nalpha + a) * npad * npad;
            for (int y = ty; y < 64; y += 4) {
                const int i = i0 + tx, j = j0 + y;
                if (i < npad && j < npad) {
                    const double v = (i < n && j < n) ? (1.0 - al) * s0[y][tx] + al * s1[y][tx] : 0.0;
                    o[i + (size_t)j * npad] = (T)v;
                }
                const int i2 = j0 + tx, j2 = i0 + y;
                if (ti != tj && i2 < npad && j2 < npad) {
                    const double v = (i2 < n && j2 < n) ? (1.0 - al) * s0[tx][y] + al * s1[tx][y] : 0.0;
                    o[i2 + (size_t)j2 * npad] = (T)v;
                }
            }
        }
        __syncthreads();
    }
}

// W of the batching form: problem b = k nalpha + a, system k's weights over
// its own order n (P is padded to npad); diag + 2 x strictly lower.
template <class T>
__global__ __launch_bounds__(256) void dot_batch_kernel(const BatchSys* sys, int nalpha, const T* P, int npad,
                                                        double* partial)
{
    __shared__ double red[4];
    const int b = blockIdx.y;
    const BatchSys sy = sys[b / nalpha];
    const int n = sy.n;
    const T* pm = P + (size_t)b * npad * npad;
    const int64_t total = (int64_t)n * n;
    double acc = 0.0;
    for (int64_t idx = (int64_t)blockIdx.x * 256 + threadIdx.x; idx < total; idx += (int64_t)gridDim.x * 256) {
        const int i = (int)(idx % n), j = (int)(idx / n);
        const double g = sy.gw[idx], pv = (double)pm[i + (size_t)j * npad];
        if (i > j) acc += 2.0 * g * pv;
        else if (i == j) acc += g * pv;
    }
    for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) partial[(size_t)b * gridDim.x + blockIdx.x] = red[0] + red[1] + red[2] + red[3];
}

__global__ void sum_partials_batch_kernel(const double* partial, int np, const BatchSys* sys, int nalpha, double* W,
                                          int nb)
{
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nb) return;
    double s = 0.0;
    for (int i = 0; i < np; i++) s += partial[(size_t)b * np + i];
    W[b] = s + sys[b / nalpha].wdiag;
}

// x[0..n) = 0 (a kernel, not a memset: the solve sequence may be captured)
__global__ void zero_int_kernel(int* x, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) x[i] = 0;
}

// V' = V |lambda|^(1/4) column-wise; count negative eigenvalues of A+.
template <class T>
__global__ void sqrt_scale_kernel(T* V, const T* lam, int n, int* nneg)
{
    const int b = blockIdx.z, k = blockIdx.y;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const T l = lam[(size_t)b * n + k];
    const T f = sqrt(sqrt(fabs(l)));
    if (i < n) V[(size_t)b * n * n + i + (size_t)k * n] *= f;
    if (i == 0 && l < T(0)) atomicAdd(&nneg[b], 1);
}

// M_ji <- M_ij for i > j: the strict upper triangle from the lower one, in
// 64 x 64 tiles through LDS so reads and writes are both contiguous.  Only
// tiles on or below the diagonal are read and only the transposed tile is
// written, so the workgroups do not overlap.
template <class T>
__global__ __launch_bounds__(256) void mirror_lower_kernel(T* M, int n)
{
    __shared__ T tile[64][65];
    const int ti = blockIdx.x, tj = blockIdx.y;
    if (ti < tj) return;
    T* m = M + (size_t)blockIdx.z * n * n;
    const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
    const int i0 = ti * 64, j0 = tj * 64;
    for (int y = ty; y < 64; y += 4) {
        const int i = i0 + tx, j = j0 + y;
        if (i < n && j < n) tile[y][tx] = m[i + (size_t)j * n];
    }
    __syncthreads();
    for (int y = ty; y < 64; y += 4) {
        const int i = i0 + y, j = j0 + tx;
        if (i < n && j < n && i > j) m[j + (size_t)i * n] = tile[tx][y];
    }
}

// Normalisation of interpa.f:1748-1776 (2/omega Y^T A- Y = 2 omega for Y = S z)
// and the filter of acfd.f:4976, 5013.  Yk <- Y scaled, Yf <- kept columns only.
template <class T>
__global__ void ycols_kernel(T* Y, T* Yf, const T* om2, double* eig, int n)
{
    const int b = blockIdx.z, k = blockIdx.y;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const double w2 = (double)om2[(size_t)b * n + k];
    double e = w2, sc = 1.0;
    if (w2 > 1.0e-6) {
        e = sqrt(w2);
        sc = 1.0 / sqrt(2.0 * e);
    }
    const bool keep = e > 1.0e-3 && e < 1.0e8;
    if (i < n) {
        const size_t at = (size_t)b * n * n + i + (size_t)k * n;
        const T y = (T)((double)Y[at] * sc);
        Y[at] = y;
        Yf[at] = keep ? y : T(0);
    }
    if (i == 0) eig[(size_t)b * n + k] = e;
}

// W = sum_IJ Gw_IJ P_IJ with P lower-triangular valid: diag + 2 * strictly lower.
template <class T>
__global__ __launch_bounds__(256) void dot_lower_kernel(const double* Gw, const T* P, int n, double* partial)
{
    __shared__ double red[4];
    const int b = blockIdx.y;
    const T* pm = P + (size_t)b * n * n;
    const int64_t total = (int64_t)n * n;
    double acc = 0.0;
    for (int64_t idx = (int64_t)blockIdx.x * 256 + threadIdx.x; idx < total; idx += (int64_t)gridDim.x * 256) {
        const int i = (int)(idx % n), j = (int)(idx / n);
        if (i > j) acc += 2.0 * Gw[idx] * (double)pm[idx];
        else if (i == j) acc += Gw[idx] * (double)pm[idx];
    }
    for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) partial[(size_t)b * gridDim.x + blockIdx.x] = red[0] + red[1] + red[2] + red[3];
}

__global__ void sum_partials_kernel(const double* partial, int np, const double* wdiag, double* W, int nb)
{
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nb) return;
    double s = 0.0;
    for (int i = 0; i < np; i++) s += partial[(size_t)b * np + i];
    W[b] = s + *wdiag;
}

template <class T>
int erpa_plan(int n, int batch, ErpaPlan<T>* p)
{
    p->n = n;
    p->batch = batch;
    const size_t nn = (size_t)n * n;
    const size_t mat = batch * nn * sizeof(T), vec = batch * n * sizeof(T);
    p->ws_bytes = carve_bytes({mat, mat, mat, mat, vec, vec, batch * sizeof(double),
                               (size_t)batch * 1024 * sizeof(double)});
    return 0;
}

template <class T>
int erpa_batched(hipStream_t st, ErpaPlan<T>& p, const AbPair& ab, const double* alphas, int nalpha,
                 const double* Gw, const double* wdiag, void* scratch, ErpaOut& out, double* Y_out, double* S_out,
                 double* Am_out, const ErpaBatchTables* bt)
{
    GMC_CHECK(nalpha >= 1 && nalpha <= p.batch, "ERPA: %d alpha points for a plan of %d", nalpha, p.batch);
    GMC_CHECK(p.eig && p.eig->bound && p.eig->batch == p.batch && p.eig->n == p.n, "ERPA: eigensolver plan unbound");
    const int n = p.n, B = nalpha;
    const size_t nn = (size_t)n * n;
    Arena ar{static_cast<char*>(scratch), p.ws_bytes};
    T* E = ar.take<T>(p.batch * nn);
    T* Mb = ar.take<T>(p.batch * nn);
    T* Sb = ar.take<T>(p.batch * nn);
    T* Tb = ar.take<T>(p.batch * nn);
    T* lam = ar.take<T>(p.batch * n);
    T* om2 = ar.take<T>(p.batch * n);
    double* d_al = ar.take<double>(p.batch);
    const int npart = 1024;
    double* partial = ar.take<double>((size_t)p.batch * npart);
    GMC_CHECK(ar.fits(), "ERPA: scratch layout overflow");
    const int nprob = bt ? bt->nsys * nalpha : nalpha;   // problems of this call
    GMC_CHECK(nprob <= p.batch, "ERPA: %d problems for a plan of %d", nprob, p.batch);
    hipLaunchKernelGGL(zero_int_kernel, dim3((nprob + 255) / 256), dim3(256), 0, st, out.nneg, nprob);

    const dim3 blk(256);
    // A+(alpha) -> E, A-(alpha) -> Tb
    const int nt64f = (n + 63) / 64;
    if (bt) {
        // the batching form: alphas are already on the device
        hipLaunchKernelGGL(form_pad_kernel<T>, dim3(nt64f * (nt64f + 1) / 2, bt->nsys), blk, 0, st, bt->sys, alphas,
                           nalpha, E, Tb, n);
    } else {
        GMC_HIP(hipMemcpyAsync(d_al, alphas, nalpha * sizeof(double), hipMemcpyHostToDevice, st));
        hipLaunchKernelGGL(form_kernel<T>, dim3(nt64f * (nt64f + 1) / 2), blk, 0, st, ab.p0, ab.p1, ab.m0, ab.m1,
                           d_al, B, E, Tb, n);
    }
    GMC_HIP(hipGetLastError());
    const int Bp = nprob;   // the solve below runs over every problem
    const dim3 gsq((n + 255) / 256, n, Bp);
    if (Am_out) {
        GMC_CHECK(sizeof(T) == 8, "ERPA: A- output is available from the FP64 solve");
        GMC_HIP(hipMemcpyAsync(Am_out, Tb, Bp * nn * sizeof(T), hipMemcpyDeviceToDevice, st));
    }
    EigPlan<T>& ep = *p.eig;
    GMC_CHECK(ep.batch == Bp, "ERPA: plan batch %d != %d problems", ep.batch, Bp);
    GMC_TRY(syevd_batched<T>(st, ep, E, lam));
    hipLaunchKernelGGL(sqrt_scale_kernel<T>, gsq, blk, 0, st, E, lam, n, out.nneg);
    GMC_HIP(hipGetLastError());
    // S = V|L|^(1/2)V^T
    // S is symmetric: its lower triangle, mirrored
    GemmShape gS{n, n, n, false, true, n, n, n, (int64_t)nn, (int64_t)nn, (int64_t)nn, Bp, 1};
    GMC_TRY(gemm_big<T>(st, gS, T(1), E, E, T(0), Sb));
    const int nt64 = (n + 63) / 64;
    hipLaunchKernelGGL(mirror_lower_kernel<T>, dim3(nt64, nt64, Bp), blk, 0, st, Sb, n);
    if (S_out) {
        GMC_CHECK(sizeof(T) == 8, "ERPA: A+^(1/2) output is available from the FP64 solve");
        GMC_HIP(hipMemcpyAsync(S_out, Sb, Bp * nn * sizeof(T), hipMemcpyDeviceToDevice, st));
    }
    // M = S A- S
    GemmShape gN{n, n, n, false, false, n, n, n, (int64_t)nn, (int64_t)nn, (int64_t)nn, Bp, 0};
    GMC_TRY(gemm_big<T>(st, gN, T(1), Tb, Sb, T(0), Mb));
    GemmShape gNl = gN;
    gNl.uplo = 1;   // M is symmetric: its lower triangle, mirrored
    GMC_TRY(gemm_big<T>(st, gNl, T(1), Sb, Mb, T(0), Tb));
    hipLaunchKernelGGL(mirror_lower_kernel<T>, dim3(nt64, nt64, Bp), blk, 0, st, Tb, n);
    GMC_HIP(hipGetLastError());
    GMC_TRY(syevd_batched<T>(st, ep, Tb, om2));
    // Y = S Z
    GMC_TRY(gemm_big<T>(st, gN, T(1), Sb, Tb, T(0), Mb));
    hipLaunchKernelGGL(ycols_kernel<T>, gsq, blk, 0, st, Mb, E, om2, out.eig, n);
    GMC_HIP(hipGetLastError());
    // P = Yf Yf^T (lower) and W (skipped without integrand weights)
    if (Gw || bt) {
        GemmShape gP{n, n, n, false, true, n, n, n, (int64_t)nn, (int64_t)nn, (int64_t)nn, Bp, 1};
        GMC_TRY(gemm_big<T>(st, gP, T(1), E, E, T(0), Sb));
        if (bt) {
            hipLaunchKernelGGL(dot_batch_kernel<T>, dim3(npart, Bp), blk, 0, st, bt->sys, nalpha, Sb, n, partial);
            hipLaunchKernelGGL(sum_partials_batch_kernel, dim3((Bp + 63) / 64), dim3(64), 0, st, partial, npart,
                               bt->sys, nalpha, out.W, Bp);
        } else {
            hipLaunchKernelGGL(dot_lower_kernel<T>, dim3(npart, B), blk, 0, st, Gw, Sb, n, partial);
            hipLaunchKernelGGL(sum_partials_kernel, dim3(1), dim3(64), 0, st, partial, npart, wdiag, out.W, B);
        }
        GMC_HIP(hipGetLastError());
    }
    if (Y_out) {
        GMC_CHECK(sizeof(T) == 8, "ERPA: Y output is available from the FP64 solve");
        GMC_HIP(hipMemcpyAsync(Y_out, Mb, Bp * nn * sizeof(T), hipMemcpyDeviceToDevice, st));
    }
    return 0;
}


// ------------------------------------------------------------ mixed precision
__global__ void f64_to_f32_kernel(const double* a, float* b, int64_t count)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += (int64_t)gridDim.x * blockDim.x)
        b[i] = (float)a[i];
}
__global__ void f32_to_f64_kernel(const float* a, double* b, int64_t count)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += (int64_t)gridDim.x * blockDim.x)
        b[i] = (double)a[i];
}

// Ogita-Aishima: lambda_i = s_ii / (1 - r_ii) with 1 - r_ii = (X^T X)_ii, and
// the Frobenius norms for the cluster threshold delta.  S and G = X^T X are
// valid on and below the diagonal.
__global__ __launch_bounds__(256) void oa_stats_kernel(const double* S, const double* G, int n, double* lam,
                                                       double* partial)
{
    __shared__ double red[3][4];
    const int b = blockIdx.y;
    const double* s = S + (size_t)b * n * n;
    const double* g = G + (size_t)b * n * n;
    const int64_t total = (int64_t)n * n;
    double a0 = 0, a1 = 0, a2 = 0;
    for (int64_t idx = (int64_t)blockIdx.x * 256 + threadIdx.x; idx < total; idx += (int64_t)gridDim.x * 256) {
        const int i = (int)(idx % n), j = (int)(idx / n);
        if (i > j) {
            a0 += 2.0 * s[idx] * s[idx];
            a1 += 2.0 * g[idx] * g[idx];
        } else if (i == j) {
            const double l = s[idx] / g[idx];
            const double d = s[idx] - l;
            a0 += d * d;
            a1 += (1.0 - g[idx]) * (1.0 - g[idx]);
            a2 = fmax(a2, fabs(l));
            lam[(size_t)b * n + i] = l;
        }
    }
    for (int o = 32; o > 0; o >>= 1) {
        a0 += __shfl_xor(a0, o, 64);
        a1 += __shfl_xor(a1, o, 64);
        a2 = fmax(a2, __shfl_xor(a2, o, 64));
    }
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) {
        red[0][w] = a0;
        red[1][w] = a1;
        red[2][w] = a2;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double* pp = partial + ((size_t)b * gridDim.x + blockIdx.x) * 3;
        pp[0] = red[0][0] + red[0][1] + red[0][2] + red[0][3];
        pp[1] = red[1][0] + red[1][1] + red[1][2] + red[1][3];
        pp[2] = fmax(fmax(red[2][0], red[2][1]), fmax(red[2][2], red[2][3]));
    }
}

__global__ void oa_delta_kernel(const double* partial, int np, double* delta, int nb)
{
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nb) return;
    double s0 = 0, s1 = 0, mx = 0;
    for (int i = 0; i < np; i++) {
        const double* pp = partial + ((size_t)b * np + i) * 3;
        s0 += pp[0];
        s1 += pp[1];
        mx = fmax(mx, pp[2]);
    }
    delta[b] = 64.0 * 2.220446049250313e-16 * mx;
}

// E_ij = (s_ij + l_j r_ij) / (l_j - l_i), r = I - X^T X (Ogita-Aishima).
// The step contracts the error of a pair quadratically with a constant of
// about max|l| / |l_i - l_j|, so a pair whose gap is small against the
// spectrum and whose FP32 start is mixed by more than that constant allows
// does not converge by these corrections; nor does a near-degenerate group of
// three or more by 2x2 rotations computed from one S.  Such pairs are joined
// into clusters (connected components), and each cluster is solved exactly:
// the FP64 eigenvectors of its projected matrix H = S_CC + (l_i + l_j) r_ij / 2
// (off the diagonal; l_i on it) rotate its columns after the update, which
// keeps only the orthogonalising r_ij / 2 within a cluster (Ogita and Aishima,
// "Iterative refinement for symmetric eigenvalue decomposition II: clustered
// eigenvalues", Japan J. Indust. Appl. Math. 36 (2019)).
//
// A pair (i > j) joins when it is close (|l_i - l_j| not above 8 max(|num|)
// or the rounding level delta = 64 eps max|l|), whatever its weights: the
// modes the ERPA filter drops must be resolved among themselves too, or the
// step of a kept mode against them (which takes their block of S as
// diagonal) does not converge.  A separated pair joins when it is mixed by
// |E| >= OA_JOIN and the mix matters: |E| rel >= OA_RHO, rel = |w_i - w_j| /
// max w the weight difference (oa_weight) through which a rotation within
// the pair moves the result.  (To first order the step leaves pair (i, j)
// the error sum_k E_ki E_kj (l_j - l_k) / (l_j - l_i); with E_ki about
// c / (l_k - l_i) from the FP32 start that is about |E_ij| times the largest
// |E_ki| of the column, so once the strongly mixed pairs are solved as
// clusters every other pair contracts by OA_JOIN or better per step.)
// Clusters of up to OA_CMAX columns are solved by a workgroup Jacobi, larger
// ones by the batched FP64 eigensolver.  At most `cap` pairs (<= OA_MAXPAIRS)
// are recorded per problem; npair keeps counting beyond, and a problem with
// more pairs than recorded, or with a cluster of more than half its columns,
// is treated as not converged (oa_policy.hpp).
constexpr int OA_MAXPAIRS = 262144;
constexpr int OA_CMAX = 64;          // largest cluster the workgroup Jacobi solves; larger ones go to the eigensolver
constexpr double OA_JOIN = 1.0 / 64;   // |E| from which a separated pair joins a cluster
constexpr double OA_RHO = 1e-7;        // |E| rel from which such a pair matters to the result
constexpr int OA_EMAX_SLOTS = 64;    // per problem, spreads the atomics of the max
constexpr int OA_ET = 32;            // tile of the E kernel
constexpr float OA_TOL = 1e-8f;      // a step whose corrections are all below this ends the refinement

// The weight of eigenvalue l in what the solve makes of the eigenvectors:
// wmode 1, sqrt|l| (S = V |L|^(1/2) V^T of A+); wmode 2, the ERPA integrand's
// 1 / (2 omega) for a mode the filter keeps (omega^2 > 1e-6 and 1e-3 < omega
// < 1e8, interpa.f:1748-1776, acfd.f:4976, 5013) and 0 for the others.
__device__ inline double oa_weight(double l, int wmode)
{
    if (wmode == 1) return sqrt(fabs(l));
    if (l <= 1e-6) return 0.0;
    const double om = sqrt(l);
    return om > 1e-3 && om < 1e8 ? 0.5 / om : 0.0;
}

// The pair terms of (i, j), i != j, from the lower-triangle values s_ij and
// g_ij: the E_ij numerator and the gap, whether the pair is separated, its
// weight difference rel and whether either mode has weight (used).  The
// separated-or-close test is symmetric in (i, j) (the larger of
// |s_ij + l_j r_ij| and |s_ij + l_i r_ij|).
struct OaPair {
    double num, den, nsym, rel;
    bool sep, used;
};
__device__ inline OaPair oa_pair(double sij, double gij, const double* lamb, double delta, int i, int j, int wmode)
{
    OaPair q;
    const double rij = -gij;
    const double li = lamb[i], lj = lamb[j];
    q.num = sij + lj * rij;
    q.den = lj - li;
    q.nsym = fmax(fabs(q.num), fabs(sij + li * rij));
    const double wi = wmode ? oa_weight(li, wmode) : 1.0, wj = wmode ? oa_weight(lj, wmode) : 1.0;
    const double wm = fmax(wi, wj);
    q.used = wm > 0.0;
    q.rel = wmode == 0 ? 1.0 : (q.used ? fabs(wi - wj) / wm : 0.0);
    q.sep = fabs(q.den) > fmax(8.0 * q.nsym, delta);
    return q;
}

// Pairs that join clusters, one thread per element of the strict lower
// triangle (tiles of OA_ET x OA_ET, read along columns).
__global__ __launch_bounds__(256) void oa_pairs_kernel(const double* S, const double* G, const double* lam,
                                                       const double* delta, int n, unsigned* npair, int2* pairs,
                                                       unsigned cap, int wmode)
{
    const int b = blockIdx.y, t = blockIdx.x;
    int I = (int)((sqrt(8.0 * t + 1.0) - 1.0) * 0.5);
    while ((I + 1) * (I + 2) / 2 <= t) I++;
    while (I * (I + 1) / 2 > t) I--;
    const int J = t - I * (I + 1) / 2;
    const size_t base = (size_t)b * n * n;
    const double dl = delta[b];
    const double* lamb = lam + (size_t)b * n;
    for (int idx = threadIdx.x; idx < OA_ET * OA_ET; idx += 256) {
        const int i = OA_ET * I + idx % OA_ET, j = OA_ET * J + idx / OA_ET;
        if (i >= n || j >= n || i <= j) continue;
        const OaPair q = oa_pair(S[base + i + (size_t)j * n], G[base + i + (size_t)j * n], lamb, dl, i, j, wmode);
        // (a NaN fails the separation test and joins)
        const double es = q.nsym / fabs(q.den);
        const bool join = !q.sep || (q.used && es >= OA_JOIN && es * q.rel >= OA_RHO);
        if (join) {
            const unsigned slot = atomicAdd(&npair[b], 1u);
            if (slot < cap) pairs[(size_t)b * OA_MAXPAIRS + slot] = make_int2(j, i);
        }
    }
}

// Clusters of one problem from its recorded pairs (union-find in LDS, one
// workgroup per problem; dynamic LDS 2 n ints).  Out: ccl[i] the cluster of
// column i (-1 for none), the members of every cluster in ascending order
// (cmem, from ctab[k].x, ctab[k].y of them) and, for one of at most cmax
// (<= OA_CMAX) columns, the offset ctab[k].z of its m x m rotation in qbuf;
// ncl = number of clusters, cinfo = (largest cluster, 1 if one holds more
// than half the columns, clusters above cmax, 0).
__global__ __launch_bounds__(256) void oa_cluster_kernel(const unsigned* npair, const int2* pairs, unsigned cap, int n,
                                                         int* ccl, int* cmem, int4* ctab, int* ncl, int* cinfo,
                                                         int cmax)
{
    extern __shared__ int csh[];
    int* par = csh;
    int* root = csh + n;
    const int b = blockIdx.x;
    for (int i = threadIdx.x; i < n; i += 256) {
        par[i] = i;
        ccl[(size_t)b * n + i] = -1;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        const unsigned np = min(npair[b], cap);
        auto find = [&](int x) {
            while (par[x] != x) {
                par[x] = par[par[x]];   // path halving
                x = par[x];
            }
            return x;
        };
        for (unsigned k = 0; k < np; k++) {
            const int2 pq = pairs[(size_t)b * OA_MAXPAIRS + k];
            const int a = find(pq.x), c = find(pq.y);
            if (a != c) par[max(a, c)] = min(a, c);   // the root is the smallest member
        }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < n; i += 256) {
        int r = i;
        while (par[r] != r) r = par[r];
        root[i] = r;
    }
    __syncthreads();
    for (int i = threadIdx.x; i < n; i += 256) par[i] = 0;
    __syncthreads();
    for (int i = threadIdx.x; i < n; i += 256) atomicAdd(&par[root[i]], 1);   // par: component sizes
    __syncthreads();
    if (threadIdx.x == 0) {
        // ascending sweep: a root comes before its members, so it opens the
        // cluster (par[root] = -(k + 1), root[root] = the fill count) and
        // the members follow in ascending order
        int nc = 0, tot = 0, qt = 0, big = 0, over = 0, nbig = 0;
        for (int i = 0; i < n; i++) {
            const int r = root[i];
            int v = par[r];
            if (r == i && v > 1) {
                ctab[(size_t)b * n + nc] = make_int4(tot, v, qt, 0);
                big = max(big, v);
                over |= 2 * v > n;
                nbig += v > cmax;
                tot += v;
                qt += v <= cmax ? v * v : 0;
                par[r] = v = -(nc + 1);
                root[r] = 0;
                nc++;
            }
            if (v < 0) {
                const int k = -v - 1;
                cmem[(size_t)b * n + ctab[(size_t)b * n + k].x + root[r]++] = i;
                ccl[(size_t)b * n + i] = k;
            }
        }
        ncl[b] = nc;
        cinfo[4 * b] = big;
        cinfo[4 * b + 1] = over;
        cinfo[4 * b + 2] = nbig;
        cinfo[4 * b + 3] = 0;
    }
}

// One element of E; returns (|E_ij|, the predicted next correction of the
// pair) for i != j ((0, 0) on the diagonal and within a cluster), +inf when
// not finite.  A separated pair takes the step's correction num / gap and
// predicts |E_ij| min(1, |E_ij| max|l| / gap) rel (the error the step leaves,
// quadratic where it contracts, weighted by how the solve uses the two
// modes); a pair within a cluster, or a close pair
// (one that could not be recorded, or of two modes of weight 0), keeps only
// the orthogonalising r_ij / 2.  Pairs of two modes of weight 0 are not
// counted.
__device__ float2 oa_e_one(double sij, double gij, const double* lamb, double dl, double* E, const int* cclb, int n,
                           size_t base, int i, int j, int wmode)
{
    const double rij = (i == j ? 1.0 : 0.0) - gij;
    if (i == j) {
        E[base + i + (size_t)j * n] = 0.5 * rij;
        return make_float2(0.f, 0.f);
    }
    const OaPair q = oa_pair(sij, gij, lamb, dl, i, j, wmode);
    const int ci = cclb[i];
    double e, pred = 0.0;
    bool count = q.used;
    if (ci >= 0 && ci == cclb[j]) {
        e = 0.5 * rij;
        count = false;
    } else if (q.sep) {
        e = q.num / q.den;
        if (wmode && q.used) {
            const double kap = (dl * (1.0 / (64.0 * 2.220446049250313e-16))) / fabs(q.den);
            pred = fabs(e) * fmin(1.0, fabs(e) * kap) * q.rel;
        }
    } else {
        e = 0.5 * rij;
    }
    E[base + i + (size_t)j * n] = e;
    const float r = (float)(count || !(fabs(e) <= 1e300) ? fabs(e) : 0.0), p = (float)pred;
    return make_float2(r <= 3.4028235e38f ? r : INFINITY, p <= 3.4028235e38f ? p : INFINITY);
}

// E of one problem by OA_ET x OA_ET tile pairs (I >= J): S and G are valid
// on and below the diagonal, so the lower tile (rows of block I, columns of
// block J) is read once, along its columns, into LDS and serves both E(I, J)
// and its mirror E(J, I) (E is not symmetric); E is written along columns
// too.  One workgroup per tile pair and problem (blockIdx.y).
__global__ __launch_bounds__(256) void oa_e_kernel(const double* S, const double* G, const double* lam,
                                                   const double* delta, double* E, int n, const int* ccl,
                                                   unsigned* emax, unsigned* pmax, int wmode)
{
    __shared__ double st[OA_ET][OA_ET + 1], gt[OA_ET][OA_ET + 1];
    __shared__ float wmax[4], wpred[4];
    const int b = blockIdx.y, t = blockIdx.x;
    int I = (int)((sqrt(8.0 * t + 1.0) - 1.0) * 0.5);
    while ((I + 1) * (I + 2) / 2 <= t) I++;
    while (I * (I + 1) / 2 > t) I--;
    const int J = t - I * (I + 1) / 2;
    const int r0 = OA_ET * I, c0 = OA_ET * J;
    const size_t base = (size_t)b * n * n;
    const double* lamb = lam + (size_t)b * n;
    const int* cclb = ccl + (size_t)b * n;
    const double dl = delta[b];
    for (int idx = threadIdx.x; idx < OA_ET * OA_ET; idx += 256) {
        const int r = idx % OA_ET, c = idx / OA_ET;
        const int gi = r0 + r, gj = c0 + c;
        const bool ok = gi < n && gj < n && gi >= gj;
        st[c][r] = ok ? S[base + gi + (size_t)gj * n] : 0.0;
        gt[c][r] = ok ? G[base + gi + (size_t)gj * n] : 0.0;
    }
    __syncthreads();
    // (oa_e_one returns +inf for a NaN element, so fmaxf keeps it)
    float2 my = make_float2(0.f, 0.f);
    auto take = [&](float2 v) {
        my.x = fmaxf(my.x, v.x);
        my.y = fmaxf(my.y, v.y);
    };
    for (int idx = threadIdx.x; idx < OA_ET * OA_ET; idx += 256) {
        const int r = idx % OA_ET, c = idx / OA_ET;
        // the lower tile: (i, j) = (r0 + r, c0 + c), i >= j
        int i = r0 + r, j = c0 + c;
        if (i < n && j < n && i >= j) take(oa_e_one(st[c][r], gt[c][r], lamb, dl, E, cclb, n, base, i, j, wmode));
        // its mirror: (i, j) = (c0 + r, r0 + c), i < j, from the lower element (j, i)
        i = c0 + r;
        j = r0 + c;
        if (i < n && j < n && i < j) take(oa_e_one(st[r][c], gt[r][c], lamb, dl, E, cclb, n, base, i, j, wmode));
    }
    for (int o = 32; o > 0; o >>= 1) {
        my.x = fmaxf(my.x, __shfl_xor(my.x, o, 64));
        my.y = fmaxf(my.y, __shfl_xor(my.y, o, 64));
    }
    if ((threadIdx.x & 63) == 0) {
        wmax[threadIdx.x >> 6] = my.x;
        wpred[threadIdx.x >> 6] = my.y;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        const size_t slot = (size_t)b * OA_EMAX_SLOTS + (t % OA_EMAX_SLOTS);
        atomicMax(&emax[slot], __float_as_uint(fmaxf(fmaxf(wmax[0], wmax[1]), fmaxf(wmax[2], wmax[3]))));
        atomicMax(&pmax[slot], __float_as_uint(fmaxf(fmaxf(wpred[0], wpred[1]), fmaxf(wpred[2], wpred[3]))));
    }
}

// Each cluster's projected matrix H (m x m, m <= OA_CMAX) diagonalised by
// cyclic Jacobi in FP64 (round-robin order: m/2 disjoint rotations at once,
// H <- J^T H J, Q <- Q J), rotations below a hundredth of the spectrum's
// rounding level skipped; its eigenvalues, ascending, become the members'
// eigenvalue estimates (in member order) and the matching columns of Q are
// stored at qbuf + ctab.z (column-major m x m).  Workgroups stride over the
// clusters of problem blockIdx.y.
__global__ __launch_bounds__(256) void oa_cluster_solve_kernel(const double* S, const double* G, double* lam,
                                                               const double* delta, int n, const int* ncl,
                                                               const int4* ctab, const int* cmem, double* qbuf,
                                                               int cmax)
{
    __shared__ double H[OA_CMAX][OA_CMAX + 1], Q[OA_CMAX][OA_CMAX + 1];
    __shared__ double rc[OA_CMAX / 2], rs[OA_CMAX / 2];
    __shared__ int rp[OA_CMAX / 2], rq[OA_CMAX / 2], idx[OA_CMAX], rank[OA_CMAX];
    __shared__ int any;
    const int b = blockIdx.y, tid = threadIdx.x;
    const size_t base = (size_t)b * n * n;
    double* lamb = lam + (size_t)b * n;
    const double tol = 0.01 * delta[b] / 64.0;   // eps max|l| / 100
    for (int k = blockIdx.x; k < ncl[b]; k += gridDim.x) {
        const int4 ct = ctab[(size_t)b * n + k];
        const int m = ct.y;
        if (m > cmax) continue;
        __syncthreads();
        for (int u = tid; u < m; u += 256) idx[u] = cmem[(size_t)b * n + ct.x + u];
        __syncthreads();
        for (int e = tid; e < m * m; e += 256) {
            const int r = e % m, c = e / m;
            const int i = idx[r], j = idx[c];
            double h;
            if (r == c) {
                h = lamb[i];
            } else {
                const int hi = max(i, j), lo = min(i, j);
                h = S[base + hi + (size_t)lo * n] - 0.5 * (lamb[i] + lamb[j]) * G[base + hi + (size_t)lo * n];
            }
            H[r][c] = h;
            Q[r][c] = r == c ? 1.0 : 0.0;
        }
        const int mp = m + (m & 1);   // even: index m (if any) is a dummy that never rotates
        for (int sweep = 0; sweep < 30; sweep++) {
            __syncthreads();   // every thread has read `any` of the last sweep
            if (tid == 0) any = 0;
            __syncthreads();
            for (int rd = 0; rd < mp - 1; rd++) {
                // round rd: positions 1..mp-1 rotated by rd, position 0 fixed
                if (tid < mp / 2) {
                    const int a = tid == 0 ? 0 : 1 + (tid - 1 + rd) % (mp - 1);
                    const int c = 1 + (mp - 2 - tid + rd) % (mp - 1);
                    const int p = min(a, c), q = max(a, c);
                    double cs = 1.0, sn = 0.0;
                    if (q < m) {
                        const double hpq = H[p][q];
                        if (fabs(hpq) > tol) {
                            const double tau = (H[q][q] - H[p][p]) / (2.0 * hpq);
                            const double tt = (tau >= 0.0 ? 1.0 : -1.0) / (fabs(tau) + sqrt(1.0 + tau * tau));
                            cs = 1.0 / sqrt(1.0 + tt * tt);
                            sn = tt * cs;
                            any = 1;
                        }
                    }
                    rp[tid] = p;
                    rq[tid] = q < m ? q : p;
                    rc[tid] = cs;
                    rs[tid] = sn;
                }
                __syncthreads();
                // rows p, q of H (the pairs are disjoint)
                for (int e = tid; e < (mp / 2) * m; e += 256) {
                    const int t2 = e / m, j = e % m;
                    const double sn = rs[t2];
                    if (sn == 0.0) continue;
                    const int p = rp[t2], q = rq[t2];
                    const double cs = rc[t2], hp = H[p][j], hq = H[q][j];
                    H[p][j] = cs * hp - sn * hq;
                    H[q][j] = sn * hp + cs * hq;
                }
                __syncthreads();
                // columns p, q of H and Q
                for (int e = tid; e < (mp / 2) * m; e += 256) {
                    const int t2 = e / m, j = e % m;
                    const double sn = rs[t2];
                    if (sn == 0.0) continue;
                    const int p = rp[t2], q = rq[t2];
                    const double cs = rc[t2];
                    const double hp = H[j][p], hq = H[j][q];
                    H[j][p] = cs * hp - sn * hq;
                    H[j][q] = sn * hp + cs * hq;
                    const double qp = Q[j][p], qq = Q[j][q];
                    Q[j][p] = cs * qp - sn * qq;
                    Q[j][q] = sn * qp + cs * qq;
                }
                __syncthreads();
            }
            const int go = any;
            if (!go) break;
        }
        // ascending order of the eigenvalues (ties by position)
        for (int u = tid; u < m; u += 256) {
            const double v = H[u][u];
            int rk = 0;
            for (int w = 0; w < m; w++) rk += H[w][w] < v || (H[w][w] == v && w < u);
            rank[u] = rk;
        }
        __syncthreads();
        for (int u = tid; u < m; u += 256) lamb[idx[rank[u]]] = H[u][u];
        double* qk = qbuf + (size_t)b * OA_CMAX * n + ct.z;
        for (int e = tid; e < m * m; e += 256) {
            const int r = e % m, c = e / m;
            qk[r + (size_t)rank[c] * m] = Q[r][c];
        }
    }
}

// X_C <- X_C Q_C for every cluster of problem blockIdx.y, on the 64 rows
// from 64 blockIdx.x (the cluster's columns staged in LDS, then written).
__global__ __launch_bounds__(256) void oa_cluster_apply_kernel(double* X, int n, const int* ncl, const int4* ctab,
                                                               const int* cmem, const double* qbuf, int cmax)
{
    __shared__ double xs[64][OA_CMAX + 1];
    __shared__ double qs[OA_CMAX * OA_CMAX];
    __shared__ int idx[OA_CMAX];
    const int b = blockIdx.y, r0 = 64 * blockIdx.x, tid = threadIdx.x;
    const int nr = min(64, n - r0);
    double* x = X + (size_t)b * n * n;
    for (int k = 0; k < ncl[b]; k++) {
        const int4 ct = ctab[(size_t)b * n + k];
        const int m = ct.y;
        if (m > cmax) continue;
        __syncthreads();
        for (int u = tid; u < m; u += 256) idx[u] = cmem[(size_t)b * n + ct.x + u];
        const double* qk = qbuf + (size_t)b * OA_CMAX * n + ct.z;
        for (int e = tid; e < m * m; e += 256) qs[e] = qk[e];
        __syncthreads();
        for (int e = tid; e < 64 * m; e += 256) {
            const int r = e % 64, c = e / 64;
            xs[r][c] = r < nr ? x[r0 + r + (size_t)idx[c] * n] : 0.0;
        }
        __syncthreads();
        for (int e = tid; e < 64 * m; e += 256) {
            const int r = e % 64, c = e / 64;
            if (r >= nr) continue;
            double acc = 0.0;
            for (int u = 0; u < m; u++) acc += xs[r][u] * qs[u + c * m];
            x[r0 + r + (size_t)idx[c] * n] = acc;
        }
    }
}

// Clusters above OA_CMAX columns, solved by the batched FP64 eigensolver in
// chunks of at most OA_BIGCHUNK of one padded order mb.  bigl[s] = (problem,
// cluster, 0, 0) of chunk slot s.
constexpr int OA_BIGCHUNK = 256;

// H of chunk slot blockIdx.x (mb x mb at H + slot mb^2): the cluster's
// projected matrix, padded with the diagonal value c above its Gershgorin
// bound so the pad eigenvalues come last and the first m eigenvectors have
// zero pad rows.
__global__ __launch_bounds__(256) void oa_big_h_kernel(const double* S, const double* G, const double* lam, int n,
                                                       const int4* bigl, const int4* ctab, const int* cmem, double* H,
                                                       int mb)
{
    __shared__ double red[4];
    const int4 L = bigl[blockIdx.x];
    const int b = L.x;
    const int4 ct = ctab[(size_t)b * n + L.y];
    const int m = ct.y;
    const int* idx = cmem + (size_t)b * n + ct.x;
    const size_t base = (size_t)b * n * n;
    const double* lamb = lam + (size_t)b * n;
    double* h = H + (size_t)blockIdx.x * mb * mb;
    auto hval = [&](int r, int c) -> double {
        const int i = idx[r], j = idx[c];
        if (r == c) return lamb[i];
        const int hi = max(i, j), lo = min(i, j);
        return S[base + hi + (size_t)lo * n] - 0.5 * (lamb[i] + lamb[j]) * G[base + hi + (size_t)lo * n];
    };
    double gm = 0.0;
    for (int c = threadIdx.x; c < m; c += 256) {
        double cs = 0.0;
        for (int r = 0; r < m; r++) cs += fabs(hval(r, c));
        gm = fmax(gm, cs);
    }
    for (int o = 32; o > 0; o >>= 1) gm = fmax(gm, __shfl_xor(gm, o, 64));
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = gm;
    __syncthreads();
    gm = fmax(fmax(red[0], red[1]), fmax(red[2], red[3]));
    const double pad = gm * (1.0 + 0.0009765625) + 2.2250738585072014e-308;
    for (int64_t e = threadIdx.x; e < (int64_t)mb * mb; e += 256) {
        const int r = (int)(e % mb), c = (int)(e / mb);
        h[e] = (r < m && c < m) ? hval(r, c) : (r == c ? pad : 0.0);
    }
}

// The eigenvalues of chunk slot blockIdx.x (ascending, pads last) become the
// members' estimates.
__global__ void oa_big_lam_kernel(const double* w, int mb, const int4* bigl, const int4* ctab, const int* cmem,
                                  double* lam, int n)
{
    const int4 L = bigl[blockIdx.x];
    const int b = L.x;
    const int4 ct = ctab[(size_t)b * n + L.y];
    for (int k = threadIdx.x; k < ct.y; k += blockDim.x)
        lam[(size_t)b * n + cmem[(size_t)b * n + ct.x + k]] = w[(size_t)blockIdx.x * mb + k];
}

// XC (n x m) = the cluster's columns of X (to = 0), or back (to = 1).
__global__ void oa_big_move_kernel(double* X, int n, const int* idx, int m, double* XC, int to)
{
    const int64_t total = (int64_t)n * m;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
        const int r = (int)(e % n), c = (int)(e / n);
        double* xp = X + r + (size_t)idx[c] * n;
        if (to) *xp = XC[e];
        else XC[e] = *xp;
    }
}

// One Ogita-Aishima step: X <- X + X E, then the cluster rotations.
// Buffers: AX, S, G, Xn (B x n x n).
struct OaPairs {
    unsigned* npair;    // B
    int2* pairs;        // B x OA_MAXPAIRS: (j, i), i > j
    int* ccl;           // B x n: cluster of each column, -1 for none
    int* cmem;          // B x n: members of the clusters
    int4* ctab;         // B x n: (first member, size, rotation offset) per cluster
    int* ncl;           // B
    int* cinfo;         // B x 4: largest cluster, one above half the columns, clusters above cmax
    int4* bigl;         // OA_BIGCHUNK: the clusters above cmax of one eigensolver chunk
    double* qbuf;       // B x OA_CMAX x n: the clusters' rotations
    unsigned* emax;     // B x OA_EMAX_SLOTS: float bits of max |E_ij| (i != j) of the last step
    unsigned* pmax;     // B x OA_EMAX_SLOTS: float bits of the largest predicted next correction
    unsigned cap = OA_MAXPAIRS;   // pairs recorded per problem (<= OA_MAXPAIRS)
    int cmax = OA_CMAX;           // largest cluster the workgroup Jacobi takes (2 .. OA_CMAX)
};

// Xn = X + s_b P, s_b the scale of problem b = i / nn (1 without scales).
__global__ void oa_add_kernel(const double* X, const float* P, double* Xn, int64_t count, int64_t nn,
                              const double* s)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += (int64_t)gridDim.x * blockDim.x)
        Xn[i] = X[i] + (s ? s[i / nn] : 1.0) * (double)P[i];
}

// X, and E / s_b per problem, in IEEE half precision (|X_ij| <= 1; E scaled
// to |E| <= ~1).  inv: 1 / s_b per problem (NULL: no scaling).
__global__ void f64_to_f16_kernel(const double* a, __half* b, int64_t count, int64_t nn, const double* inv)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += (int64_t)gridDim.x * blockDim.x)
        b[i] = __float2half((float)(a[i] * (inv ? inv[i / nn] : 1.0)));   // via float: 2^-24 against half's 2^-11
}

// hi + 2^-11 lo split of a / s_b in IEEE half precision: the lo part carries
// the next 11 bits, scaled by 2^11 so it stays in half's normal range; the
// pair represents the value to 2^-22 relative (1.5e-11 absolute below half's
// normal range).
__global__ void f64_to_f16x2_kernel(const double* a, __half* hi, __half* lo, int64_t count, int64_t nn,
                                    const double* inv)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += (int64_t)gridDim.x * blockDim.x) {
        const float v = (float)(a[i] * (inv ? inv[i / nn] : 1.0));
        const __half h = __float2half(v);
        hi[i] = h;
        lo[i] = __float2half((v - __half2float(h)) * 2048.0f);
    }
}

// The update X E is a correction of X: its rounding adds about u |X||E| to
// an error the step leaves at about max|E|^2 (the step converges
// quadratically).  A step with max|E| >= half_from forms it from IEEE half
// inputs (u = 4.9e-4) with FP32 accumulation, which keeps u max|E| at or
// below a tenth of max|E|^2.  Smaller corrections split X and E / max|E|
// into hi + 2^-11 lo half pairs and sum three half products (hi hi, hi lo,
// lo hi; lo lo lies below 2^-22): close to FP32 accuracy (2^-22 against
// 2^-24) on the half MFMA rate.  The residual terms (A X, X^T A X, X^T X) are FP64.
// pe receives each problem's correction of the step (oa_effective: +inf for
// a non-finite step or more close pairs than recorded).
static int oa_refine(hipStream_t st, int n, int B, int wmode, const double* A, double*& X, double*& Xn, double* lam,
                     double* AX,
                     double* S, double* G, double* partial, int npart, double* delta, const OaPairs& op, float* X32,
                     float* E32, float* P32, float half_from, double* dscale, std::vector<float>& pe,
                     const EigSolveAt* cluster_eig)
{
    const size_t nn = (size_t)n * n;
    GemmShape gnn{n, n, n, false, false, n, n, n, (int64_t)nn, (int64_t)nn, (int64_t)nn, B, 0};
    GemmShape gtl{n, n, n, true, false, n, n, n, (int64_t)nn, (int64_t)nn, (int64_t)nn, B, 1};
    GMC_TRY(gemm_big<double>(st, gnn, 1.0, A, X, 0.0, AX));
    GMC_TRY(gemm_big<double>(st, gtl, 1.0, X, AX, 0.0, S));
    GMC_TRY(gemm_big<double>(st, gtl, 1.0, X, X, 0.0, G));
    hipLaunchKernelGGL(oa_stats_kernel, dim3(npart, B), dim3(256), 0, st, S, G, n, lam, partial);
    hipLaunchKernelGGL(oa_delta_kernel, dim3(1), dim3(64), 0, st, partial, npart, delta, B);
    GMC_HIP(hipMemsetAsync(op.npair, 0, B * sizeof(unsigned), st));
    GMC_HIP(hipMemsetAsync(op.emax, 0, (size_t)B * OA_EMAX_SLOTS * sizeof(unsigned), st));
    GMC_HIP(hipMemsetAsync(op.pmax, 0, (size_t)B * OA_EMAX_SLOTS * sizeof(unsigned), st));
    const int nte = (n + OA_ET - 1) / OA_ET;
    const dim3 gtri(nte * (nte + 1) / 2, B);
    // the pairs that join clusters, the clusters, then E with them
    hipLaunchKernelGGL(oa_pairs_kernel, gtri, dim3(256), 0, st, S, G, lam, delta, n, op.npair, op.pairs, op.cap,
                       wmode);
    const size_t clds = 2 * (size_t)n * sizeof(int);
    GMC_CHECK(clds <= 160 * 1024, "ERPA mixed: order %d too large for the cluster kernel's LDS", n);
    static bool clds_set = false;
    if (!clds_set) {
        GMC_HIP(hipFuncSetAttribute((const void*)oa_cluster_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                                    160 * 1024));
        clds_set = true;
    }
    hipLaunchKernelGGL(oa_cluster_kernel, dim3(B), dim3(256), clds, st, op.npair, op.pairs, op.cap, n, op.ccl,
                       op.cmem, op.ctab, op.ncl, op.cinfo, op.cmax);
    hipLaunchKernelGGL(oa_e_kernel, gtri, dim3(256), 0, st, S, G, lam, delta, AX, n, op.ccl, op.emax, op.pmax, wmode);
    GMC_HIP(hipGetLastError());
    // each problem's largest off-diagonal correction and prediction
    // (non-negative floats order as their bits; +inf marks a non-finite
    // element), pair count and clusters
    std::vector<unsigned> em((size_t)B * OA_EMAX_SLOTS), pm((size_t)B * OA_EMAX_SLOTS), np(B);
    std::vector<int> ci(4 * (size_t)B), ncl(B);
    GMC_HIP(hipMemcpyAsync(em.data(), op.emax, em.size() * sizeof(unsigned), hipMemcpyDeviceToHost, st));
    GMC_HIP(hipMemcpyAsync(pm.data(), op.pmax, pm.size() * sizeof(unsigned), hipMemcpyDeviceToHost, st));
    GMC_HIP(hipMemcpyAsync(np.data(), op.npair, B * sizeof(unsigned), hipMemcpyDeviceToHost, st));
    GMC_HIP(hipMemcpyAsync(ci.data(), op.cinfo, ci.size() * sizeof(int), hipMemcpyDeviceToHost, st));
    GMC_HIP(hipMemcpyAsync(ncl.data(), op.ncl, B * sizeof(int), hipMemcpyDeviceToHost, st));
    GMC_HIP(hipStreamSynchronize(st));
    pe.assign(B, 0.f);
    // Clusters above cmax columns: (problem, cluster, size, first member),
    // largest first, in eigensolver chunks of one padded order that fit the
    // old X's space (free once the update has swapped X and Xn).  A problem
    // whose cluster cannot be solved so stays unconverged.
    std::vector<char> bad(B, 0);
    std::vector<int4> bigs;
    for (int b = 0; b < B; b++) {
        unsigned mxb = 0, pxb = 0;
        for (int k = 0; k < OA_EMAX_SLOTS; k++) {
            mxb = std::max(mxb, em[(size_t)b * OA_EMAX_SLOTS + k]);
            pxb = std::max(pxb, pm[(size_t)b * OA_EMAX_SLOTS + k]);
        }
        // (+inf bits or above: a non-finite step, whose clusters are not solved)
        if (ci[4 * b + 1] || mxb >= 0x7f800000u || pxb >= 0x7f800000u) bad[b] = 1;
        if (bad[b] || ci[4 * b + 2] == 0) continue;
        if (!cluster_eig) {
            bad[b] = 1;
            continue;
        }
        std::vector<int4> tab(ncl[b]);
        GMC_HIP(hipMemcpy(tab.data(), op.ctab + (size_t)b * n, tab.size() * sizeof(int4), hipMemcpyDeviceToHost));
        for (int k = 0; k < ncl[b]; k++)
            if (tab[k].y > op.cmax) bigs.push_back(make_int4(b, k, tab[k].y, tab[k].x));
    }
    std::stable_sort(bigs.begin(), bigs.end(), [](const int4& u, const int4& v) { return u.z > v.z; });
    struct BigChunk {
        int mb;
        size_t i0, c;
    };
    std::vector<BigChunk> chunks;
    {
        const size_t avail = (size_t)B * nn * sizeof(double);
        auto need = [&](int mb, int c) -> size_t {
            const size_t ws = eig_workspace_bytes<double>(mb, c);
            if (ws == SIZE_MAX) return SIZE_MAX;
            return align_up((size_t)c * mb * mb * sizeof(double)) + align_up((size_t)c * mb * sizeof(double)) + ws;
        };
        size_t i0 = 0;
        while (i0 < bigs.size()) {
            const int mb = (bigs[i0].z + 31) / 32 * 32;
            int c = (int)std::min<size_t>(OA_BIGCHUNK, bigs.size() - i0);
            while (c > 1 && need(mb, c) > avail) c = (c + 1) / 2;
            if (need(mb, c) > avail) {
                bad[bigs[i0].x] = 1;
                i0++;
                continue;
            }
            chunks.push_back({mb, i0, (size_t)c});
            i0 += c;
        }
    }
    // Each problem's update is scaled by its own largest correction: one
    // batch holds problems at every stage of convergence (the batching form
    // mixes systems), and a batch-wide scale would push a converging
    // problem's small corrections below half's normal range.  A problem at
    // +inf is solved again or fails the call; its update is moot.
    std::vector<float> eb(B);
    float emin = INFINITY;
    for (int b = 0; b < B; b++) {
        unsigned mxb = 0, pxb = 0;
        for (int k = 0; k < OA_EMAX_SLOTS; k++) {
            mxb = std::max(mxb, em[(size_t)b * OA_EMAX_SLOTS + k]);
            pxb = std::max(pxb, pm[(size_t)b * OA_EMAX_SLOTS + k]);
        }
        float e, q;
        std::memcpy(&e, &mxb, 4);
        std::memcpy(&q, &pxb, 4);
        pe[b] = bad[b] ? INFINITY : oa_effective(oa_measure(e, q), np[b], op.cap);
        eb[b] = e <= 3.4028235e38f ? e : 1.0f;
        if (e <= 3.4028235e38f) emin = std::min(emin, e);
    }
    // plain half inputs only when every problem's correction is large enough
    // for their rounding (about 4.9e-4 max|E|) to stay under a tenth of the
    // max|E|^2 the step leaves; else the split form for the whole batch
    const bool plain = emin >= half_from;
    std::vector<double> sc(2 * (size_t)B);
    for (int b = 0; b < B; b++) {
        // the split form's floor 1e-4: the diagonal E_ii = r_ii / 2 (not in
        // max|E|) is a few 1e-6 at most from the FP32 start on (orthogonality
        // ~3e-6), so E / s stays far inside half's range
        const double sb = plain ? (eb[b] > 0.f ? eb[b] : 1.0) : std::max(eb[b], 1e-4f);
        sc[b] = sb;
        sc[B + b] = 1.0 / sb;
    }
    GMC_HIP(hipMemcpyAsync(dscale, sc.data(), sc.size() * sizeof(double), hipMemcpyHostToDevice, st));
    const double* s_b = dscale;
    const double* inv_b = dscale + B;
    const unsigned gcopy = (unsigned)std::min<int64_t>(65536, ((int64_t)B * nn + 255) / 256);
    const int64_t cnt = (int64_t)B * nn, snn = (int64_t)nn;
    if (plain) {
        // half inputs in the FP32 buffers' space, E scaled by 1 / s_b
        __half* X16 = reinterpret_cast<__half*>(X32);
        __half* E16 = reinterpret_cast<__half*>(E32);
        hipLaunchKernelGGL(f64_to_f16_kernel, dim3(gcopy), dim3(256), 0, st, X, X16, cnt, snn, nullptr);
        hipLaunchKernelGGL(f64_to_f16_kernel, dim3(gcopy), dim3(256), 0, st, AX, E16, cnt, snn, inv_b);
        GMC_HIP(hipGetLastError());
        GMC_TRY(gemm_f16(st, gnn, 1.0f, X16, E16, 0.0f, P32));
    } else {
        // hi and lo halves of X in X32's space, of E / s_b in E32's
        __half* Xh = reinterpret_cast<__half*>(X32);
        __half* Xl = Xh + (size_t)B * nn;
        __half* Eh = reinterpret_cast<__half*>(E32);
        __half* El = Eh + (size_t)B * nn;
        hipLaunchKernelGGL(f64_to_f16x2_kernel, dim3(gcopy), dim3(256), 0, st, X, Xh, Xl, cnt, snn, nullptr);
        hipLaunchKernelGGL(f64_to_f16x2_kernel, dim3(gcopy), dim3(256), 0, st, AX, Eh, El, cnt, snn, inv_b);
        GMC_HIP(hipGetLastError());
        const float lo = 1.0f / 2048.0f;
        GMC_TRY(gemm_f16(st, gnn, 1.0f, Xh, Eh, 0.0f, P32));
        GMC_TRY(gemm_f16(st, gnn, lo, Xh, El, 1.0f, P32));
        GMC_TRY(gemm_f16(st, gnn, lo, Xl, Eh, 1.0f, P32));
    }
    // the clusters' rotations (from S, G and the step's eigenvalues) and
    // eigenvalue estimates, applied after the update
    hipLaunchKernelGGL(oa_cluster_solve_kernel, dim3(64, B), dim3(256), 0, st, S, G, lam, delta, n, op.ncl, op.ctab,
                       op.cmem, op.qbuf, op.cmax);
    hipLaunchKernelGGL(oa_add_kernel, dim3(gcopy), dim3(256), 0, st, X, P32, Xn, cnt, snn, s_b);
    GMC_HIP(hipGetLastError());
    std::swap(X, Xn);
    hipLaunchKernelGGL(oa_cluster_apply_kernel, dim3((n + 63) / 64, B), dim3(256), 0, st, X, n, op.ncl, op.ctab,
                       op.cmem, op.qbuf, op.cmax);
    GMC_HIP(hipGetLastError());
    // the large clusters: H in the old X's space, its eigenvectors rotate
    // the cluster's columns (moved through AX, free after the update)
    for (const BigChunk& ch : chunks) {
        if (bad[bigs[ch.i0].x]) continue;
        std::vector<int4> list(ch.c);
        for (size_t u = 0; u < ch.c; u++) list[u] = make_int4(bigs[ch.i0 + u].x, bigs[ch.i0 + u].y, 0, 0);
        GMC_HIP(hipMemcpyAsync(op.bigl, list.data(), ch.c * sizeof(int4), hipMemcpyHostToDevice, st));
        const int mb = ch.mb, c = (int)ch.c;
        double* H = Xn;
        double* w = H + align_up((size_t)c * mb * mb * sizeof(double)) / sizeof(double);
        void* ws = w + align_up((size_t)c * mb * sizeof(double)) / sizeof(double);
        const size_t wsb = eig_workspace_bytes<double>(mb, c);
        hipLaunchKernelGGL(oa_big_h_kernel, dim3(c), dim3(256), 0, st, S, G, lam, n, op.bigl, op.ctab, op.cmem, H, mb);
        GMC_HIP(hipGetLastError());
        GMC_TRY((*cluster_eig)(H, w, mb, c, ws, wsb));
        hipLaunchKernelGGL(oa_big_lam_kernel, dim3(c), dim3(256), 0, st, w, mb, op.bigl, op.ctab, op.cmem, lam, n);
        for (size_t u = 0; u < ch.c; u++) {
            const int4 bg = bigs[ch.i0 + u];
            const int m = bg.z;
            const int* idx = op.cmem + (size_t)bg.x * n + bg.w;
            double* Xb = X + (size_t)bg.x * nn;
            double* XC = AX;
            double* XO = AX + (size_t)n * m;
            const unsigned gm = (unsigned)std::min<int64_t>(65536, ((int64_t)n * m + 255) / 256);
            hipLaunchKernelGGL(oa_big_move_kernel, dim3(gm), dim3(256), 0, st, Xb, n, idx, m, XC, 0);
            GemmShape g{n, m, m, false, false, n, mb, n, 0, 0, 0, 1, 0};
            GMC_TRY(gemm_big<double>(st, g, 1.0, XC, H + (size_t)u * mb * mb, 0.0, XO));
            hipLaunchKernelGGL(oa_big_move_kernel, dim3(gm), dim3(256), 0, st, Xb, n, idx, m, XO, 1);
            GMC_HIP(hipGetLastError());
        }
        // `list` must outlive its copy
        GMC_HIP(hipStreamSynchronize(st));
    }
    return 0;
}

__global__ void sqrt_scale64_kernel(double* V, const double* lam, int n, int* nneg)
{
    const int b = blockIdx.z, k = blockIdx.y;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const double l = lam[(size_t)b * n + k];
    if (i < n) V[(size_t)b * n * n + i + (size_t)k * n] *= sqrt(sqrt(fabs(l)));
    if (i == 0 && l < 0.0) atomicAdd(&nneg[b], 1);
}

size_t erpa_mixed_bytes(int n, int batch)
{
    const size_t nn = (size_t)n * n, B = batch;
    return carve_bytes({B * nn * 8, B * nn * 8, B * nn * 8, B * nn * 8, B * nn * 8, B * nn * 8, B * nn * 8,
                        B * nn * 4, B * n * 4, B * n * 8, B * n * 8, B * 8, B * 1024 * 3 * 8, B * 8, B * 4,
                        B * (size_t)OA_MAXPAIRS * 8, B * n * 4, B * n * 4, B * n * 16, B * 4, 4 * B * 4,
                        B * (size_t)OA_CMAX * n * 8, (size_t)OA_BIGCHUNK * 16, B * nn * 4, B * nn * 4,
                        B * (size_t)OA_EMAX_SLOTS * 4, B * (size_t)OA_EMAX_SLOTS * 4, B * n * 8,
                        potrf_work_bytes(n, batch), B * 4, 2 * B * 8});
}

// ------------------------------------------------------------ Cholesky route
// The reference's S = |A+|^(1/2) (interpa.f:1693-1716) enters only through
// M = S A- S and Y = S Z.  Any factor L of |A+| = L L^T serves as well: L =
// S Q with Q orthogonal, L^T A- L = Q^T M Q has the eigenvectors Q^T Z, and
// L Q^T Z = S Z.  |A+| = A+ - 2 sum_{lambda < 0} lambda x x^T needs only the
// few eigenpairs of A+ below zero to FP64 accuracy: the K smallest columns of
// the FP32 eigenvector matrix (every negative eigenvalue and at least 8 more)
// are refined by Ogita-Aishima steps restricted to those columns, the FP32
// basis and eigenvalues standing in for the rest.  Each step costs O(n^2 K)
// and gains about the FP32 basis' accuracy, so two or three steps reach the
// FP64 rounding floor, where the full refinement of A+ costs three O(n^3)
// steps.

// Negative FP32 eigenvalues per problem (ascending order).
__global__ void count_neg_kernel(const float* w, int n, int B, int* cnt)
{
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= B) return;
    int c = 0;
    while (c < n && w[(size_t)b * n + c] < 0.f) c++;
    cnt[b] = c;
}

// lambda_j = s_jj / g_jj of the refined columns (S = X^T A X_J, G = X^T X_J,
// n x K, ld n); delta_b = 64 eps max |lambda| (the cluster threshold).
__global__ void oac_lam_kernel(const double* S, const double* G, const float* w32, int n, int K, double* lamJ,
                               double* delta)
{
    const int b = blockIdx.x;
    const size_t nn = (size_t)n * n;
    double mx = fmax(fabs((double)w32[(size_t)b * n]), fabs((double)w32[(size_t)b * n + n - 1]));
    for (int j = threadIdx.x; j < K; j += 64) {
        const size_t o = (size_t)b * nn + (size_t)j * n + j;
        const double l = S[o] / G[o];
        lamJ[(size_t)b * K + j] = l;
        mx = fmax(mx, fabs(l));
    }
    for (int o = 32; o > 0; o >>= 1) mx = fmax(mx, __shfl_xor(mx, o, 64));
    if (threadIdx.x == 0) delta[b] = 64.0 * 2.220446049250313e-16 * mx;
}

// E_ij of the restricted step (i < n, j < K): (s_ij - l_j g_ij) / (l_j - l_i)
// for separated pairs, -g_ij / 2 for close ones (the same rule as the full
// step), (1 - g_jj) / 2 on the diagonal; l_i of an unrefined column is its
// FP32 eigenvalue.  emax[b]: the largest |E_ij| off the diagonal over the
// columns j with l_j < 0, the only ones |A+| takes (the others are refined
// along but need not converge: a positive column next to the unrefined part
// of the spectrum can converge slowly and does not matter), as float bits
// (+inf for a non-finite element).
__global__ __launch_bounds__(256) void oac_e_kernel(const double* S, const double* G, const double* lamJ,
                                                    const float* w32, const double* delta, int n, int K, double* E,
                                                    unsigned* emax)
{
    const int b = blockIdx.z, j = blockIdx.y;
    const int i = blockIdx.x * 256 + threadIdx.x;
    const size_t o = (size_t)b * n * n + (size_t)j * n + i;
    float ea = 0.f;
    if (i < n) {
        const double s = S[o], g = G[o];
        double e;
        if (i == j) {
            e = 0.5 * (1.0 - g);
        } else {
            const double lj = lamJ[(size_t)b * K + j];
            const double li = i < K ? lamJ[(size_t)b * K + i] : (double)w32[(size_t)b * n + i];
            const double gap = lj - li;
            e = (s - lj * g) / gap;
            if (!(fabs(gap) > delta[b]) || !(fabs(e) < 0.125)) e = -0.5 * g;
            if (!(lj >= 0.0)) ea = isfinite(e) ? (float)fabs(e) : INFINITY;
        }
        E[o] = e;
    }
    for (int off = 32; off > 0; off >>= 1) ea = fmaxf(ea, __shfl_xor(ea, off, 64));
    if ((threadIdx.x & 63) == 0) atomicMax(&emax[b], __float_as_uint(ea));
}

// X_J += P over the first K columns of each problem.
__global__ void oac_add_kernel(double* X, const double* P, int n, int K)
{
    const size_t o = (size_t)blockIdx.y * n * n;
    const int64_t cnt = (int64_t)n * K;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < cnt; i += (int64_t)gridDim.x * blockDim.x)
        X[o + i] += P[o + i];
}

// T = X_J diag(-2 min(lambda_j, 0)) (n x K, ld n); nneg counts lambda_j < 0.
__global__ void oac_reflect_kernel(const double* X, const double* lamJ, int n, int K, double* T, int* nneg)
{
    const int b = blockIdx.z, j = blockIdx.y;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const double l = lamJ[(size_t)b * K + j];
    const double d = l < 0.0 ? -2.0 * l : 0.0;
    const size_t o = (size_t)b * n * n + (size_t)j * n + i;
    if (i < n) T[o] = d * X[o];
    if (i == 0 && l < 0.0) atomicAdd(&nneg[b], 1);
}

// The batching form in the mixed path: the padded modes of A+ and of M get
// diagonal values apart from the system's spectrum, so that the FP32
// eigensolves, whose close-eigenvalue rotations act at FP32 rounding (about
// 1e-7 of the norm), cannot mix them with the system's modes near zero.
// A+'s pads take +max_i A+_ii and M's pads -max_i M_ii (i over the system's
// own order): neither exceeds the largest eigenvalue, so the deflation
// tolerances (8 eps max|d|) stay as they are.  With the pad value c in A+,
// |A+| has c on its pad diagonal, its factor sqrt(c), and M's pad block is
// zero before its pads are set.  The pad modes keep zero weight in the
// integrand: only the system's block is read, and M's pads have omega^2 < 0,
// which the filter drops.  One workgroup (256 threads) per problem.
__global__ void pad_diag_set_kernel(const BatchSys* sys, int nalpha, double* A, int npad, double sign)
{
    __shared__ double red[4];
    const int b = blockIdx.x, k = b / nalpha;
    const int n = sys[k].n;
    double* Ab = A + (size_t)b * npad * npad;
    double mx = 0.0;
    for (int i = threadIdx.x; i < n; i += blockDim.x) mx = fmax(mx, Ab[(size_t)i * npad + i]);
    for (int o = 32; o > 0; o >>= 1) mx = fmax(mx, __shfl_xor(mx, o, 64));
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = mx;
    __syncthreads();
    mx = fmax(fmax(red[0], red[1]), fmax(red[2], red[3]));
    if (!(mx > 0.0)) mx = 1.0;
    for (int i = n + threadIdx.x; i < npad; i += blockDim.x) Ab[(size_t)i * npad + i] = sign * mx;
}

// Mixed ERPA: eigen solves of A+ and of M in FP32.  A+ gives the factor F
// of |A+| = F F^T: its Cholesky factor (the Cholesky route above), or S =
// |A+|^(1/2) from the full refinement when that route fails; M = F^T A- F.
// M's eigenpairs are refined in FP64 by n_min to n_max Ogita-Aishima
// iterations (n_min >= 1), problems that do not converge solved again in
// FP64 when `resolve` is given; F, M, Y and the integrand in FP64.
int erpa_batched_mixed(hipStream_t st, EigPlan<float>& ep, int n, const AbPair& ab, const double* alphas, int nalpha,
                       const double* Gw, const double* wdiag, void* scratch, size_t scratch_bytes, ErpaOut& out,
                       int n_min, int n_max, float oa_extend, const EigSolve64* resolve, float half_from,
                       const ErpaBatchTables* bt, const EigSolveAt* cluster_eig)
{
    // B problems: the alpha points, or (batching form) every system at every alpha point
    const int B = bt ? bt->nsys * nalpha : nalpha;
    GMC_CHECK(ep.bound && ep.batch == B && ep.n == n, "ERPA mixed: eigensolver plan mismatch");
    const size_t nn = (size_t)n * n;
    Arena ar{static_cast<char*>(scratch), scratch_bytes};
    double* Ap = ar.take<double>(B * nn);
    double* Am = ar.take<double>(B * nn);
    double* X = ar.take<double>(B * nn);
    double* AX = ar.take<double>(B * nn);
    double* Sb = ar.take<double>(B * nn);
    double* Gb = ar.take<double>(B * nn);
    double* Xn = ar.take<double>(B * nn);
    float* E32 = ar.take<float>(B * nn);
    float* w32 = ar.take<float>(B * n);
    double* lam = ar.take<double>(B * n);
    double* om2 = ar.take<double>(B * n);
    double* d_al = ar.take<double>(B);
    const int npart = 1024;
    double* partial = ar.take<double>((size_t)B * npart * 3);
    double* delta = ar.take<double>(B);
    OaPairs op;
    op.npair = ar.take<unsigned>(B);
    op.pairs = ar.take<int2>((size_t)B * OA_MAXPAIRS);
    op.ccl = ar.take<int>((size_t)B * n);
    op.cmem = ar.take<int>((size_t)B * n);
    op.ctab = ar.take<int4>((size_t)B * n);
    op.ncl = ar.take<int>(B);
    op.cinfo = ar.take<int>(4 * (size_t)B);
    op.qbuf = ar.take<double>((size_t)B * OA_CMAX * n);
    op.bigl = ar.take<int4>(OA_BIGCHUNK);
    float* X32 = ar.take<float>(B * nn);
    float* P32 = ar.take<float>(B * nn);
    op.emax = ar.take<unsigned>((size_t)B * OA_EMAX_SLOTS);
    op.pmax = ar.take<unsigned>((size_t)B * OA_EMAX_SLOTS);
    // GMC_OA_MAXPAIRS lowers the number of close pairs recorded per problem
    // (diagnostics: forces the over-the-cap path)
    if (const char* e = getenv("GMC_OA_MAXPAIRS")) op.cap = (unsigned)std::min(std::max(0, atoi(e)), OA_MAXPAIRS);
    // GMC_OA_CMAX lowers the largest cluster the workgroup Jacobi solves
    // (diagnostics: sends small clusters to the batched eigensolver)
    if (const char* e = getenv("GMC_OA_CMAX")) op.cmax = std::min(std::max(2, atoi(e)), OA_CMAX);
    double* wre = ar.take<double>(B * n);
    char* pwork = ar.take<char>(potrf_work_bytes(n, B));
    int* info = ar.take<int>(B);
    double* dscale = ar.take<double>(2 * (size_t)B);
    GMC_CHECK(ar.fits(), "ERPA mixed: scratch layout overflow");
    const bool stats = getenv("GMC_OA_STATS") != nullptr;
    // (the first nb problems of the batch: the ones the last step refined)
    auto report = [&](const char* what, int it, int nb) -> int {
        if (!stats) return 0;
        std::vector<unsigned> h(nb);
        GMC_HIP(hipMemcpyAsync(h.data(), op.npair, nb * sizeof(unsigned), hipMemcpyDeviceToHost, st));
        GMC_HIP(hipStreamSynchronize(st));
        unsigned tot = 0;
        for (unsigned v : h) tot += v;
        std::vector<unsigned> em((size_t)nb * OA_EMAX_SLOTS);
        GMC_HIP(hipMemcpyAsync(em.data(), op.emax, em.size() * sizeof(unsigned), hipMemcpyDeviceToHost, st));
        GMC_HIP(hipStreamSynchronize(st));
        float mx = 0.f;
        for (unsigned u : em) {
            float f;
            std::memcpy(&f, &u, 4);
            mx = std::max(mx, f);
        }
        std::vector<int> nc(nb), ci(4 * (size_t)nb);
        GMC_HIP(hipMemcpyAsync(nc.data(), op.ncl, nb * sizeof(int), hipMemcpyDeviceToHost, st));
        GMC_HIP(hipMemcpyAsync(ci.data(), op.cinfo, ci.size() * sizeof(int), hipMemcpyDeviceToHost, st));
        GMC_HIP(hipStreamSynchronize(st));
        int ctot = 0, big = 0, nlarge = 0;
        for (int b = 0; b < nb; b++) {
            ctot += nc[b];
            big = std::max(big, ci[4 * b]);
            nlarge += ci[4 * b + 2];
        }
        fprintf(stderr,
                "[gmc] OA %s step %d: %u pairs in %d clusters (largest %d, %d to the eigensolver) over %d problems, "
                "max |E| %.3e\n", what, it, tot, ctot, big, nlarge, nb, mx);
        return 0;
    };
    // Largest correction of the last step, per problem (oa_effective)
    std::vector<float> pe(B);
    // Refinement of X and lamv against the FP64 matrices A, step by step until
    // oa_decide (oa_policy.hpp) stops it; with `resolve`, the problems it
    // leaves unconverged go to `left`.  Without one, a problem whose
    // corrections are not finite or whose close pairs overflowed the record
    // fails the call.
    //
    // After the first n_min steps a problem whose correction is below
    // OA_LEAVE has converged and leaves the batch: the problems
    // still refined are moved to the front (their slots of A, X and lamv
    // swapped with converged ones), so the next steps' products and kernels
    // run on them alone; the original order is restored at the end.
    auto refine = [&](const char* what, int wmode, double* A, double* lamv, double* axbuf,
                      std::vector<int>& left) -> int {
        left.clear();
        std::vector<int> prob(B);   // prob[slot] = the problem held in that slot
        for (int b = 0; b < B; b++) prob[b] = b;
        int nact = B;
        // swap the data of slots a and c (Xn, free between steps, as the temporary)
        auto swap_slots = [&](int a, int c) -> int {
            auto sw = [&](double* base, size_t len) -> int {
                GMC_HIP(hipMemcpyAsync(Xn, base + (size_t)a * len, len * sizeof(double), hipMemcpyDeviceToDevice, st));
                GMC_HIP(hipMemcpyAsync(base + (size_t)a * len, base + (size_t)c * len, len * sizeof(double),
                                       hipMemcpyDeviceToDevice, st));
                GMC_HIP(hipMemcpyAsync(base + (size_t)c * len, Xn, len * sizeof(double), hipMemcpyDeviceToDevice, st));
                return 0;
            };
            GMC_TRY(sw(X, nn));
            GMC_TRY(sw(A, nn));
            GMC_TRY(sw(lamv, (size_t)n));
            std::swap(prob[a], prob[c]);
            return 0;
        };
        int rc = 0;
        for (int it = 0; it < n_max; it++) {
            GMC_TRY(oa_refine(st, n, nact, wmode, A, X, Xn, lamv, axbuf, Sb, Gb, partial, npart, delta, op, X32, E32,
                              P32, half_from, dscale, pe, cluster_eig));
            // the converged problems' eigenvectors into the new X buffer
            if (nact < B)
                GMC_HIP(hipMemcpyAsync(X + (size_t)nact * nn, Xn + (size_t)nact * nn, (size_t)(B - nact) * nn *
                                       sizeof(double), hipMemcpyDeviceToDevice, st));
            GMC_TRY(report(what, it, nact));
            if (stats && atoi(getenv("GMC_OA_STATS")) >= 2) {
                // every refined problem's correction of the step
                fprintf(stderr, "[gmc] OA %s step %d per problem:", what, it);
                for (int b = 0; b < nact; b++) fprintf(stderr, " %d:%.1e", prob[b], pe[b]);
                fprintf(stderr, "\n");
            }
            const float leave = std::min(OA_LEAVE, oa_extend);
            const OaDecision d =
                oa_decide(pe, it, n_min, n_max, OA_TOL, oa_extend, OA_GIVEUP, resolve != nullptr, leave);
            if (d.next == OaNext::Continue) {
                if (it + 1 >= n_min) {
                    // converged problems leave; the others move to the front
                    int lo = 0, hi = nact - 1;
                    while (true) {
                        while (lo <= hi && !(pe[lo] < leave)) lo++;
                        while (hi >= lo && pe[hi] < leave) hi--;
                        if (lo >= hi) break;
                        GMC_TRY(swap_slots(lo, hi));
                        std::swap(pe[lo], pe[hi]);
                    }
                    nact = lo;
                }
                continue;
            }
            if (d.next == OaNext::Resolve)
                for (int sl : d.left) left.push_back(prob[sl]);
            if (d.next == OaNext::Fail) {
                set_error("ERPA mixed: refinement of %s left %d problem(s) (first %d) with non-finite corrections, "
                          "more than %u close pairs or a cluster of more than half its columns after %d steps; "
                          "GMC_MIXED solves such problems again in FP64", what, (int)d.left.size(),
                          d.left.empty() ? -1 : prob[d.left[0]], op.cap, n_max);
                rc = -2;
            }
            break;
        }
        // the original order
        for (int sl = 0; sl < B; sl++)
            while (prob[sl] != sl) GMC_TRY(swap_slots(sl, prob[sl]));
        return rc;
    };
    // The listed problems solved again in FP64: their matrices side by side
    // in Xn (free after the refinement), eigenvectors and eigenvalues back
    // into X and lamv.
    auto resolve_left = [&](const char* what, const std::vector<int>& left, const double* A, double* lamv) -> int {
        if (left.empty()) return 0;
        const int nl = (int)left.size();
        for (int k = 0; k < nl; k++)
            GMC_HIP(hipMemcpyAsync(Xn + (size_t)k * nn, A + (size_t)left[k] * nn, nn * sizeof(double),
                                   hipMemcpyDeviceToDevice, st));
        GMC_TRY((*resolve)(Xn, wre, nl));
        for (int k = 0; k < nl; k++) {
            GMC_HIP(hipMemcpyAsync(X + (size_t)left[k] * nn, Xn + (size_t)k * nn, nn * sizeof(double),
                                   hipMemcpyDeviceToDevice, st));
            GMC_HIP(hipMemcpyAsync(lamv + (size_t)left[k] * n, wre + (size_t)k * n, n * sizeof(double),
                                   hipMemcpyDeviceToDevice, st));
        }
        if (stats) fprintf(stderr, "[gmc] OA %s: %d of %d problems solved again in FP64\n", what, nl, B);
        return 0;
    };
    std::vector<int> left;
    hipLaunchKernelGGL(zero_int_kernel, dim3((B + 255) / 256), dim3(256), 0, st, out.nneg, B);
    const dim3 gsq((n + 255) / 256, n, B), blk(256);
    const unsigned gcopy = (unsigned)std::min<int64_t>(65536, ((int64_t)B * nn + 255) / 256);

    const int nt64f = (n + 63) / 64;
    if (bt) {
        // the batching form: alphas are already on the device
        hipLaunchKernelGGL(form_pad_kernel<double>, dim3(nt64f * (nt64f + 1) / 2, bt->nsys), blk, 0, st, bt->sys,
                           alphas, nalpha, Ap, Am, n);
        hipLaunchKernelGGL(pad_diag_set_kernel, dim3(B), blk, 0, st, bt->sys, nalpha, Ap, n, 1.0);
    } else {
        GMC_HIP(hipMemcpyAsync(d_al, alphas, nalpha * sizeof(double), hipMemcpyHostToDevice, st));
        hipLaunchKernelGGL(form_kernel<double>, dim3(nt64f * (nt64f + 1) / 2), blk, 0, st, ab.p0, ab.p1, ab.m0,
                           ab.m1, d_al, B, Ap, Am, n);
    }
    // eig(A+) in FP32, refined in FP64
    hipLaunchKernelGGL(f64_to_f32_kernel, dim3(gcopy), blk, 0, st, Ap, E32, (int64_t)B * nn);
    GMC_HIP(hipGetLastError());
    GMC_TRY(syevd_batched<float>(st, ep, E32, w32));
    hipLaunchKernelGGL(f32_to_f64_kernel, dim3(gcopy), blk, 0, st, E32, X, (int64_t)B * nn);
    GMC_HIP(hipGetLastError());
    const int nt64 = (n + 63) / 64;
    // The Cholesky route: L L^T = |A+| into AX.  ok stays false when the
    // restricted refinement does not reach OAC_TOL in OAC_MAX_STEPS steps or a
    // pivot is not positive; the eigen route below then takes over from the
    // (partly refined) X.
    auto chol_route = [&](bool& ok) -> int {
        ok = false;
        ProfScope ps(st, "mixed_chol_route", 0.0);
        hipLaunchKernelGGL(count_neg_kernel, dim3((B + 255) / 256), dim3(256), 0, st, w32, n, B, info);
        std::vector<int> hc(B);
        GMC_HIP(hipMemcpyAsync(hc.data(), info, B * sizeof(int), hipMemcpyDeviceToHost, st));
        GMC_HIP(hipStreamSynchronize(st));
        const int maxneg = *std::max_element(hc.begin(), hc.end());
        const int K = std::min(n, (maxneg + 8 + 15) / 16 * 16);
        const int64_t snn = (int64_t)nn;
        GemmShape gAX{n, K, n, false, false, n, n, n, snn, snn, snn, B, 0};
        GemmShape gTX{n, K, n, true, false, n, n, n, snn, snn, snn, B, 0};
        const dim3 gk((n + 255) / 256, K, B);
        const unsigned gadd = (unsigned)std::min<int64_t>(4096, ((int64_t)n * K + 255) / 256);
        std::vector<unsigned> em(B);
        bool conv = false;
        for (int it = 0; it < OAC_MAX_STEPS && !conv; it++) {
            GMC_TRY(gemm_big<double>(st, gAX, 1.0, Ap, X, 0.0, AX));   // A X_J
            GMC_TRY(gemm_big<double>(st, gTX, 1.0, X, AX, 0.0, Sb));   // X^T A X_J
            GMC_TRY(gemm_big<double>(st, gTX, 1.0, X, X, 0.0, Gb));    // X^T X_J
            hipLaunchKernelGGL(oac_lam_kernel, dim3(B), dim3(64), 0, st, Sb, Gb, w32, n, K, lam, delta);
            hipLaunchKernelGGL(zero_int_kernel, dim3((B + 255) / 256), dim3(256), 0, st,
                               reinterpret_cast<int*>(op.emax), B);
            hipLaunchKernelGGL(oac_e_kernel, gk, blk, 0, st, Sb, Gb, lam, w32, delta, n, K, Xn, op.emax);
            GMC_TRY(gemm_big<double>(st, gAX, 1.0, X, Xn, 0.0, AX));   // X E_J
            hipLaunchKernelGGL(oac_add_kernel, dim3(gadd, B), blk, 0, st, X, AX, n, K);
            GMC_HIP(hipGetLastError());
            GMC_HIP(hipMemcpyAsync(em.data(), op.emax, B * sizeof(unsigned), hipMemcpyDeviceToHost, st));
            GMC_HIP(hipStreamSynchronize(st));
            float mx = 0.f;
            for (unsigned u : em) {
                float f;
                std::memcpy(&f, &u, 4);
                mx = std::max(mx, f);   // +inf stays +inf
            }
            if (stats) fprintf(stderr, "[gmc] Cholesky route: restricted step %d on %d columns, max |E| %.3e\n", it, K, mx);
            conv = mx < OAC_TOL;
        }
        if (!conv) return 0;
        // |A+| (lower triangle) into AX, then its Cholesky factor
        GMC_HIP(hipMemcpyAsync(AX, Ap, (size_t)B * nn * sizeof(double), hipMemcpyDeviceToDevice, st));
        hipLaunchKernelGGL(oac_reflect_kernel, gk, blk, 0, st, X, lam, n, K, Sb, out.nneg);
        GemmShape gR{n, n, K, false, true, n, n, n, snn, snn, snn, B, 1};
        GMC_TRY(gemm<double>(st, gR, 1.0, Sb, X, 1.0, AX));   // rank K: one lower-triangle launch
        GMC_HIP(hipGetLastError());
        GMC_TRY(potrf_batched(st, n, B, AX, pwork, info));
        GMC_HIP(hipMemcpyAsync(hc.data(), info, B * sizeof(int), hipMemcpyDeviceToHost, st));
        GMC_HIP(hipStreamSynchronize(st));
        int bad = 0;
        for (int v : hc) bad += v != 0;
        if (bad) {
            if (stats) fprintf(stderr, "[gmc] Cholesky route: %d of %d factorisations failed\n", bad, B);
            hipLaunchKernelGGL(zero_int_kernel, dim3((B + 255) / 256), dim3(256), 0, st, out.nneg, B);
            return 0;
        }
        ok = true;
        return 0;
    };
    bool chol_ok = false;
    const char* route = getenv("GMC_MIXED_ROUTE");
    if (!(route && strcmp(route, "eigen") == 0)) GMC_TRY(chol_route(chol_ok));
    if (!chol_ok) {
        // the eigen route of the reference: A+ refined in full, S = V|L|^(1/2)V^T (into AX)
        ProfScope ps(st, "mixed_eigen_route", 0.0);
        GMC_TRY(refine("A+", 1, Ap, lam, AX, left));
        GMC_TRY(resolve_left("A+", left, Ap, lam));
        hipLaunchKernelGGL(sqrt_scale64_kernel, gsq, blk, 0, st, X, lam, n, out.nneg);
        GMC_HIP(hipGetLastError());
        // S is symmetric: its lower triangle, mirrored
        GemmShape gS{n, n, n, false, true, n, n, n, (int64_t)nn, (int64_t)nn, (int64_t)nn, B, 1};
        GMC_TRY(gemm_big<double>(st, gS, 1.0, X, X, 0.0, AX));
        hipLaunchKernelGGL(mirror_lower_kernel<double>, dim3(nt64, nt64, B), blk, 0, st, AX, n);
    }
    // M = F^T A- F (into Ap) for the factor F in AX (L, or S = S^T).  With L
    // the products run over TRI_SLICES slices that skip its zero triangle:
    // A- L by column slices (columns [c0, c1) need rows c0.. of L), L^T Sb
    // by row slices of the lower triangle (rows [r0, r1) need rows r0.. of
    // Sb and columns up to r1), L Z by row slices (rows [r0, r1) need
    // columns up to r1 of L): 9/16, 15/64 and 9/16 of the full products.
    const int64_t snn = (int64_t)nn;
    constexpr int TRI_SLICES = 8;
    auto cut = [&](int p) { return (int)((int64_t)n * p / TRI_SLICES); };
    GemmShape gN{n, n, n, false, false, n, n, n, snn, snn, snn, B, 0};
    if (chol_ok) {
        for (int p = 0; p < TRI_SLICES; p++) {
            const int c0 = cut(p), c1 = cut(p + 1);
            if (c1 <= c0) continue;
            GemmShape g{n, c1 - c0, n - c0, false, false, n, n, n, snn, snn, snn, B, 0};
            GMC_TRY(gemm_big<double>(st, g, 1.0, Am + (size_t)c0 * n, AX + c0 + (size_t)c0 * n, 0.0,
                                     Sb + (size_t)c0 * n));
        }
        for (int p = 0; p < TRI_SLICES; p++) {
            const int r0 = cut(p), r1 = cut(p + 1);
            if (r1 <= r0) continue;
            GemmShape g{r1 - r0, r1, n - r0, true, false, n, n, n, snn, snn, snn, B, 0};
            GMC_TRY(gemm_big<double>(st, g, 1.0, AX + r0 + (size_t)r0 * n, Sb + r0, 0.0, Ap + r0));
        }
    } else {
        GMC_TRY(gemm_big<double>(st, gN, 1.0, Am, AX, 0.0, Sb));
        GemmShape gTl{n, n, n, true, false, n, n, n, snn, snn, snn, B, 1};
        GMC_TRY(gemm_big<double>(st, gTl, 1.0, AX, Sb, 0.0, Ap));   // M is symmetric: its lower triangle
    }
    hipLaunchKernelGGL(mirror_lower_kernel<double>, dim3(nt64, nt64, B), blk, 0, st, Ap, n);
    if (bt) hipLaunchKernelGGL(pad_diag_set_kernel, dim3(B), blk, 0, st, bt->sys, nalpha, Ap, n, -1.0);
    // eig(M) in FP32, refined in FP64 (Am is free now and serves as the AX buffer)
    hipLaunchKernelGGL(f64_to_f32_kernel, dim3(gcopy), blk, 0, st, Ap, E32, (int64_t)B * nn);
    GMC_HIP(hipGetLastError());
    GMC_TRY(syevd_batched<float>(st, ep, E32, w32));
    hipLaunchKernelGGL(f32_to_f64_kernel, dim3(gcopy), blk, 0, st, E32, X, (int64_t)B * nn);
    GMC_HIP(hipGetLastError());
    GMC_TRY(refine("M", 2, Ap, om2, Am, left));
    GMC_TRY(resolve_left("M", left, Ap, om2));
    if (stats && atoi(getenv("GMC_OA_STATS")) >= 2) {
        // per problem: modes the integrand keeps (omega^2 > 1e-6, 1e-3 < omega < 1e8)
        std::vector<double> h((size_t)B * n);
        GMC_HIP(hipMemcpyAsync(h.data(), om2, h.size() * sizeof(double), hipMemcpyDeviceToHost, st));
        GMC_HIP(hipStreamSynchronize(st));
        fprintf(stderr, "[gmc] M kept modes per problem:");
        for (int b = 0; b < B; b++) {
            int c = 0;
            for (int i = 0; i < n; i++) {
                const double o = h[(size_t)b * n + i];
                c += o > 1e-6 && sqrt(o) > 1e-3 && sqrt(o) < 1e8;
            }
            fprintf(stderr, " %d", c);
        }
        fprintf(stderr, "\n");
    }
    // Y = F Z (Sb), filtered copy (Gb), P = Yf Yf^T (Am), integrand
    if (chol_ok) {
        for (int p = 0; p < TRI_SLICES; p++) {
            const int r0 = cut(p), r1 = cut(p + 1);
            if (r1 <= r0) continue;
            GemmShape g{r1 - r0, n, r1, false, false, n, n, n, snn, snn, snn, B, 0};
            GMC_TRY(gemm_big<double>(st, g, 1.0, AX + r0, X, 0.0, Sb + r0));
        }
    } else {
        GMC_TRY(gemm_big<double>(st, gN, 1.0, AX, X, 0.0, Sb));
    }
    hipLaunchKernelGGL(ycols_kernel<double>, gsq, blk, 0, st, Sb, Gb, om2, out.eig, n);
    GemmShape gP{n, n, n, false, true, n, n, n, (int64_t)nn, (int64_t)nn, (int64_t)nn, B, 1};
    GMC_TRY(gemm_big<double>(st, gP, 1.0, Gb, Gb, 0.0, Am));
    if (bt) {
        hipLaunchKernelGGL(dot_batch_kernel<double>, dim3(npart, B), blk, 0, st, bt->sys, nalpha, Am, n, partial);
        hipLaunchKernelGGL(sum_partials_batch_kernel, dim3((B + 63) / 64), dim3(64), 0, st, partial, npart, bt->sys,
                           nalpha, out.W, B);
    } else {
        hipLaunchKernelGGL(dot_lower_kernel<double>, dim3(npart, B), blk, 0, st, Gw, Am, n, partial);
        hipLaunchKernelGGL(sum_partials_kernel, dim3(1), dim3(64), 0, st, partial, npart, wdiag, out.W, B);
    }
    GMC_HIP(hipGetLastError());
    return 0;
}

template int erpa_plan<double>(int, int, ErpaPlan<double>*);
template int erpa_plan<float>(int, int, ErpaPlan<float>*);
template int erpa_batched<double>(hipStream_t, ErpaPlan<double>&, const AbPair&, const double*, int, const double*,
                                  const double*, void*, ErpaOut&, double*, double*, double*, const ErpaBatchTables*);
template int erpa_batched<float>(hipStream_t, ErpaPlan<float>&, const AbPair&, const double*, int, const double*,
                                 const double*, void*, ErpaOut&, double*, double*, double*, const ErpaBatchTables*);

// ------------------------------------------------------------ integrand from given Y
// Columns of Y whose Eig lies in (1e-3, 1e8) (acfd.f:4976, 5013), others zero.
__global__ void filter_cols_kernel(const double* Y, const double* eig, double* Yf, int n)
{
    const int k = blockIdx.y;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double e = eig[k];
    const size_t at = i + (size_t)k * n;
    Yf[at] = (e > 1.0e-3 && e < 1.0e8) ? Y[at] : 0.0;
}

size_t integrand_from_y_bytes(int n) { return carve_bytes({(size_t)n * n * 8, (size_t)n * n * 8, 1024 * 8}); }

int integrand_from_y(hipStream_t st, int n, const double* Y, const double* eig, const double* Gw, const double* wdiag,
                     void* scratch, double* W)
{
    const size_t nn = (size_t)n * n;
    Arena ar{static_cast<char*>(scratch), integrand_from_y_bytes(n)};
    double* Yf = ar.take<double>(nn);
    double* P = ar.take<double>(nn);
    const int npart = 1024;
    double* partial = ar.take<double>(npart);
    GMC_CHECK(ar.fits(), "ACEneERPA: scratch layout overflow");
    hipLaunchKernelGGL(filter_cols_kernel, dim3((n + 255) / 256, n), dim3(256), 0, st, Y, eig, Yf, n);
    GemmShape gP{n, n, n, false, true, n, n, n, 0, 0, 0, 1, 1};
    GMC_TRY(gemm_big<double>(st, gP, 1.0, Yf, Yf, 0.0, P));
    hipLaunchKernelGGL(dot_lower_kernel<double>, dim3(npart, 1), dim3(256), 0, st, Gw, P, n, partial);
    hipLaunchKernelGGL(sum_partials_kernel, dim3(1), dim3(64), 0, st, partial, npart, wdiag, W, 1);
    GMC_HIP(hipGetLastError());
    return 0;
}

}  // namespace gmc
