// Batched FP64 Cholesky factorisation A = L L^T (right-looking, panels of
// CHOL_NB columns).  The mixed ERPA path factors |A+| with it (the reference
// takes |A+|^(1/2) from an eigendecomposition instead, interpa.f:1693-1716).
//
// Per panel k: one workgroup per problem factors the diagonal block in LDS,
// writes it back and inverts it in place (column by column from the right,
// each column a triangular product with the part already inverted); the
// panel below becomes L21 = A21 L11^-T (a product with the small inverse)
// and the trailing lower triangle takes A22 -= L21 L21^T, both on the
// library's own MFMA tile.  The update is split with look-ahead: the next
// panel's columns first, then the next diagonal block is factored on the
// caller's stream while the lower triangle past that panel is updated on a
// second stream (one lower-triangle launch).
#include "chol.hpp"
#include "gemm.hpp"

namespace gmc {

namespace {

constexpr int LDSW = CHOL_NB + 1;   // padded column of the LDS block
constexpr size_t DIAG_LDS = ((size_t)CHOL_NB * LDSW + 4 * CHOL_NB) * sizeof(double);
static_assert(CHOL_NB == 128, "potrf_diag_kernel maps 256 threads to 128 rows");

// Factor of the kb x kb diagonal block at (k, k), written back, and its
// inverse into Linv.  a[c][r] holds block element (r, c).  A problem with
// info != 0 is skipped.  Threads own rows: thread t works on row t % 128 and
// on the columns of parity t / 128, so no index is divided and a column's
// entries are read by consecutive threads.
__global__ __launch_bounds__(256) void potrf_diag_kernel(double* A, int n, int k, int kb, double* Linv, int* info)
{
    extern __shared__ double lds[];
    double (*a)[LDSW] = reinterpret_cast<double (*)[LDSW]>(lds);
    double* tmp = lds + CHOL_NB * LDSW;
    __shared__ int bad;
    const int b = blockIdx.x, t = threadIdx.x;
    const int r = t & (CHOL_NB - 1), half = t / CHOL_NB;   // blockDim.x == 2 * CHOL_NB
    if (info[b] != 0) return;
    double* Ab = A + (size_t)b * n * n;
    if (r < kb)
        for (int c = half; c < kb; c += 2) a[c][r] = r >= c ? Ab[(size_t)(k + c) * n + k + r] : 0.0;
    if (t == 0) bad = 0;
    __syncthreads();
    // two columns per step: the 2 x 2 pivot block from values every thread
    // reads alike (so every branch is uniform), columns j and j+1 of L, then
    // the rank-2 update of the block's remaining lower triangle
    int j = 0;
    for (; j + 1 < kb; j += 2) {
        const double p0 = a[j][j];
        if (!(p0 > 0.0 && p0 < INFINITY)) {
            if (t == 0) bad = k + j + 1;
            break;
        }
        const double l0 = sqrt(p0), l10 = a[j][j + 1] / l0;
        const double p1 = a[j + 1][j + 1] - l10 * l10;
        if (!(p1 > 0.0 && p1 < INFINITY)) {
            if (t == 0) bad = k + j + 2;
            break;
        }
        const double l1 = sqrt(p1);
        const double x0 = r < kb ? a[j][r] : 0.0, x1 = r < kb ? a[j + 1][r] : 0.0;
        const double L0 = r == j ? l0 : x0 / l0;
        const double L1 = r == j + 1 ? l1 : (x1 - L0 * l10) / l1;
        __syncthreads();
        if (half == 0 && r >= j && r < kb) a[j][r] = L0;
        if (half == 1 && r >= j + 1 && r < kb) a[j + 1][r] = L1;
        __syncthreads();
        if (r > j + 1 && r < kb)
            for (int c = j + 2 + half; c <= r; c += 2) a[c][r] -= L0 * a[j][c] + L1 * a[j + 1][c];
        __syncthreads();
    }
    if (j == kb - 1 && !bad) {
        // an odd block's last column
        const double p = a[j][j];
        if (!(p > 0.0 && p < INFINITY)) {
            if (t == 0) bad = k + j + 1;
        } else {
            __syncthreads();
            if (t == 0) a[j][j] = sqrt(p);
        }
    }
    __syncthreads();
    if (bad) {
        if (t == 0) info[b] = bad;
        return;
    }
    if (r < kb)
        for (int c = half; c < kb; c += 2) Ab[(size_t)(k + c) * n + k + r] = r >= c ? a[c][r] : 0.0;
    // in-place inverse, columns from the right: column j of L^-1 below the
    // diagonal is -(1/l_jj) Linv[j+1:, j+1:] l[j+1:, j].  Two columns per
    // step: one pass forms both sums over the columns already inverted,
    //   S_j(r) = sum_{q>j} Linv(r,q) l(q,j),  S_i(r) = sum_{q>j} Linv(r,q) l(q,i),
    // i = j - 1; then Linv(r,j) = -S_j(r)/l_jj and Linv(r,i) =
    // -(Linv(r,j) l(j,i) + S_i(r))/l_ii.  The two threads of a row split the
    // sums by parity and combine through tmp2/tmp4 (the lone column of an
    // odd block takes the one-column form)
    double* tmp2 = tmp + CHOL_NB;
    double* tmp3 = tmp2 + CHOL_NB;
    double* tmp4 = tmp3 + CHOL_NB;
    j = kb - 1;
    for (; j >= 1; j -= 2) {
        const int i = j - 1;
        __syncthreads();
        const double djj = 1.0 / a[j][j], dii = 1.0 / a[i][i], lji = a[i][j];
        if (half == 0 && r > j && r < kb) tmp[r] = a[j][r];
        if (half == 1 && r > j && r < kb) tmp3[r] = a[i][r];
        __syncthreads();
        double sj = 0.0, si = 0.0;
        if (r > j && r < kb)
            for (int q = j + 1 + half; q <= r; q += 2) {
                const double v = a[q][r];
                sj += v * tmp[q];
                si += v * tmp3[q];
            }
        if (half == 1) {
            tmp2[r] = sj;
            tmp4[r] = si;
        }
        __syncthreads();
        if (half == 0 && r > j && r < kb) {
            const double lrj = -djj * (sj + tmp2[r]);
            a[j][r] = lrj;
            a[i][r] = -dii * (lrj * lji + si + tmp4[r]);
        }
        if (t == 0) {
            a[j][j] = djj;
            a[i][j] = -dii * djj * lji;
            a[i][i] = dii;
        }
    }
    if (j == 0) {
        __syncthreads();
        const double d00 = 1.0 / a[0][0];
        if (half == 0 && r > 0 && r < kb) tmp[r] = a[0][r];
        __syncthreads();
        double s = 0.0;
        if (r > 0 && r < kb)
            for (int q = 1 + half; q <= r; q += 2) s += a[q][r] * tmp[q];
        if (half == 1) tmp2[r] = s;
        __syncthreads();
        if (half == 0 && r > 0 && r < kb) a[0][r] = -d00 * (s + tmp2[r]);
        if (t == 0) a[0][0] = d00;
    }
    __syncthreads();
    double* Lb = Linv + (size_t)b * CHOL_NB * CHOL_NB;
    if (r < kb)
        for (int c = half; c < kb; c += 2) Lb[(size_t)c * CHOL_NB + r] = r >= c ? a[c][r] : 0.0;
}

// P = A21: rows k + kb .. n-1, columns k .. k + kb - 1 (P: ld n, stride n CHOL_NB).
__global__ void potrf_copy_panel_kernel(const double* A, int n, int k, int kb, double* P)
{
    const int b = blockIdx.y, m = n - k - kb;
    const double* Ab = A + (size_t)b * n * n;
    double* Pb = P + (size_t)b * n * CHOL_NB;
    for (int idx = blockIdx.x * blockDim.x + threadIdx.x; idx < m * kb; idx += gridDim.x * blockDim.x) {
        const int r = idx % m, c = idx / m;
        Pb[(size_t)c * n + r] = Ab[(size_t)(k + c) * n + k + kb + r];
    }
}

__global__ void zero_strict_upper_kernel(double* A, int n)
{
    const int b = blockIdx.y;
    double* Ab = A + (size_t)b * n * n;
    const int64_t nn = (int64_t)n * n;
    for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < nn; idx += (int64_t)gridDim.x * blockDim.x) {
        const int r = (int)(idx % n), c = (int)(idx / n);
        if (r < c) Ab[idx] = 0.0;
    }
}

__global__ void zero_info_kernel(int* info, int B)
{
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b < B) info[b] = 0;
}

// The look-ahead's second stream and events, made on a device's first
// factorisation and kept until potrf_release
constexpr int CHOL_MAX_DEV = 16;
bool ready[CHOL_MAX_DEV] = {};
hipStream_t s2s[CHOL_MAX_DEV];
hipEvent_t evn[CHOL_MAX_DEV], evr[CHOL_MAX_DEV];

}  // namespace

void potrf_release()
{
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= CHOL_MAX_DEV || !ready[dev]) return;
    (void)hipStreamSynchronize(s2s[dev]);
    (void)hipEventDestroy(evn[dev]);
    (void)hipEventDestroy(evr[dev]);
    (void)hipStreamDestroy(s2s[dev]);
    ready[dev] = false;
}

size_t potrf_work_bytes(int n, int B)
{
    return ((size_t)B * CHOL_NB * CHOL_NB + (size_t)B * n * CHOL_NB) * sizeof(double);
}

int potrf_batched(hipStream_t s, int n, int B, double* A, void* work, int* info)
{
    GMC_CHECK(n >= 1 && B >= 1, "potrf: n=%d, batch=%d", n, B);
    double* Linv = static_cast<double*>(work);
    double* P = Linv + (size_t)B * CHOL_NB * CHOL_NB;
    const int64_t nn = (int64_t)n * n;
    // look-ahead: the trailing update past the next panel runs on a second
    // stream (default priority) while the next diagonal block is factored;
    // the stream and its two events are made once per device
    int dev = 0;
    GMC_HIP(hipGetDevice(&dev));
    GMC_CHECK(dev >= 0 && dev < CHOL_MAX_DEV, "potrf: device %d out of range", dev);
    if (!ready[dev]) {
        GMC_HIP(hipFuncSetAttribute((const void*)potrf_diag_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                                    (int)DIAG_LDS));
        GMC_HIP(hipStreamCreateWithFlags(&s2s[dev], hipStreamNonBlocking));
        GMC_HIP(hipEventCreateWithFlags(&evn[dev], hipEventDisableTiming));
        GMC_HIP(hipEventCreateWithFlags(&evr[dev], hipEventDisableTiming));
        ready[dev] = true;
    }
    hipStream_t s2 = s2s[dev];
    hipEvent_t ev_next = evn[dev], ev_rest = evr[dev];
    hipStreamCaptureStatus cst = hipStreamCaptureStatusNone;
    GMC_HIP(hipStreamIsCapturing(s, &cst));
    const bool ahead = cst == hipStreamCaptureStatusNone;   // (a capture keeps to one stream)
    bool rest_pending = false;
    hipLaunchKernelGGL(zero_info_kernel, dim3((B + 255) / 256), dim3(256), 0, s, info, B);
    hipLaunchKernelGGL(potrf_diag_kernel, dim3(B), dim3(256), DIAG_LDS, s, A, n, 0, std::min(CHOL_NB, n), Linv,
                       info);
    GMC_HIP(hipGetLastError());
    for (int k = 0; k < n; k += CHOL_NB) {
        // the diagonal block at k is factored (and inverted into Linv)
        const int kb = std::min(CHOL_NB, n - k), m = n - k - kb;
        if (m == 0) break;
        const unsigned gp = (unsigned)std::min<int64_t>(4096, ((int64_t)m * kb + 255) / 256);
        hipLaunchKernelGGL(potrf_copy_panel_kernel, dim3(gp, B), dim3(256), 0, s, A, n, k, kb, P);
        GMC_HIP(hipGetLastError());
        double* A21 = A + (k + kb) + (size_t)k * n;
        // L21 = A21 L11^-T
        GemmShape gp21{m, kb, kb, false, true, n, CHOL_NB, n, (int64_t)n * CHOL_NB, (int64_t)CHOL_NB * CHOL_NB, nn, B,
                       0};
        GMC_TRY(gemm<double>(s, gp21, 1.0, P, Linv, 0.0, A21));
        // A22 -= L21 L21^T: the next panel's columns first (all m rows, so
        // the strict upper part of its diagonal block takes garbage that the
        // final clear removes), then the lower triangle past them
        const int kb2 = std::min(CHOL_NB, m), m2 = m - kb2;
        if (rest_pending) GMC_HIP(hipStreamWaitEvent(s, ev_rest, 0));   // the columns the last rest updated
        rest_pending = false;
        GemmShape gnx{m, kb2, kb, false, true, n, n, n, nn, nn, nn, B, 0};
        GMC_TRY(gemm<double>(s, gnx, -1.0, A21, A21, 1.0, A21 + (size_t)kb * n));
        if (m2 > 0) {
            double* A21r = A21 + kb2;
            GemmShape g22{m2, m2, kb, false, true, n, n, n, nn, nn, nn, B, 1};
            if (ahead) {
                GMC_HIP(hipEventRecord(ev_next, s));
                GMC_HIP(hipStreamWaitEvent(s2, ev_next, 0));
                GMC_TRY(gemm<double>(s2, g22, -1.0, A21r, A21r, 1.0, A21r + (size_t)(kb + kb2) * n));
                GMC_HIP(hipEventRecord(ev_rest, s2));
                rest_pending = true;
            } else {
                GMC_TRY(gemm<double>(s, g22, -1.0, A21r, A21r, 1.0, A21r + (size_t)(kb + kb2) * n));
            }
        }
        // the next diagonal block, while the rest of the update runs
        hipLaunchKernelGGL(potrf_diag_kernel, dim3(B), dim3(256), DIAG_LDS, s, A, n, k + kb, kb2, Linv, info);
        GMC_HIP(hipGetLastError());
    }
    if (rest_pending) GMC_HIP(hipStreamWaitEvent(s, ev_rest, 0));
    const unsigned gz = (unsigned)std::min<int64_t>(65536, (nn + 255) / 256);
    hipLaunchKernelGGL(zero_strict_upper_kernel, dim3(gz, B), dim3(256), 0, s, A, n);
    GMC_HIP(hipGetLastError());
    return 0;
}

}  // namespace gmc
