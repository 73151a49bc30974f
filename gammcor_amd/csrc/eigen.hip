// Batched symmetric eigensolver (see eigen.hpp for the structure).
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdlib>
#include <functional>
#include <limits>

#include "eigen.hpp"
#include "gemm_tile.hpp"
#include "prof.hpp"

namespace gmc {

template <class T> struct Eps;
template <> struct Eps<double> { static constexpr double v = 2.220446049250313e-16; };
template <> struct Eps<float> { static constexpr float v = 1.1920929e-07f; };

__device__ inline double wave_sum(double v)
{
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}
__device__ inline float wave_sum(float v)
{
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}
// The sums over the 64 lanes of a wave of KV values per lane (KV a power of
// two, at most 64).  The values are first halved between lane pairs (a lane
// keeps one half of its columns, adds its partner's share of them, and sends
// the other half), so the KV sums take KV - 1 + 6 - log2(KV) exchanges
// instead of 6 KV.  Returns the sum of column lane >> (6 - log2 KV): every
// lane of each group of 64 / KV consecutive lanes holds its group's column.
// The value of lane ^ O (O < 32 by ds_swizzle's xor mode, which takes the
// pattern as an immediate and needs no address register).
template <int O>
__device__ inline float xor_lane(float v)
{
    if constexpr (O == 32) return __shfl_xor(v, 32, 64);
    else return __int_as_float(__builtin_amdgcn_ds_swizzle(__float_as_int(v), (O << 10) | 0x1f));
}
template <int O>
__device__ inline double xor_lane(double v)
{
    if constexpr (O == 32) {
        return __shfl_xor(v, 32, 64);
    } else {
        const long long b = __double_as_longlong(v);
        const int lo = __builtin_amdgcn_ds_swizzle((int)(b & 0xffffffff), (O << 10) | 0x1f);
        const int hi = __builtin_amdgcn_ds_swizzle((int)(b >> 32), (O << 10) | 0x1f);
        return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
    }
}
// One halving step (H values per lane kept, exchanged across lane bit O),
// then the next; below one value per lane, plain xor sums.
template <int H, int O, int KV, class T>
__device__ inline T wave_sum_step(T (&s)[KV], int lane)
{
    if constexpr (H >= 1) {
        const bool up = (lane & O) != 0;
#pragma unroll
        for (int k = 0; k < H; k++) {
            const T send = up ? s[k] : s[k + H];
            const T keep = up ? s[k + H] : s[k];
            s[k] = keep + xor_lane<O>(send);
        }
        return wave_sum_step<H / 2, O / 2>(s, lane);
    } else if constexpr (O >= 1) {
        s[0] += xor_lane<O>(s[0]);
        return wave_sum_step<0, O / 2>(s, lane);
    } else {
        return s[0];
    }
}
template <int KV, class T>
__device__ inline T wave_sum_cols(T (&s)[KV], int lane)
{
    return wave_sum_step<KV / 2, 32>(s, lane);
}
template <int KV>
constexpr int ilog2()
{
    if constexpr (KV <= 1) return 0;
    else return 1 + ilog2<KV / 2>();
}

template <class T>
__device__ inline T wave_prod(T v)
{
    for (int o = 32; o > 0; o >>= 1) v *= __shfl_xor(v, o, 64);
    return v;
}
template <class T>
__device__ inline T wave_max(T v)
{
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
    return v;
}

// Sum over the threads of a workgroup (blockDim.x / 64 waves, at most 16);
// every thread gets the result.
template <class T>
__device__ T block_sum(T v, T* red)
{
    v = wave_sum(v);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
    __syncthreads();
    if (lane == 0) red[wave] = v;
    __syncthreads();
    T s = 0;
    for (int w = 0; w < nw; w++) s += red[w];
    return s;
}
// Two such sums at once, over one pair of barriers (red: 2 x the waves);
// each comes out as block_sum would give it.
template <class T>
__device__ void block_sum2(T& a, T& b, T* red)
{
    a = wave_sum(a);
    b = wave_sum(b);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
    __syncthreads();
    if (lane == 0) {
        red[wave] = a;
        red[nw + wave] = b;
    }
    __syncthreads();
    T sa = 0, sb = 0;
    for (int w = 0; w < nw; w++) {
        sa += red[w];
        sb += red[nw + w];
    }
    a = sa;
    b = sb;
}
template <class T>
__device__ T block_max(T v, T* red)
{
    v = wave_max(v);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) red[wave] = v;
    __syncthreads();
    return fmax(fmax(red[0], red[1]), fmax(red[2], red[3]));
}

// Barrier among the G workgroups that cooperate on one problem.  Each
// workgroup publishes its epoch in a flag of its own (64 bytes apart, no
// contended atomic); threads 0..G-1 each wait for one flag.  Release fence
// before the publish, acquire fence after the wait (agent scope: the
// workgroups may sit on different XCDs with separate L2 caches).  A bounded
// spin turns a lost arrival into an error flag instead of a hang.
constexpr int BAR_STRIDE = 16;   // unsigned words between two flags
__device__ inline void group_barrier(unsigned* flags, int gi, int G, unsigned& epoch, int* err)
{
    __syncthreads();
    epoch += 1;
    if (threadIdx.x == 0) {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        __hip_atomic_store(&flags[gi * BAR_STRIDE], epoch, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    if ((int)threadIdx.x < G) {
        const unsigned* f = &flags[threadIdx.x * BAR_STRIDE];
        long spins = 0;
        while (__hip_atomic_load(f, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < epoch) {
            if (__hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) break;
            if (++spins > (1L << 22)) {   // seconds of polling: a lost workgroup
                atomicExch(err, 1);
                break;
            }
            __builtin_amdgcn_s_sleep(1);
        }
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    }
    __syncthreads();
}

// ||A_b||_F of each column-major n x n problem (both triangles, as the
// matrix arrives), one workgroup of 256 threads per problem: every thread
// sums a fixed set of elements in a fixed order, the workgroup sum follows a
// fixed tree, so the result has the same bits on every call.
template <class T>
__global__ void frob_kernel(const T* A, int n, T* out)
{
    __shared__ double part[256];
    const T* a = A + (size_t)blockIdx.x * n * n;
    const int64_t nn = (int64_t)n * n;
    double acc = 0.0;
    for (int64_t i = threadIdx.x; i < nn; i += 256) acc += (double)a[i] * (double)a[i];
    part[threadIdx.x] = acc;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h) part[threadIdx.x] += part[threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[blockIdx.x] = (T)sqrt(part[0]);
}

namespace {
std::atomic<int> g_det{-1};   // -1: not read yet from GMC_DETERMINISTIC
}

bool deterministic()
{
    int v = g_det.load();
    if (v < 0) {
        const char* e = getenv("GMC_DETERMINISTIC");
        v = (e && atoi(e) != 0) ? 1 : 0;
        int expect = -1;
        g_det.compare_exchange_strong(expect, v);
        v = g_det.load();
    }
    return v != 0;
}

void set_deterministic(bool on) { g_det.store(on ? 1 : 0); }

__global__ void zero_u32(unsigned* p, int n)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0u;
}

// ====================================================================== sytrd
template <class T>
struct PanelArgs {
    T *A, *W, *X, *d, *e, *tau, *part;
    T* Pu;   // B x G x n: per-workgroup partial sums of the lower-triangle SYMV
    T* vg;   // B x G x n: per-workgroup copy of v when it does not fit in LDS
    unsigned* bar;
    int* err;
    const T* fro;   // deterministic mode: Frobenius norm of each problem's matrix
    int n, k, kb, nb, G;
    unsigned epoch0;
    int vec;   // 16-byte SYMV path: n a multiple of 16 / sizeof(T) and A 16-byte aligned
};

// Rows owned by workgroup gi: one unit of SY_OB rows per round of G units,
// taken in snake order (round t owns unit tG + gi for even t, tG + G-1-gi for
// odd t).  The lower-triangle SYMV costs n - r per row r and the workgroups
// wait for the slowest at every group barrier, so the units are small: at
// column j the unit the sweep is inside is the only part of a round that is
// owned unevenly, which bounds the imbalance by SY_OB (n - j) elements.
// Slot q -> row; rows grow with the slot and row >= q.
constexpr int SY_OB = 8;
__device__ inline int owned_unit(int t, int gi, int G) { return t * G + ((t & 1) ? G - 1 - gi : gi); }
__device__ inline int owned_row(int q, int gi, int G) { return SY_OB * owned_unit(q / SY_OB, gi, G) + q % SY_OB; }
// Number of slots (SY_OB per owned unit) of workgroup gi over nunit units.
__device__ inline int owned_slots(int nunit, int gi, int G)
{
    const int T = nunit / G, rem = nunit - T * G;
    const int off = (T & 1) ? G - 1 - gi : gi;
    return SY_OB * (T + (off < rem ? 1 : 0));
}
// First slot of the first owned unit that reaches row `row`.
__device__ inline int first_slot(int row, int gi, int G)
{
    // unit(t) lies in [tG, tG + G - 1]: start a round early and step forward
    int t = (row - (SY_OB - 1)) / (SY_OB * G) - 1;
    if (t < 0) t = 0;
    while (SY_OB * owned_unit(t, gi, G) + SY_OB - 1 < row) t++;
    return SY_OB * t;
}

constexpr int SY_THREADS = 1024, SY_WAVES = SY_THREADS / 64;

// LDS atomic add (ds_add_f64 / ds_add_f32 on gfx950)
__device__ inline void lds_add(double* a, double v) { __hip_atomic_fetch_add(a, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP); }
__device__ inline void lds_add(float* a, float v) { __hip_atomic_fetch_add(a, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP); }

// Deterministic mode: the SYMV's scatter sums arrive in any order, so they
// are accumulated as fixed-point integers (an integer sum has the same bits
// in any order).  Every partial sum of |A_rc v_c| is at most ||A||_F ||v||_2
// (Cauchy-Schwarz on the trailing matrix, an orthogonal transform of the
// start), so with the scale s = 2^29 / (||A||_F ||v||_2) (FP32, int32) or
// 2^60 / (...) (FP64, int64) no sum overflows, and each term is rounded to
// 2^-29 (2^-60) of that bound: below FP32's (FP64's) own rounding of the
// floating-point sum.  The integers share the float slot of the same width.
__device__ inline void det_add(float* a, float v, float sc)
{
    __hip_atomic_fetch_add(reinterpret_cast<int*>(a), __float2int_rn(v * sc), __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_WORKGROUP);
}
__device__ inline void det_add(double* a, double v, double sc)
{
    __hip_atomic_fetch_add(reinterpret_cast<long long*>(a), __double2ll_rn(v * sc), __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_WORKGROUP);
}
template <bool DET, class T>
__device__ inline void scat(T* a, T v, T sc)
{
    if constexpr (DET) det_add(a, v, sc);
    else lds_add(a, v);
}
__device__ inline float det_value(const float* a, float sc) { return (float)*reinterpret_cast<const int*>(a) / sc; }
__device__ inline double det_value(const double* a, double sc)
{
    return (double)*reinterpret_cast<const long long*>(a) / sc;
}
template <class T>
constexpr T det_range() { return sizeof(T) == 8 ? T(1152921504606846976.0) : T(536870912.0); }   // 2^60, 2^29

// VLDS: keep v in LDS (n <= a few thousand); otherwise in a per-workgroup
// global copy.  The SYMV reads only the lower triangle of the trailing
// matrix: column r gives the dot sum_{c>=r} A[c,r] v[c] for y_r and scatters
// A[c,r] v[r] (c > r) into an LDS accumulator u, whose per-workgroup
// partials are summed after the next group barrier.
// 16-byte aligned vector of T (global_load_dwordx4 / ds_read_b128)
template <class T>
struct alignas(16) Vec16 {
    T x[16 / sizeof(T)];
};
template <class T>
__device__ inline Vec16<T> ld16(const T* p) { return *reinterpret_cast<const Vec16<T>*>(p); }
// 16-byte buffer load (buffer_load_dwordx4 ... offen): the descriptor lives in
// SGPRs, the row offset in one VGPR and the column offset in an SGPR, so the
// columns of a group share one address register.  Word 3 0x00020000 is the
// gfx9 raw-buffer format (32-bit data, untyped).
constexpr int BUF_WORD3 = 0x00020000;
template <class T>
__device__ inline Vec16<T> bld16(__amdgpu_buffer_rsrc_t r, int voff, int soff)
{
    return __builtin_bit_cast(Vec16<T>, __builtin_amdgcn_raw_buffer_load_b128(r, voff, soff, 0));
}

template <class T, bool VLDS, bool PLDS, int KV, int UV, bool DET>
__global__ __launch_bounds__(SY_THREADS) void sytrd_panel_kernel(PanelArgs<T> g)
{
    // columns per wave in the SYMV: FP32 halves the HBM time per element, so
    // it groups columns to keep the LDS atomic unit off the critical path
    constexpr int KC = sizeof(T) == 8 ? 1 : 4;
    // 16-byte path (n a multiple of the vector width): KV columns per wave,
    // UV row chunks of 64 vectors in flight per lane
    constexpr int VN = 16 / sizeof(T);
    extern __shared__ char smem[];
    const int n = g.n, k = g.k, kb = g.kb, nb = g.nb, G = g.G;
    // Workgroups are dispatched round-robin over the 8 XCDs.  When G divides
    // the workgroups per XCD, number them XCD by XCD so the G workgroups of a
    // problem share one XCD (and its L2) for their barriers and partial sums.
    const int nwg = (int)gridDim.x;
    const int lin = (nwg % 8 == 0 && (nwg / 8) % G == 0) ? (int)(blockIdx.x & 7) * (nwg / 8) + (int)(blockIdx.x >> 3)
                                                      : (int)blockIdx.x;
    const int p = lin / G, gi = lin % G;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int nunit = (n + SY_OB - 1) / SY_OB;
    const int nslots = owned_slots(nunit, gi, G);
    const int nsm = SY_OB * ((nunit + G - 1) / G);   // slots of the largest workgroup

    T* A = g.A + (size_t)p * n * n;
    T* W = g.W + (size_t)p * n * nb;   // W'(r, c) = W(r, c) + a2(c) V(r, c), the corrected panel
    T* X = g.X + (size_t)p * n * 3 * nb;
    T* part = g.part + (size_t)p * (2 + G * (2 + 2 * nb));
    T* p_alpha = part;
    T* p_ydot = part + 1;                 // dot part of y at row j+1 (its owner)
    T* p_norm = part + 2;
    T* p_dot = part + 2 + G;
    T* p_vy = part + 2 + G + 2 * nb * G;  // v^T y partials
    unsigned* bar = g.bar + (size_t)p * G * BAR_STRIDE;
    int* err = g.err;

    T* u = reinterpret_cast<T*>(smem);             // n
    T* v = VLDS ? u + n : g.vg + ((size_t)p * G + gi) * n;
    T* y = VLDS ? u + 2 * n : u + n;                // nsm
    T* Pu = g.Pu + (size_t)p * G * n;
    T* Vp = y + nsm;                                // PLDS: V(r, c) of the owned rows, nb x nsm
    T* Wp = Vp + (PLDS ? (size_t)nb * nsm : 0);     // PLDS: W'(r, c) of the owned rows
    T* Ar = Wp + (PLDS ? (size_t)nb * nsm : 0);     // V(j+1, c), nb
    T* Wr = Ar + nb;                                // W'(j+1, c), nb
    T* a2 = Wr + nb;
    T* t12 = a2 + nb;                               // 2 nb
    T* red = t12 + 2 * nb;                          // 2 SY_WAVES
    T* sc = red + 2 * SY_WAVES;                     // [0]=tau [1]=scal [2]=fixed-point scale (DET)

    // panel entries of an owned row (slot q, row r): LDS copies or the global panel
    auto vpan = [&](int c, int q, int r) -> T {
        if constexpr (PLDS) return Vp[(size_t)c * nsm + q];
        else return A[r + (size_t)(k + c) * n];
    };
    auto wpan = [&](int c, int q, int r) -> T {
        if constexpr (PLDS) return Wp[(size_t)c * nsm + q];
        else return W[r + (size_t)c * n];
    };

    // Column k arrives updated by the trailing product: its diagonal, alpha
    // and norm partial.  Every later column is updated and normed at the end
    // of the column before it.
    {
        T nrm = 0;
        for (int q = first_slot(k, gi, G) + tid; q < nslots; q += SY_THREADS) {
            const int r = owned_row(q, gi, G);
            if (r < k || r >= n) continue;
            const T a = A[r + (size_t)k * n];
            if (r == k) g.d[(size_t)p * n + k] = a;
            if (r == k + 1) *p_alpha = a;
            if (r >= k + 2) nrm += a * a;
        }
        nrm = block_sum(nrm, red);
        if (tid == 0) p_norm[gi] = nrm;
    }

    // Two group barriers per column:
    //   1. the column's norm partials and its updated entries are published:
    //      reflector, v, SYMV, panel dots and the v^T y partial;
    //   2. the SYMV partials are published: every workgroup forms the panel
    //      coefficients t1, t2, a2 and row j+1 of W' redundantly, then its own
    //      rows of y, W' and V, and the update and norm of column j+1.
    unsigned epoch = g.epoch0;
    for (int i = 0; i < kb; i++) {
        const int j = k + i;
        group_barrier(bar, gi, G, epoch, err);

        // reflector: the norm partials and the column are loaded together,
        // the column scaled once tau is known
        {
            const T xpart = tid < G ? p_norm[tid] : T(0);
            for (int c0 = j + 2 + tid; c0 < n; c0 += 8 * SY_THREADS) {
                T tmp[8];
#pragma unroll
                for (int e = 0; e < 8; e++) {
                    const int c = c0 + e * SY_THREADS;
                    tmp[e] = c < n ? A[c + (size_t)j * n] : T(0);
                }
#pragma unroll
                for (int e = 0; e < 8; e++) {
                    const int c = c0 + e * SY_THREADS;
                    if (c < n) v[c] = tmp[e];
                }
            }
            // the G norm partials: one wave's sum when they fit in it (the
            // same value block_sum gives, without its two barriers)
            const T xn2 = G <= 64 ? (wave == 0 ? wave_sum(xpart) : T(0)) : block_sum(xpart, red);
            if (tid == 0) {
                const T alpha = *p_alpha;
                T tau, beta, scal;
                if (xn2 == T(0)) {
                    tau = 0;
                    beta = alpha;
                    scal = 0;
                } else {
                    beta = -copysign(sqrt(alpha * alpha + xn2), alpha);
                    tau = (beta - alpha) / beta;
                    scal = T(1) / (alpha - beta);
                }
                sc[0] = tau;
                sc[1] = scal;
                if constexpr (DET) {
                    // ||v||_2 with v(j+1) = 1 and the rest scal * x
                    const T bound = g.fro[p] * sqrt(T(1) + scal * scal * xn2);
                    sc[2] = bound > T(0) && bound < T(INFINITY) ? det_range<T>() / bound : T(1);
                }
                if (gi == 0) {
                    g.tau[(size_t)p * n + j] = tau;
                    g.e[(size_t)p * n + j] = beta;
                }
            }
            __syncthreads();
        }
        const T tau = sc[0], scal = sc[1];
        const T dsc = DET ? sc[2] : T(0);
        for (int c = j + 1 + tid; c < n; c += SY_THREADS) {
            v[c] = c == j + 1 ? T(1) : scal * v[c];
            u[c] = T(0);
        }
        for (int c = tid; c < 2 * i; c += SY_THREADS) t12[c] = T(0);
        __syncthreads();

        // SYMV on the lower triangle: dot part -> y, scatter part -> u (LDS atomics)
        if (g.vec) {
            // 16-byte path: a wave takes KV consecutive owned columns; lane l
            // reads rows c..c+VN-1 (c = chunk + VN*l) of each, UV chunks in
            // flight.  The first chunk (diagonal block) is masked, the rest
            // are not: every row there lies below all KV columns.
            static_assert(SY_OB % KV == 0, "a column group must lie inside one unit");
            // groups in snake order over the waves: group work falls with the row
            for (int t = 0;; t++) {
                const int qg = t * SY_WAVES + ((t & 1) ? SY_WAVES - 1 - wave : wave);
                if (KV * qg >= nslots) break;
                const int r0 = owned_row(KV * qg, gi, G);
                if (r0 + KV - 1 <= j || r0 >= n) continue;
                T vk[KV], s[KV];
                const T* col[KV];
#pragma unroll
                for (int kk = 0; kk < KV; kk++) {
                    const int r = r0 + kk;
                    const bool okk = r > j && r < n;
                    vk[kk] = okk ? v[r] : T(0);
                    col[kk] = A + (size_t)(okk ? r : r0) * n;
                    s[kk] = 0;
                }
                const int base = r0 & ~63;
                {
                    const int c = base + VN * lane;
                    if (c < n) {
                        const Vec16<T> vv = ld16(v + c);
                        T w[VN];
#pragma unroll
                        for (int e = 0; e < VN; e++) w[e] = 0;
#pragma unroll
                        for (int kk = 0; kk < KV; kk++) {
                            const int r = r0 + kk;
                            if (r > j && r < n) {
                                const Vec16<T> a = ld16(col[kk] + c);
#pragma unroll
                                for (int e = 0; e < VN; e++) {
                                    const int row = c + e;
                                    if (row >= r) s[kk] += a.x[e] * vv.x[e];
                                    if (row > r) w[e] += a.x[e] * vk[kk];
                                }
                            }
                        }
#pragma unroll
                        for (int e = 0; e < VN; e++)
                            if (c + e > j && c + e > r0) scat<DET>(&u[c + e], w[e], dsc);
                    }
                }
                // rows below the diagonal block through one buffer descriptor
                // over the group's KV columns (KV n elements from column r0);
                // a column past n reads column r0 instead, as above
                const int r0s = __builtin_amdgcn_readfirstlane(r0);
                const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(
                    A + (size_t)r0s * n, 0, KV * n * (int)sizeof(T), BUF_WORD3);
                int soff[KV];
#pragma unroll
                for (int kk = 0; kk < KV; kk++) soff[kk] = (r0s + kk < n ? kk : 0) * n * (int)sizeof(T);
                int c = base + 64 * VN + VN * lane;
                for (; c + (UV - 1) * 64 * VN < n; c += UV * 64 * VN) {
                    Vec16<T> a[UV][KV], vv[UV];
#pragma unroll
                    for (int uu = 0; uu < UV; uu++) {
#pragma unroll
                        for (int kk = 0; kk < KV; kk++)
                            a[uu][kk] = bld16<T>(rs, (c + uu * 64 * VN) * (int)sizeof(T), soff[kk]);
                        vv[uu] = ld16(v + c + uu * 64 * VN);
                    }
#pragma unroll
                    for (int uu = 0; uu < UV; uu++) {
                        T w[VN];
#pragma unroll
                        for (int e = 0; e < VN; e++) w[e] = 0;
#pragma unroll
                        for (int kk = 0; kk < KV; kk++)
#pragma unroll
                            for (int e = 0; e < VN; e++) {
                                s[kk] += a[uu][kk].x[e] * vv[uu].x[e];
                                w[e] += a[uu][kk].x[e] * vk[kk];
                            }
#pragma unroll
                        for (int e = 0; e < VN; e++) scat<DET>(&u[c + uu * 64 * VN + e], w[e], dsc);
                    }
                }
                for (; c < n; c += 64 * VN) {
                    const Vec16<T> vv = ld16(v + c);
                    T w[VN];
#pragma unroll
                    for (int e = 0; e < VN; e++) w[e] = 0;
#pragma unroll
                    for (int kk = 0; kk < KV; kk++) {
                        const Vec16<T> a = bld16<T>(rs, c * (int)sizeof(T), soff[kk]);
#pragma unroll
                        for (int e = 0; e < VN; e++) {
                            s[kk] += a.x[e] * vv.x[e];
                            w[e] += a.x[e] * vk[kk];
                        }
                    }
#pragma unroll
                    for (int e = 0; e < VN; e++) scat<DET>(&u[c + e], w[e], dsc);
                }
                {
                    const T t = wave_sum_cols<KV>(s, lane);
                    const int kk = lane >> (6 - ilog2<KV>());
                    const int r = r0 + kk;
                    if ((lane & (64 / KV - 1)) == 0 && r > j && r < n) y[KV * qg + kk] = t;
                }
            }
        } else
        // each wave takes KC consecutive owned columns r0..r0+KC-1 (a unit
        // holds SY_OB/KC such groups): one LDS atomic per row for their scatters
        for (int t = 0;; t++) {
            static_assert(SY_OB % KC == 0, "a column group must lie inside one unit");
            const int qg = t * SY_WAVES + ((t & 1) ? SY_WAVES - 1 - wave : wave);
            if (KC * qg >= nslots) break;
            const int r0 = owned_row(KC * qg, gi, G);
            if (r0 + KC - 1 <= j || r0 >= n) continue;
            bool ok[KC];
            T vk[KC], s[KC];
            const T* col[KC];
#pragma unroll
            for (int kk = 0; kk < KC; kk++) {
                const int r = r0 + kk;
                ok[kk] = r > j && r < n;
                vk[kk] = ok[kk] ? v[r] : T(0);
                col[kk] = A + (size_t)(ok[kk] ? r : r0) * n;
                s[kk] = 0;
            }
            if (KC == 1) {
                // single column: four row chunks in flight per lane
                T s1 = 0, s2 = 0, s3 = 0;
                const T* c0p = col[0];
                const T vr = vk[0];
                int c = r0 + lane;
                for (; c + 192 < n; c += 256) {
                    const T x0 = c0p[c], x1 = c0p[c + 64], x2 = c0p[c + 128], x3 = c0p[c + 192];
                    s[0] += x0 * v[c];
                    s1 += x1 * v[c + 64];
                    s2 += x2 * v[c + 128];
                    s3 += x3 * v[c + 192];
                    if (c > r0) scat<DET>(&u[c], x0 * vr, dsc);
                    scat<DET>(&u[c + 64], x1 * vr, dsc);
                    scat<DET>(&u[c + 128], x2 * vr, dsc);
                    scat<DET>(&u[c + 192], x3 * vr, dsc);
                }
                for (; c < n; c += 64) {
                    const T x0 = c0p[c];
                    s[0] += x0 * v[c];
                    if (c > r0) scat<DET>(&u[c], x0 * vr, dsc);
                }
                s[0] += (s1 + s2) + s3;
            } else {
                for (int c = r0 + lane; c < n; c += 64) {
                    // v[c] is only defined for c > j: touch it only under a valid column
                    T w = 0;
#pragma unroll
                    for (int kk = 0; kk < KC; kk++) {
                        if (ok[kk] && c >= r0 + kk) {
                            const T a = col[kk][c];
                            s[kk] += a * v[c];
                            if (c > r0 + kk) w += a * vk[kk];
                        }
                    }
                    if (c > j && c > r0) scat<DET>(&u[c], w, dsc);
                }
            }
            {
                const T t = wave_sum_cols<KC>(s, lane);
                const int kk = lane >> (6 - ilog2<KC>());
#pragma unroll
                for (int k2 = 0; k2 < KC; k2++)
                    if (k2 == kk && (lane & (64 / KC - 1)) == 0 && ok[k2]) y[KC * qg + k2] = t;
            }
        }
        __syncthreads();
        if constexpr (DET) {
            // the fixed-point scatter sums back to T
            for (int c = j + 1 + tid; c < n; c += SY_THREADS) u[c] = det_value(&u[c], dsc);
            __syncthreads();
        }
        // row j+1 of the V and W' panels (final since the last column's
        // barriers), loaded ahead of barrier 2
        for (int c = tid; c < i; c += SY_THREADS) {
            Ar[c] = A[j + 1 + (size_t)(k + c) * n];
            Wr[c] = W[j + 1 + (size_t)c * n];
        }
        {
            // v^T y partial of this workgroup: own rows' dot parts and its scatter partial
            T vy = 0;
            for (int q = tid; q < nslots; q += SY_THREADS) {
                const int r = owned_row(q, gi, G);
                if (r > j && r < n) vy += v[r] * y[q];
                if (r == j + 1 && r < n) *p_ydot = y[q];
            }
            for (int c = j + 1 + tid; c < n; c += SY_THREADS) vy += v[c] * u[c];
            vy = block_sum(vy, red);
            if (tid == 0) p_vy[gi] = vy;
        }
        for (int c = j + 1 + tid; c < n; c += SY_THREADS) Pu[(size_t)gi * n + c] = u[c];
        // panel dots over the own rows: c < i with W'(:, c), c >= i with V(:, c - i)
        {
            const int qb = first_slot(j + 1, gi, G);
            for (int c = wave; c < 2 * i; c += SY_WAVES) {
                const int cc = c < i ? c : c - i;
                T s = 0;
                for (int q = qb + lane; q < nslots; q += 64) {
                    const int r = owned_row(q, gi, G);
                    if (r <= j || r >= n) continue;
                    s += (c < i ? wpan(cc, q, r) : vpan(cc, q, r)) * v[r];
                }
                s = wave_sum(s);
                if (lane == 0) p_dot[gi * 2 * nb + c] = s;
            }
        }
        group_barrier(bar, gi, G, epoch, err);

        // coefficients shared by all workgroups, from one round of loads:
        // t12 = [W'^T v; V^T v] summed over the partials, v^T y, the partials of
        // y(j+1), and row j+1 of the panel
        const bool more = i + 1 < kb;   // column j+1 belongs to this panel
        {
            const T pvy = tid < G ? p_vy[tid] : T(0);
            const T pyr = tid < G ? Pu[(size_t)tid * n + j + 1] : T(0);
            const T ydot1 = *p_ydot;
            if (i > 0) {
                const int nc = 2 * i;
                const int PP = DET ? 1 : max(1, min(G, SY_THREADS / nc));   // DET: one summing thread
                for (int idx = tid; idx < nc * PP; idx += SY_THREADS) {
                    const int c = idx % nc, part = idx / nc;
                    T acc = 0;
                    for (int q = part; q < G; q += PP) acc += p_dot[q * 2 * nb + c];
                    lds_add(&t12[c], acc);
                }
            }
            __syncthreads();
            // a2(i) = -tau/2 w^T v with w^T v = tau (v^T y - 2 t1.t2), and
            // W'(j+1, i) = tau y(j+1) + a2(i) v(j+1), v(j+1) = 1
            T x = pvy, yr = pyr;
            if (tid < i) {
                x -= T(2) * t12[tid] * t12[i + tid];
                yr -= Ar[tid] * t12[tid] + Wr[tid] * t12[i + tid];
            }
            block_sum2(x, yr, red);
            const T wv_over_tau = x, ysum = yr;
            if (tid == 0) {
                const T a2i = T(-0.5) * tau * (tau * wv_over_tau);
                a2[i] = a2i;
                Ar[i] = T(1);
                Wr[i] = tau * (ysum + ydot1) + a2i;
            }
        }
        // own rows of y: the G scatter partials and the panel terms, each
        // row's G + i terms split over P threads
        const int qb = first_slot(j + 1, gi, G), nbq = nslots - qb;
        {
            const int nterm = G + i;
            const int P = (nbq > 0 && !DET) ? max(1, min(nterm, SY_THREADS / nbq)) : 1;
            for (int idx = tid; idx < nbq * P; idx += SY_THREADS) {
                const int q = qb + idx % nbq, part = idx / nbq;
                const int r = owned_row(q, gi, G);
                if (r <= j || r >= n) continue;
                T acc = 0;
                for (int t = part; t < nterm; t += P) {
                    if (t < G) {
                        acc += Pu[(size_t)t * n + r];
                    } else {
                        const int c = t - G;
                        acc -= vpan(c, q, r) * t12[c] + wpan(c, q, r) * t12[i + c];
                    }
                }
                lds_add(&y[q], acc);
            }
            __syncthreads();
        }
        // W'(r, i) and V(r, i) = v(r) of the own rows; u[0, nbq) cleared for
        // the update of column j+1 (slot q - qb <= row < n)
        {
            const T a2i = a2[i];
            for (int q = qb + tid; q < nslots; q += SY_THREADS) {
                const int r = owned_row(q, gi, G);
                if (r >= n) continue;
                u[q - qb] = T(0);
                if (r <= j) continue;
                const T vr = v[r], wr = tau * y[q] + a2i * vr;
                W[r + (size_t)i * n] = wr;
                A[r + (size_t)j * n] = vr;
                if constexpr (PLDS) {
                    Vp[(size_t)i * nsm + q] = vr;
                    Wp[(size_t)i * nsm + q] = wr;
                }
            }
            __syncthreads();
        }
        if (!more) break;
        // column j+1 of the own rows: A(r, j+1) -= sum_{c<=i} V(r,c) W'(j+1,c) + W'(r,c) V(j+1,c),
        // each row's i+1 terms split over P threads into u
        {
            const int nterm = i + 1;
            const int P = (nbq > 0 && !DET) ? max(1, min(nterm, SY_THREADS / nbq)) : 1;
            for (int idx = tid; idx < nbq * P; idx += SY_THREADS) {
                const int qq = idx % nbq, part = idx / nbq;
                const int q = qb + qq;
                const int r = owned_row(q, gi, G);
                if (r <= j || r >= n) continue;
                T acc = 0;
                for (int c = part; c < nterm; c += P) acc += vpan(c, q, r) * Wr[c] + wpan(c, q, r) * Ar[c];
                lds_add(&u[qq], acc);
            }
            __syncthreads();
        }
        {
            const int j1 = j + 1;
            T nrm = 0;
            for (int q = qb + tid; q < nslots; q += SY_THREADS) {
                const int r = owned_row(q, gi, G);
                if (r < j1 || r >= n) continue;
                const T a = A[r + (size_t)j1 * n] - u[q - qb];
                A[r + (size_t)j1 * n] = a;
                if (r == j1) g.d[(size_t)p * n + j1] = a;
                if (r == j1 + 1) *p_alpha = a;
                if (r >= j1 + 2) nrm += a * a;
            }
            nrm = block_sum(nrm, red);
            if (tid == 0) p_norm[gi] = nrm;
        }
    }
    // X = [V | W' | V] for the trailing rows owned here
    const int k1 = k + kb;
    for (int q = tid; q < nslots; q += SY_THREADS) {
        const int r = owned_row(q, gi, G);
        if (r < k1 || r >= n) continue;
        for (int c = 0; c < kb; c++) {
            const T vr = vpan(c, q, r), wr = wpan(c, q, r);
            X[r + (size_t)c * n] = vr;
            X[r + (size_t)(kb + c) * n] = wr;
            X[r + (size_t)(2 * kb + c) * n] = vr;
        }
    }
}

// After the last panel: d[n-1], and clear everything on/above the diagonal so
// the strictly lower triangle holds the unit-lower reflectors.
template <class T>
__global__ void sytrd_finish_kernel(const T* A, T* d, int n, int B)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < B) d[(size_t)p * n + n - 1] = A[(size_t)p * n * n + (size_t)(n - 1) * n + (n - 1)];
}

template <class T>
__global__ void clear_upper_kernel(T* A, int n)
{
    const int p = blockIdx.y;
    T* a = A + (size_t)p * n * n;
    const int64_t total = (int64_t)n * n;
    for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
         idx += (int64_t)gridDim.x * blockDim.x) {
        int r = (int)(idx % n), c = (int)(idx / n);
        if (r <= c) a[idx] = T(0);
    }
}

// ====================================================================== ormtr
// T factor (upper triangular, ld nbb) of H(k)...H(k+kb-1) = I - V T V^T from
// the Gram matrix Gm = V^T V (ld nbb) and tau: column i of T is
// -tau_i T(0:i, 0:i) G(0:i, i), diagonal tau_i.  T lives in LDS as its
// upper triangle packed by columns (column c at c (c + 1) / 2); column i of
// G is read into LDS at step i, with tau (dynamic LDS: kb (kb + 1) / 2 + 2 kb
// elements).  Workgroup (p, q) forms the factor of problem p's block
// k = k0 + q nbb (kb = min(nbb, nu - 1 - k)): Gram matrix and factor at
// (q B + p) nbb^2, so one launch can form every block's factor of a
// back-transformation.
template <class T>
constexpr int bt_max() { return sizeof(T) == 4 ? 256 : 128; }   // largest back-transformation block
template <class T>
__global__ __launch_bounds__(256) void larft_kernel(const T* Gm, const T* tau, T* Tf, int n, int nu, int k0, int nbb)
{
    extern __shared__ char smem[];
    const int p = blockIdx.x, q = blockIdx.y;
    const int k = k0 + q * nbb, kb = min(nbb, nu - 1 - k);
    const size_t at0 = ((size_t)q * gridDim.x + p) * nbb * nbb;
    const T* G = Gm + at0;
    T* Tm = Tf + at0;
    const int tri = kb * (kb + 1) / 2;
    T* t = reinterpret_cast<T*>(smem);
    T* gcol = t + tri;
    T* ts = gcol + kb;
    for (int i = threadIdx.x; i < kb; i += 256) ts[i] = tau[(size_t)p * n + k + i];
    for (int i = 0; i < kb; i++) {
        const int ci = i * (i + 1) / 2;
        for (int c = threadIdx.x; c < i; c += 256) gcol[c] = G[c + (size_t)i * nbb];
        __syncthreads();
        if ((int)threadIdx.x < i) {
            const int c = threadIdx.x;
            // four independent terms per pass keep LDS loads in flight
            T acc0 = 0, acc1 = 0, acc2 = 0, acc3 = 0;
            int c2 = c, at = c * (c + 1) / 2 + c;   // at = position of t(c, c2)
            for (; c2 + 3 < i; c2 += 4) {
                acc0 += t[at] * gcol[c2];
                acc1 += t[at + c2 + 1] * gcol[c2 + 1];
                acc2 += t[at + 2 * c2 + 3] * gcol[c2 + 2];
                acc3 += t[at + 3 * c2 + 6] * gcol[c2 + 3];
                at += 4 * c2 + 10;
            }
            for (; c2 < i; c2++) {
                acc0 += t[at] * gcol[c2];
                at += c2 + 1;
            }
            t[ci + c] = -ts[i] * ((acc0 + acc1) + (acc2 + acc3));
        }
        if (threadIdx.x == 0) t[ci + i] = ts[i];
        __syncthreads();
    }
    for (int idx = threadIdx.x; idx < nbb * nbb; idx += 256) {
        const int r = idx % nbb, c = idx / nbb;
        Tm[idx] = (c < kb && r <= c) ? t[c * (c + 1) / 2 + r] : T(0);
    }
}

// ====================================================================== D&C
template <class T>
__global__ void dc_init_kernel(const T* d, const T* e, const unsigned char* split, T* lam, T* Q, int n)
{
    const int p = blockIdx.y;
    const int64_t total = (int64_t)n * n;
    T* q = Q + (size_t)p * n * n;
    for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
         idx += (int64_t)gridDim.x * blockDim.x) {
        int r = (int)(idx % n), c = (int)(idx / n);
        q[idx] = (r == c) ? T(1) : T(0);
        if (r == 0) {
            const T* dd = d + (size_t)p * n;
            const T* ee = e + (size_t)p * n;
            T v = dd[c];
            if (c > 0 && split[c - 1]) v -= fabs(ee[c - 1]);
            if (c < n - 1 && split[c]) v -= fabs(ee[c]);
            lam[(size_t)p * n + c] = v;
        }
    }
}

struct DcArgs {
    int n, nmerge, batch;
    const int* desc;    // s, m1, m2, tile offset
};

template <class T>
struct DcBufs {
    const T* e;
    T *lam, *Q, *QK, *QN, *UB;
    T *dsort, *zsort, *dK, *zK, *taus, *zh, *EV, *rotc, *rots;
    int *perm, *lst, *lstD, *org, *rank, *rotp, *rotj, *kcnt, *rcnt;
    // Column grouping of the merge product: a non-deflated column t of a
    // merge comes from the first child (rows s..s+m1 only), from both (a
    // deflation rotation mixed one of each) or from the second (rows
    // s+m1..s+m only); kpos[t] is its place with the first kind first and the
    // second last, kc1 / kc2 (per merge) the counts of the first two kinds.
    int *kpos, *kc1, *kc2;
};

__device__ inline int find_merge(const int* desc, int nmerge, int x)
{
    int lo = 0, hi = nmerge - 1, ans = -1;
    while (lo <= hi) {
        int mid = (lo + hi) >> 1;
        if (desc[4 * mid] <= x) {
            ans = mid;
            lo = mid + 1;
        } else {
            hi = mid - 1;
        }
    }
    if (ans >= 0 && x < desc[4 * ans] + desc[4 * ans + 1] + desc[4 * ans + 2]) return ans;
    return -1;
}

// z vector, sorted merge of the two children, deflation scan.
template <class T>
__global__ __launch_bounds__(256) void dc_prep_kernel(DcArgs a, DcBufs<T> b)
{
    __shared__ T red[8];
    __shared__ T cd[256], cz[256];
    __shared__ int cp[256];
    const int n = a.n;
    const int p = blockIdx.x / a.nmerge, q = blockIdx.x % a.nmerge;
    const int s = a.desc[4 * q], m1 = a.desc[4 * q + 1], m2 = a.desc[4 * q + 2];
    const int m = m1 + m2;
    const size_t po = (size_t)p * n;
    const T* Q = b.Q + po * n;
    const T* lam = b.lam + po;
    T* dsort = b.dsort + po + s;
    T* zsort = b.zsort + po + s;
    int* perm = b.perm + po + s;
    const T beta = b.e[po + s + m1 - 1];
    const T rho = T(2) * fabs(beta);
    const T sgn = beta >= T(0) ? T(1) : T(-1);
    const T rs2 = T(0.70710678118654752440);
    T dmax = 0, zmax = 0;
    for (int t = threadIdx.x; t < m; t += 256) {
        T D = lam[s + t], z;
        int pos;
        if (t < m1) {
            z = Q[(s + m1 - 1) + (size_t)(s + t) * n] * rs2;
            int lo = 0, hi = m2;   // count right values < D
            while (lo < hi) {
                int mid = (lo + hi) >> 1;
                if (lam[s + m1 + mid] < D) lo = mid + 1; else hi = mid;
            }
            pos = t + lo;
        } else {
            z = sgn * Q[(s + m1) + (size_t)(s + t) * n] * rs2;
            int u = t - m1, lo = 0, hi = m1;   // count left values <= D
            while (lo < hi) {
                int mid = (lo + hi) >> 1;
                if (lam[s + mid] <= D) lo = mid + 1; else hi = mid;
            }
            pos = u + lo;
        }
        dsort[pos] = D;
        zsort[pos] = z;
        perm[pos] = t;
        dmax = fmax(dmax, fabs(D));
        zmax = fmax(zmax, fabs(z));
    }
    dmax = block_max(dmax, red);
    zmax = block_max(zmax, red);
    const T tol = T(8) * Eps<T>::v * fmax(dmax, zmax);
    __syncthreads();

    // sequential deflation scan (thread 0), m values staged 256 at a time;
    // tpj is the kind of the pending column pj (1: first child, 3: second,
    // 2: mixed by a rotation), kept in kpos until the grouping below
    int k = 0, ndf = 0, nrot = 0, pj = -1, tpj = 0;
    T dpj = 0, zpj = 0;
    int* kpos = b.kpos + po + s;
    int* lst = b.lst + po + s;
    int* lstD = b.lstD + po + s;
    T* dK = b.dK + po + s;
    T* zK = b.zK + po + s;
    for (int c0 = 0; c0 < m; c0 += 256) {
        __syncthreads();
        if (c0 + (int)threadIdx.x < m) {
            cd[threadIdx.x] = dsort[c0 + threadIdx.x];
            cz[threadIdx.x] = zsort[c0 + threadIdx.x];
            cp[threadIdx.x] = perm[c0 + threadIdx.x] < m1 ? 1 : 3;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            const int ce = min(256, m - c0);
            for (int jj = 0; jj < ce; jj++) {
                const int j = c0 + jj;
                T dj = cd[jj], zj = cz[jj];
                const int tj = cp[jj];
                if (rho * fabs(zj) <= tol) {
                    lstD[ndf++] = j;
                    continue;
                }
                if (pj < 0) {
                    pj = j;
                    dpj = dj;
                    zpj = zj;
                    tpj = tj;
                    continue;
                }
                T sv = zpj, cv = zj;
                T tv = hypot(cv, sv);
                T t = dj - dpj;
                cv /= tv;
                sv = -sv / tv;
                if (fabs(t * cv * sv) <= tol) {
                    b.rotp[po + s + nrot] = pj;
                    b.rotj[po + s + nrot] = j;
                    b.rotc[po + s + nrot] = cv;
                    b.rots[po + s + nrot] = sv;
                    nrot++;
                    T t2 = dpj * cv * cv + dj * sv * sv;
                    T djn = dpj * sv * sv + dj * cv * cv;
                    dsort[pj] = t2;
                    lstD[ndf++] = pj;
                    pj = j;
                    dpj = djn;
                    zpj = tv;
                    tpj = tpj == tj ? tj : 2;   // the rotation mixes the two columns
                    dsort[j] = djn;
                } else {
                    lst[k] = pj;
                    dK[k] = dpj;
                    zK[k] = zpj;
                    kpos[k] = tpj;
                    k++;
                    pj = j;
                    dpj = dj;
                    zpj = zj;
                    tpj = tj;
                }
            }
        }
    }
    if (threadIdx.x == 0) {
        if (pj >= 0) {
            lst[k] = pj;
            dK[k] = dpj;
            zK[k] = zpj;
            kpos[k] = tpj;
            k++;
        }
        b.kcnt[po + q] = k;
        b.rcnt[po + q] = nrot;
        // the grouping: first-child columns, then mixed, then second-child
        int n1 = 0, n2 = 0;
        for (int t = 0; t < k; t++) {
            n1 += kpos[t] == 1;
            n2 += kpos[t] == 2;
        }
        int i1 = 0, i2 = n1, i3 = n1 + n2;
        for (int t = 0; t < k; t++) {
            const int kind = kpos[t];
            kpos[t] = kind == 1 ? i1++ : (kind == 2 ? i2++ : i3++);
        }
        b.kc1[po + q] = n1;
        b.kc2[po + q] = n2;
    }
}

// Apply the deflation rotations to the merge's columns of Q (one row per thread).
template <class T>
__global__ __launch_bounds__(256) void dc_rot_kernel(DcArgs a, DcBufs<T> b)
{
    const int n = a.n;
    const int p = blockIdx.y, q = blockIdx.x;
    const size_t po = (size_t)p * n;
    const int nrot = b.rcnt[po + q];
    if (nrot == 0) return;
    const int s = a.desc[4 * q], m = a.desc[4 * q + 1] + a.desc[4 * q + 2];
    T* Q = b.Q + po * n;
    const int* perm = b.perm + po + s;
    for (int r = threadIdx.x; r < m; r += 256) {
        for (int t = 0; t < nrot; t++) {
            const int cx = s + perm[b.rotp[po + s + t]], cy = s + perm[b.rotj[po + s + t]];
            const T c = b.rotc[po + s + t], sn = b.rots[po + s + t];
            T* x = Q + (s + r) + (size_t)cx * n;
            T* y = Q + (s + r) + (size_t)cy * n;
            T xv = *x, yv = *y;
            *x = c * xv + sn * yv;
            *y = c * yv - sn * xv;
        }
    }
}

// Secular equation 1/rho + sum z_j^2/(d_j - lam) = 0, one wave per root.
template <class T>
__global__ __launch_bounds__(256) void dc_secular_kernel(DcArgs a, DcBufs<T> b)
{
    const int n = a.n;
    const int lane = threadIdx.x & 63;
    const int64_t wid = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int p = (int)(wid / n), x = (int)(wid % n);
    if (p >= a.batch) return;   // the grid is rounded up to whole workgroups
    const int q = find_merge(a.desc, a.nmerge, x);
    if (q < 0) return;
    const size_t po = (size_t)p * n;
    const int s = a.desc[4 * q];
    const int k = b.kcnt[po + q];
    const int i = x - s;
    if (i >= k) return;
    const T* dK = b.dK + po + s;
    const T* zK = b.zK + po + s;
    const T beta = b.e[po + s + a.desc[4 * q + 1] - 1];
    const T rho = T(2) * fabs(beta);
    const T rinv = T(1) / rho;
    int o;
    T lo, hi;
    if (i < k - 1) {
        const T di = dK[i];
        const T mid = T(0.5) * (dK[i + 1] - di);
        T f = 0;
        for (int j = lane; j < k; j += 64) {
            T zj = zK[j];
            f += zj * zj / ((dK[j] - di) - mid);
        }
        f = wave_sum(f) + rinv;
        if (f >= T(0)) {
            o = i;
            lo = 0;
            hi = mid;
        } else {
            o = i + 1;
            lo = -mid;
            hi = 0;
        }
    } else {
        o = i;
        lo = 0;
        hi = rho;
    }
    const T dorg = dK[o];
    const T Dl = dK[i] - dorg;
    const T Dr = (i + 1 < k) ? dK[i + 1] - dorg : T(0);
    const T zl2 = zK[i] * zK[i];
    const T zr2 = (i + 1 < k) ? zK[i + 1] * zK[i + 1] : T(0);
    const T eps = Eps<T>::v;
    T t = T(0.5) * (lo + hi);
    T wprev = hi - lo;
    for (int it = 0; it < 200; it++) {
        T gsum = 0, gabs = 0;
        for (int j = lane; j < k; j += 64) {
            T zj = zK[j];
            T term = zj * zj / ((dK[j] - dorg) - t);
            gsum += term;
            gabs += fabs(term);
        }
        gsum = wave_sum(gsum);
        gabs = wave_sum(gabs);
        const T gv = gsum + rinv;
        if (gv == T(0)) break;
        if (gv < T(0)) lo = t; else hi = t;
        if (fabs(gv) <= T(8) * eps * (rinv + gabs)) break;
        if (hi - lo <= T(4) * eps * fmax(fabs(lo), fabs(hi))) break;
        const bool force = (it & 1) && (hi - lo) > T(0.5) * wprev;
        if (it & 1) wprev = hi - lo;
        // fixed-weight rational step with the two bracketing poles
        const T cres = gv - zl2 / (Dl - t) - ((i + 1 < k) ? zr2 / (Dr - t) : T(0));
        T tn = T(0.5) * (lo + hi);
        bool ok = false;
        if (i + 1 < k) {
            const T a2 = cres;
            const T a1 = -(cres * (Dl + Dr) + zl2 + zr2);
            const T a0 = cres * Dl * Dr + zl2 * Dr + zr2 * Dl;
            if (a2 == T(0)) {
                if (a1 != T(0)) {
                    T c1 = -a0 / a1;
                    if (c1 > lo && c1 < hi) { tn = c1; ok = true; }
                }
            } else {
                T disc = a1 * a1 - T(4) * a2 * a0;
                if (disc >= T(0)) {
                    T sq = sqrt(disc);
                    T qq = T(-0.5) * (a1 + copysign(sq, a1));
                    T c1 = qq / a2;
                    if (c1 > lo && c1 < hi) { tn = c1; ok = true; }
                    else if (qq != T(0)) {
                        T c2 = a0 / qq;
                        if (c2 > lo && c2 < hi) { tn = c2; ok = true; }
                    }
                }
            }
        } else if (cres > T(0)) {
            T c1 = Dl + zl2 / cres;
            if (c1 > lo && c1 < hi) { tn = c1; ok = true; }
        }
        if (!ok || force) tn = T(0.5) * (lo + hi);
        t = tn;
    }
    if (lane == 0) {
        b.org[po + s + i] = o;
        b.taus[po + s + i] = t;
    }
}

// Gu-Eisenstat z-hat, one wave per component.
template <class T>
__global__ __launch_bounds__(256) void dc_zhat_kernel(DcArgs a, DcBufs<T> b)
{
    const int n = a.n;
    const int lane = threadIdx.x & 63;
    const int64_t wid = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int p = (int)(wid / n), x = (int)(wid % n);
    if (p >= a.batch) return;   // the grid is rounded up to whole workgroups
    const int q = find_merge(a.desc, a.nmerge, x);
    if (q < 0) return;
    const size_t po = (size_t)p * n;
    const int s = a.desc[4 * q];
    const int k = b.kcnt[po + q];
    const int j = x - s;
    if (j >= k) return;
    const T* dK = b.dK + po + s;
    const T* taus = b.taus + po + s;
    const int* org = b.org + po + s;
    const T rho = T(2) * fabs(b.e[po + s + a.desc[4 * q + 1] - 1]);
    const T dj = dK[j];
    T prod = T(1);
    for (int i = lane; i < k; i += 64) {
        const T lmj = (dK[org[i]] - dj) + taus[i];
        T f;
        if (i < j) f = lmj / (dK[i] - dj);
        else if (i < k - 1) f = lmj / (dK[i + 1] - dj);
        else f = lmj / rho;
        prod *= f;
    }
    prod = wave_prod(prod);
    if (lane == 0) b.zh[po + s + j] = copysign(sqrt(fmax(prod, T(0))), b.zK[po + s + j]);
}

// Rank-one eigenvectors U (k x k, ld n at column offset s) and gather of the
// non-deflated columns of Q into QK; one wave per column.
template <class T>
__global__ __launch_bounds__(256) void dc_vec_kernel(DcArgs a, DcBufs<T> b)
{
    const int n = a.n;
    const int lane = threadIdx.x & 63;
    const int64_t wid = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int p = (int)(wid / n), x = (int)(wid % n);
    if (p >= a.batch) return;   // the grid is rounded up to whole workgroups
    const int q = find_merge(a.desc, a.nmerge, x);
    if (q < 0) return;
    const size_t po = (size_t)p * n;
    const int s = a.desc[4 * q], m = a.desc[4 * q + 1] + a.desc[4 * q + 2];
    const int k = b.kcnt[po + q];
    const int i = x - s;
    if (i >= k) return;
    const T* dK = b.dK + po + s;
    const T* zh = b.zh + po + s;
    const T dorg = dK[b.org[po + s + i]];
    const T ti = b.taus[po + s + i];
    T ss = 0;
    for (int t = lane; t < k; t += 64) {
        T u = zh[t] / ((dK[t] - dorg) - ti);
        ss += u * u;
    }
    ss = wave_sum(ss);
    const T inv = T(1) / sqrt(ss);
    // U's rows and QK's columns in the grouped order (kpos): their product is
    // the same, and the merge GEMM skips the zero blocks
    const int* kpos = b.kpos + po + s;
    T* U = b.UB + po * n + (size_t)(s + i) * n;
    for (int t = lane; t < k; t += 64) U[kpos[t]] = zh[t] / ((dK[t] - dorg) - ti) * inv;
    // gather: QK[:, s+kpos[i]] = Q[s.., s + perm[lst[i]]]
    const int col = s + b.perm[po + s + b.lst[po + s + i]];
    const T* src = b.Q + po * n + s + (size_t)col * n;
    T* dst = b.QK + po * n + (size_t)(s + kpos[i]) * n;
    for (int r = lane; r < m; r += 64) dst[r] = src[r];
}

// Grouped merge GEMM: QN[:, s..s+k) = QK[:, s..s+k) * U, per merge, as two
// products that skip QK's zero blocks (columns grouped by kpos): the first
// child's m1 rows over the first-child and mixed columns, the second child's
// m2 rows over the mixed and second-child columns.
template <class T>
__global__ __launch_bounds__(256) void dc_gemm_kernel(DcArgs a, DcBufs<T> b, int64_t tiles)
{
    const int n = a.n;
    const int p = (int)(blockIdx.x / tiles);
    const int64_t t = blockIdx.x % tiles;
    int lo = 0, hi = a.nmerge - 1, q = 0;
    while (lo <= hi) {
        int mid = (lo + hi) >> 1;
        if (a.desc[4 * mid + 3] <= t) { q = mid; lo = mid + 1; } else hi = mid - 1;
    }
    const size_t po = (size_t)p * n;
    const int s = a.desc[4 * q], m1 = a.desc[4 * q + 1], m2 = a.desc[4 * q + 2], m = m1 + m2;
    const int k = b.kcnt[po + q], k1 = b.kc1[po + q], k2 = b.kc2[po + q];
    const int ntm = (m + 63) / 64, nt1 = (m1 + 63) / 64, nt2 = (m2 + 63) / 64;
    int64_t lt = t - a.desc[4 * q + 3];
    const T* A = b.QK + po * n + (size_t)s * n;
    const T* B = b.UB + po * n + (size_t)s * n;
    T* C = b.QN + po * n + (size_t)s * n;
    int rows, inner, ti, tj;
    if (lt < (int64_t)nt1 * ntm) {
        // first child's rows over columns [0, k1 + k2)
        ti = (int)(lt % nt1);
        tj = (int)(lt / nt1);
        rows = m1;
        inner = k1 + k2;
    } else {
        // second child's rows over columns [k1, k)
        lt -= (int64_t)nt1 * ntm;
        if (nt2 == 0) return;
        ti = (int)(lt % nt2);
        tj = (int)(lt / nt2);
        rows = m2;
        inner = k - k1;
        A += m1 + (size_t)k1 * n;
        B += k1;
        C += m1;
    }
    if (tj * 64 >= k || ti * 64 >= rows) return;
    GemmShape g{rows, k, inner, false, false, (int64_t)n, (int64_t)n, (int64_t)n, 0, 0, 0, 1, 0};
    gemm_tile<T, GemmBK<T>::value>(g, ti * 64, tj * 64, T(1), A, B, T(0), C);
}

// Eigenvalues of the merged block and the deflated columns appended to QN.
template <class T>
__global__ __launch_bounds__(256) void dc_final1_kernel(DcArgs a, DcBufs<T> b)
{
    const int n = a.n;
    const int p = blockIdx.y, x = blockIdx.x;
    const int q = find_merge(a.desc, a.nmerge, x);
    if (q < 0) return;
    const size_t po = (size_t)p * n;
    const int s = a.desc[4 * q], m = a.desc[4 * q + 1] + a.desc[4 * q + 2];
    const int k = b.kcnt[po + q];
    const int t = x - s;
    if (t < k) {
        if (threadIdx.x == 0) b.EV[po + x] = b.dK[po + s + b.org[po + s + t]] + b.taus[po + s + t];
        return;
    }
    const int pos = b.lstD[po + s + (t - k)];
    if (threadIdx.x == 0) b.EV[po + x] = b.dsort[po + s + pos];
    const int col = s + b.perm[po + s + pos];
    const T* src = b.Q + po * n + s + (size_t)col * n;
    T* dst = b.QN + po * n + (size_t)x * n;
    for (int r = threadIdx.x; r < m; r += 256) dst[r] = src[r];
}

template <class T>
__global__ __launch_bounds__(256) void dc_rank_kernel(DcArgs a, DcBufs<T> b)
{
    const int n = a.n;
    const int p = blockIdx.y;
    const int x = blockIdx.x * 256 + threadIdx.x;
    if (x >= n) return;
    const int q = find_merge(a.desc, a.nmerge, x);
    if (q < 0) return;
    const size_t po = (size_t)p * n;
    const int s = a.desc[4 * q], m = a.desc[4 * q + 1] + a.desc[4 * q + 2];
    const T* EV = b.EV + po + s;
    const int t = x - s;
    const T v = EV[t];
    int r = 0;
    for (int u = 0; u < m; u++) {
        T w = EV[u];
        r += (w < v || (w == v && u < t)) ? 1 : 0;
    }
    b.rank[po + x] = s + r;
}

template <class T>
__global__ __launch_bounds__(256) void dc_final2_kernel(DcArgs a, DcBufs<T> b)
{
    const int n = a.n;
    const int p = blockIdx.y, x = blockIdx.x;
    const int q = find_merge(a.desc, a.nmerge, x);
    if (q < 0) return;
    const size_t po = (size_t)p * n;
    const int s = a.desc[4 * q], m = a.desc[4 * q + 1] + a.desc[4 * q + 2];
    const int dst_col = b.rank[po + x];
    if (threadIdx.x == 0) b.lam[po + dst_col] = b.EV[po + x];
    const T* src = b.QN + po * n + (size_t)x * n;
    T* dst = b.Q + po * n + s + (size_t)dst_col * n;
    for (int r = threadIdx.x; r < m; r += 256) dst[r] = src[r];
}

// Padded problems (EigPlan::np > n): Ap = diag(A, 0) of order np.  The
// Householder vectors of the reduction then vanish on the pad rows, so the
// pad rows and columns stay zero apart from their diagonal and the
// subdiagonal e is exactly zero from row n - 1 on.
template <class T>
__global__ void pad_in_kernel(const T* A, int n, T* Ap, int np)
{
    const int b = blockIdx.y;
    const T* a = A + (size_t)b * n * n;
    T* ap = Ap + (size_t)b * np * np;
    const int64_t total = (int64_t)np * np;
    for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
         idx += (int64_t)gridDim.x * blockDim.x) {
        const int r = (int)(idx % np), c = (int)(idx / np);
        ap[idx] = (r < n && c < n) ? a[r + (size_t)c * n] : T(0);
    }
}

// The pad diagonal of the tridiagonal matrix: c = max_i (d_i + |e_i-1| + |e_i|)
// over the n rows of the problem (Gershgorin: no eigenvalue of A lies above
// it) plus 2^-10 of the largest Gershgorin radius, so the np - n pad
// eigenvalues come last in ascending order and never deflate against an
// eigenvalue of A.  One workgroup per problem.
template <class T>
__global__ __launch_bounds__(256) void pad_tail_kernel(T* d, T* e, int n, int np)
{
    __shared__ T red[8];
    T* dd = d + (size_t)blockIdx.x * np;
    T* ee = e + (size_t)blockIdx.x * np;
    T hi = -INFINITY, mag = 0;
    for (int i = threadIdx.x; i < n; i += 256) {
        const T off = (i > 0 ? fabs(ee[i - 1]) : T(0)) + (i < n - 1 ? fabs(ee[i]) : T(0));
        hi = fmax(hi, dd[i] + off);
        mag = fmax(mag, fabs(dd[i]) + off);
    }
    hi = block_max(hi, red);
    mag = block_max(mag, red);
    const T c = hi + T(0.0009765625) * mag + std::numeric_limits<T>::min();
    for (int i = n + threadIdx.x; i < np; i += 256) dd[i] = c;
    for (int i = n - 1 + threadIdx.x; i < np - 1; i += 256) ee[i] = T(0);
}

// The reflectors of a padded reduction (strictly lower part of Ap, np x np)
// into A (n x n): their pad rows are zero, so the back-transformation reads
// them at order n.
template <class T>
__global__ void pad_vecs_kernel(const T* Ap, int np, T* A, int n)
{
    const int b = blockIdx.y;
    const T* ap = Ap + (size_t)b * np * np;
    T* a = A + (size_t)b * n * n;
    const int64_t total = (int64_t)n * n;
    for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
         idx += (int64_t)gridDim.x * blockDim.x) {
        const int r = (int)(idx % n), c = (int)(idx / n);
        a[idx] = ap[r + (size_t)c * np];
    }
}

// Eigenvectors and eigenvalues of the padded problems back to order n: the
// first n eigenpairs are those of A, and their pad rows are zero.
template <class T>
__global__ void pad_out_kernel(const T* Z, const T* wp, int np, T* A, T* w, int n)
{
    const int b = blockIdx.y;
    const T* z = Z + (size_t)b * np * np;
    T* a = A + (size_t)b * n * n;
    const int64_t total = (int64_t)n * n;
    for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
         idx += (int64_t)gridDim.x * blockDim.x) {
        const int r = (int)(idx % n), c = (int)(idx / n);
        a[idx] = z[r + (size_t)c * np];
        if (r == 0) w[(size_t)b * n + c] = wp[(size_t)b * np + c];
    }
}

template <class T>
__global__ void copy_kernel(const T* src, T* dst, int64_t count)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += (int64_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

// ====================================================================== plan
template <class T>
EigPlan<T>::~EigPlan()
{
    if (!owns_tables) return;
    for (auto& l : levels)
        if (l.d_desc) (void)hipFree(l.d_desc);
    if (d_split) (void)hipFree(d_split);
}

// Carve the workspace; with base == nullptr only ws_bytes is computed.
template <class T>
static int eig_layout(EigPlan<T>* p, char* base)
{
    const size_t B = p->batch, n = p->np, nn = n * n, nb = p->nb, nbb = p->nbb, G = p->G;
    size_t off = 0;
    auto take = [&](auto*& ptr, size_t bytes) {
        size_t o = align_up(off);
        off = o + bytes;
        ptr = base ? reinterpret_cast<std::remove_reference_t<decltype(ptr)>>(base + o) : nullptr;
    };
    take(p->W, B * n * nb * sizeof(T));
    take(p->X, B * n * 3 * nb * sizeof(T));
    take(p->d, B * n * sizeof(T));
    take(p->e, B * n * sizeof(T));
    take(p->tau, B * n * sizeof(T));
    take(p->part, B * (2 + G * (2 + 2 * nb)) * sizeof(T));
    take(p->Pu, B * G * n * sizeof(T));
    take(p->vg, B * G * n * sizeof(T));
    take(p->Q, B * nn * sizeof(T));
    take(p->QK, B * nn * sizeof(T));
    take(p->QN, B * nn * sizeof(T));
    take(p->UB, B * nn * sizeof(T));
    for (T** v : {&p->dsort, &p->zsort, &p->dK, &p->zK, &p->taus, &p->zh, &p->EV, &p->rotc, &p->rots})
        take(*v, B * n * sizeof(T));
    for (int** v : {&p->perm, &p->lst, &p->lstD, &p->org, &p->rank, &p->rotp, &p->rotj, &p->kcnt, &p->rcnt,
                    &p->kpos, &p->kc1, &p->kc2})
        take(*v, B * n * sizeof(int));
    take(p->Tf, B * 2 * nbb * nbb * sizeof(T));
    take(p->bar, (B * G + 1) * BAR_STRIDE * sizeof(unsigned));
    take(p->err, 64);
    take(p->fro, B * sizeof(T));
    // padded problems: the reduction runs in Q (free until the divide &
    // conquer), the eigenvalues of order np need their own vector
    p->Ap = p->np != p->n ? p->Q : nullptr;
    if (p->np != p->n) take(p->wp, B * n * sizeof(T));
    else p->wp = nullptr;
    p->ws_bytes = align_up(off);
    return 0;
}

// The plan's host part: G, the divide & conquer levels with their merge
// descriptors (h_desc, one vector per level) and split flags (h_split).
template <class T>
static int eig_plan_host(int n, int batch, EigPlan<T>* p, std::vector<std::vector<int>>& h_desc,
                         std::vector<unsigned char>& h_split)
{
    GMC_CHECK(n >= 1 && batch >= 1, "eig_plan: bad sizes n=%d batch=%d", n, batch);
    p->n = n;
    p->batch = batch;
    {
        constexpr int vw = 16 / (int)sizeof(T);
        p->np = (n >= EIG_PAD_MIN && n % vw != 0) ? (n + vw - 1) / vw * vw : n;
    }
    n = p->np;   // the tree, G and the workspace are those of the padded order
    // workgroups per problem: all problems resident at once (one per CU)
    int ncu = 256;
    {
        hipDeviceProp_t prop;
        int dev = 0;
        if (hipGetDevice(&dev) == hipSuccess && hipGetDeviceProperties(&prop, dev) == hipSuccess)
            ncu = prop.multiProcessorCount;
    }
    const int nblk = (n + 63) / 64;
    p->G = std::max(1, std::min(nblk, ncu / batch));
    GMC_CHECK(batch <= ncu, "eig_plan: batch %d exceeds the %d co-resident problems", batch, ncu);

    // divide & conquer tree: recursive halving down to single rows
    struct Node { int s, m1, m2, h; };
    std::vector<Node> nodes;
    std::vector<unsigned char> split(n > 1 ? n - 1 : 1, 0);
    std::function<int(int, int)> build = [&](int s, int m) -> int {
        if (m == 1) return 0;
        int m1 = m / 2;
        split[s + m1 - 1] = 1;
        int h = 1 + std::max(build(s, m1), build(s + m1, m - m1));
        nodes.push_back({s, m1, m - m1, h});
        return h;
    };
    int H = build(0, n);
    p->levels.assign(H, DcLevel{});
    for (int h = 1; h <= H; h++) {
        std::vector<Node> lv;
        for (auto& nd : nodes)
            if (nd.h == h) lv.push_back(nd);
        std::sort(lv.begin(), lv.end(), [](const Node& x, const Node& y) { return x.s < y.s; });
        std::vector<int> desc;
        int64_t off = 0;
        int maxm = 0;
        for (auto& nd : lv) {
            int m = nd.m1 + nd.m2;
            desc.insert(desc.end(), {nd.s, nd.m1, nd.m2, (int)off});
            // the merge product's tiles: rows of each child's part times the columns
            int64_t tm = (m + 63) / 64;
            off += ((nd.m1 + 63) / 64 + (nd.m2 + 63) / 64) * tm;
            maxm = std::max(maxm, m);
        }
        GMC_CHECK(off < (1LL << 31), "eig_plan: tile count overflow");
        DcLevel& L = p->levels[h - 1];
        L.nmerge = (int)lv.size();
        L.tiles = off;
        L.maxm = maxm;
        h_desc.push_back(std::move(desc));
    }
    h_split = std::move(split);

    // back-transformation block: the rank-nbb updates of Z read and write Z
    // once per block, so the block is as wide as the T factor in LDS allows
    p->nbb = std::min(bt_max<T>(), n);
    GMC_TRY(eig_layout(p, nullptr));
    return 0;
}

template <class T>
size_t eig_workspace_bytes(int n, int batch)
{
    EigPlan<T> p;
    p.owns_tables = false;
    std::vector<std::vector<int>> h_desc;
    std::vector<unsigned char> h_split;
    if (eig_plan_host(n, batch, &p, h_desc, h_split) != 0) return SIZE_MAX;
    return p.ws_bytes;
}
template size_t eig_workspace_bytes<double>(int, int);
template size_t eig_workspace_bytes<float>(int, int);

template <class T>
int eig_plan(int n, int batch, EigPlan<T>* p)
{
    std::vector<std::vector<int>> h_desc;
    std::vector<unsigned char> h_split;
    GMC_TRY(eig_plan_host(n, batch, p, h_desc, h_split));
    p->owns_tables = true;
    for (size_t h = 0; h < h_desc.size(); h++) {
        DcLevel& L = p->levels[h];
        GMC_HIP(hipMalloc(&L.d_desc, h_desc[h].size() * sizeof(int)));
        GMC_HIP(hipMemcpy(L.d_desc, h_desc[h].data(), h_desc[h].size() * sizeof(int), hipMemcpyHostToDevice));
    }
    GMC_HIP(hipMalloc(&p->d_split, h_split.size()));
    GMC_HIP(hipMemcpy(p->d_split, h_split.data(), h_split.size(), hipMemcpyHostToDevice));
    return 0;
}

template <class T>
int eig_plan_at(int n, int batch, EigPlan<T>* p, void* mem, size_t mem_bytes, size_t* used, hipStream_t s)
{
    std::vector<std::vector<int>> h_desc;
    std::vector<unsigned char> h_split;
    GMC_TRY(eig_plan_host(n, batch, p, h_desc, h_split));
    p->owns_tables = false;
    // the tables, then the workspace (bound by the caller at mem + *used)
    size_t need = 0;
    for (auto& d : h_desc) need = align_up(need) + d.size() * sizeof(int);
    need = align_up(align_up(need) + h_split.size()) + p->ws_bytes;
    GMC_CHECK(need <= mem_bytes, "syevd: n=%d, batch=%d needs %zu bytes of device scratch, more than the %zu "
              "available", n, batch, need, mem_bytes);
    char* base = static_cast<char*>(mem);
    size_t off = 0;
    for (size_t h = 0; h < h_desc.size(); h++) {
        const size_t bytes = h_desc[h].size() * sizeof(int);
        off = align_up(off);
        p->levels[h].d_desc = reinterpret_cast<int*>(base + off);
        GMC_HIP(hipMemcpyAsync(base + off, h_desc[h].data(), bytes, hipMemcpyHostToDevice, s));
        off += bytes;
    }
    off = align_up(off);
    p->d_split = reinterpret_cast<unsigned char*>(base + off);
    GMC_HIP(hipMemcpyAsync(base + off, h_split.data(), h_split.size(), hipMemcpyHostToDevice, s));
    off += h_split.size();
    // the host vectors go out of scope: the copies must have left them
    GMC_HIP(hipStreamSynchronize(s));
    *used = align_up(off);
    return 0;
}


template <class T>
int eig_bind(EigPlan<T>* p, void* ws)
{
    GMC_TRY(eig_layout(p, static_cast<char*>(ws)));
    p->bound = ws != nullptr;
    p->clear_err = true;
    return 0;
}

static inline unsigned grid1(int64_t work, int per = 256, int64_t cap = 1 << 20)
{
    int64_t g = (work + per - 1) / per;
    return (unsigned)std::max<int64_t>(1, std::min<int64_t>(g, cap));
}

static bool capturing(hipStream_t s)
{
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    return hipStreamIsCapturing(s, &cs) == hipSuccess && cs == hipStreamCaptureStatusActive;
}

// The group barriers need every workgroup of the grid resident at once; a
// cooperative launch guarantees that, on a stream being captured into a graph
// too (a kernel node with the cooperative attribute).  When the runtime
// refuses the eager cooperative launch, a plain launch is taken only if the
// occupancy calculator puts the whole grid on the device at once; otherwise
// the call fails instead of waiting for a barrier that cannot complete.
static int launch_persistent(const void* kfn, dim3 grid, dim3 block, void** args, size_t lds, hipStream_t s)
{
    if (capturing(s)) {
        hipLaunchAttribute attr{};
        attr.id = hipLaunchAttributeCooperative;
        attr.val.cooperative = 1;
        hipLaunchConfig_t cfg{};
        cfg.gridDim = grid;
        cfg.blockDim = block;
        cfg.dynamicSmemBytes = lds;
        cfg.stream = s;
        cfg.attrs = &attr;
        cfg.numAttrs = 1;
        const hipError_t e = hipLaunchKernelExC(&cfg, kfn, args);
        GMC_CHECK(e == hipSuccess, "syevd: cooperative launch not captured into the graph (%s)", hipGetErrorString(e));
        return 0;
    }
    const hipError_t e = hipLaunchCooperativeKernel(kfn, grid, block, args, (unsigned)lds, s);
    if (e == hipSuccess) return 0;
    (void)hipGetLastError();
    int dev = 0, ncu = 0, per_cu = 0;
    GMC_HIP(hipGetDevice(&dev));
    GMC_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
    GMC_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kfn, (int)(block.x * block.y * block.z), lds));
    const long resident = (long)per_cu * ncu, need = (long)grid.x * grid.y * grid.z;
    GMC_CHECK(need <= resident,
              "syevd: cooperative launch refused (%s) and %ld workgroups do not fit the %ld resident slots",
              hipGetErrorString(e), need, resident);
    GMC_HIP(hipLaunchKernel(kfn, grid, block, args, lds, s));
    return 0;
}

template <class T>
int syevd_batched(hipStream_t s, EigPlan<T>& p, T* A, T* w)
{
    GMC_CHECK(p.bound, "syevd: plan not bound to a workspace");
    const int n = p.np, B = p.batch, nb = p.nb, G = p.G;
    const size_t nn = (size_t)n * n;
    const bool padded = p.np != p.n;
    T* const Au = A;
    T* const wu = w;
    if (padded) {
        hipLaunchKernelGGL(pad_in_kernel<T>, dim3(grid1(nn), B), dim3(256), 0, s, Au, p.n, p.Ap, n);
        GMC_HIP(hipGetLastError());
        A = p.Ap;
        w = p.wp;
    }
    hipLaunchKernelGGL(zero_u32, dim3(((B * G + 1) * BAR_STRIDE + 255) / 256), dim3(256), 0, s, p.bar,
                       (B * G + 1) * BAR_STRIDE);
    if (p.clear_err) {
        hipLaunchKernelGGL(zero_u32, dim3(1), dim3(64), 0, s, (unsigned*)p.err, 16);
        p.clear_err = false;
    }
    GMC_HIP(hipGetLastError());

    // ---------------- tridiagonalisation
    const bool f64 = sizeof(T) == 8;
    ProfScope ps_tri(s, f64 ? "eig_tridiag_f64" : "eig_tridiag_f32", 0.0);
    if (n > 1) {
        const int nunit = (n + SY_OB - 1) / SY_OB;
        const int nslots_max = SY_OB * ((nunit + G - 1) / G);
        // LDS: u and v (n each; v moves to a global copy when it does not
        // fit), y and the small arrays, and the own rows of the V and W'
        // panels when they fit beside them
        constexpr size_t LDS_MAX = 160 * 1024;
        const size_t small = (size_t)(nslots_max + 5 * nb + 2 * SY_WAVES + 8) * sizeof(T);
        const size_t panels = 2 * (size_t)nb * nslots_max * sizeof(T);
        const size_t vec_n = (size_t)n * sizeof(T);
        const bool vlds = 2 * vec_n + small <= LDS_MAX;
        const bool plds = (vlds ? 2 : 1) * vec_n + small + panels <= LDS_MAX;
        const size_t lds = (vlds ? 2 : 1) * vec_n + small + (plds ? panels : 0);
        GMC_CHECK(lds <= LDS_MAX, "syevd: n=%d needs %zu bytes of LDS per workgroup", n, lds);
        // 16-byte SYMV shape per type (measured at n=2664/4524, batch 16):
        // FP32 8 columns x 1 chunk, FP64 4 columns x 2 chunks in flight
        constexpr int KV = sizeof(T) == 8 ? 4 : 8, UV = sizeof(T) == 8 ? 2 : 1;
        const bool det = deterministic();
        const void* kfn =
            det ? (vlds ? (plds ? (const void*)sytrd_panel_kernel<T, true, true, KV, UV, true>
                                : (const void*)sytrd_panel_kernel<T, true, false, KV, UV, true>)
                        : (plds ? (const void*)sytrd_panel_kernel<T, false, true, KV, UV, true>
                                : (const void*)sytrd_panel_kernel<T, false, false, KV, UV, true>))
                : (vlds ? (plds ? (const void*)sytrd_panel_kernel<T, true, true, KV, UV, false>
                                : (const void*)sytrd_panel_kernel<T, true, false, KV, UV, false>)
                        : (plds ? (const void*)sytrd_panel_kernel<T, false, true, KV, UV, false>
                                : (const void*)sytrd_panel_kernel<T, false, false, KV, UV, false>));
        if (det) {
            // the Frobenius norm of each problem, summed in a fixed order
            hipLaunchKernelGGL(frob_kernel<T>, dim3(B), dim3(256), 0, s, A, n, p.fro);
            GMC_HIP(hipGetLastError());
        }
        // (set by an eager call before a capture; not repeated inside one)
        if (!capturing(s)) GMC_HIP(hipFuncSetAttribute(kfn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
        unsigned epoch = 0;
        for (int k = 0; k < n - 1; k += nb) {
            const int kb = std::min(nb, n - 1 - k);
            const int vec = (n % (16 / (int)sizeof(T)) == 0) && (reinterpret_cast<uintptr_t>(A) % 16 == 0) &&
                            (reinterpret_cast<uintptr_t>(p.vg) % 16 == 0);
            PanelArgs<T> a{A, p.W, p.X, p.d, p.e, p.tau, p.part, p.Pu, p.vg, p.bar, p.err, p.fro, n, k, kb, nb, G,
                           epoch, vec};
            void* args[] = {&a};
            double bytes = 0.0;   // lower triangle of the trailing matrix, once per column
            for (int jj = k; jj < k + kb; jj++) bytes += 0.5 * (double)(n - jj - 1) * (double)(n - jj) * sizeof(T);
            {
                ProfScope ps(s, sizeof(T) == 8 ? "sytrd_panel_f64" : "sytrd_panel_f32", bytes * B);
                GMC_TRY(launch_persistent(kfn, dim3(B * G), dim3(SY_THREADS), args, lds, s));
            }
            epoch += 2 * kb;   // two group barriers per column
            const int k1 = k + kb;
            if (k1 < n) {
                const int mt = n - k1;
                GemmShape g{mt, mt, 2 * kb, false, true, (int64_t)n, (int64_t)n, (int64_t)n,
                            (int64_t)n * 3 * nb, (int64_t)n * 3 * nb, (int64_t)nn, B, 1};
                GMC_TRY(gemm<T>(s, g, T(-1), p.X + k1, p.X + (size_t)kb * n + k1, T(1), A + k1 + (size_t)k1 * n));
            }
        }
    }
    hipLaunchKernelGGL(sytrd_finish_kernel<T>, dim3((B + 255) / 256), dim3(256), 0, s, A, p.d, n, B);
    hipLaunchKernelGGL(clear_upper_kernel<T>, dim3(grid1(nn), B), dim3(256), 0, s, A, n);
    if (padded) {
        hipLaunchKernelGGL(pad_tail_kernel<T>, dim3(B), dim3(256), 0, s, p.d, p.e, p.n, n);
        hipLaunchKernelGGL(pad_vecs_kernel<T>, dim3(grid1((int64_t)p.n * p.n), B), dim3(256), 0, s, A, n, Au, p.n);
        A = Au;   // the reflectors at order p.n; Q is the divide & conquer's from here
    }
    GMC_HIP(hipGetLastError());
    ps_tri.close();
    ProfScope ps_dc(s, f64 ? "eig_dc_f64" : "eig_dc_f32", 0.0);

    // ---------------- divide & conquer
    DcBufs<T> b{p.e, w, p.Q, p.QK, p.QN, p.UB, p.dsort, p.zsort, p.dK, p.zK, p.taus, p.zh, p.EV, p.rotc, p.rots,
                p.perm, p.lst, p.lstD, p.org, p.rank, p.rotp, p.rotj, p.kcnt, p.rcnt, p.kpos, p.kc1, p.kc2};
    hipLaunchKernelGGL(dc_init_kernel<T>, dim3(grid1(nn), B), dim3(256), 0, s, p.d, p.e, p.d_split, w, p.Q, n);
    GMC_HIP(hipGetLastError());
    for (auto& L : p.levels) {
        DcArgs a{n, L.nmerge, B, L.d_desc};
        hipLaunchKernelGGL(dc_prep_kernel<T>, dim3(B * L.nmerge), dim3(256), 0, s, a, b);
        hipLaunchKernelGGL(dc_rot_kernel<T>, dim3(L.nmerge, B), dim3(256), 0, s, a, b);
        const unsigned wg = (unsigned)(((int64_t)B * n * 64 + 255) / 256);
        hipLaunchKernelGGL(dc_secular_kernel<T>, dim3(wg), dim3(256), 0, s, a, b);
        hipLaunchKernelGGL(dc_zhat_kernel<T>, dim3(wg), dim3(256), 0, s, a, b);
        hipLaunchKernelGGL(dc_vec_kernel<T>, dim3(wg), dim3(256), 0, s, a, b);
        {
            ProfScope ps(s, sizeof(T) == 8 ? "dc_merge_gemm_f64" : "dc_merge_gemm_f32", 0.0);
            hipLaunchKernelGGL(dc_gemm_kernel<T>, dim3((unsigned)(B * L.tiles)), dim3(256), 0, s, a, b, L.tiles);
        }
        hipLaunchKernelGGL(dc_final1_kernel<T>, dim3(n, B), dim3(256), 0, s, a, b);
        hipLaunchKernelGGL(dc_rank_kernel<T>, dim3((n + 255) / 256, B), dim3(256), 0, s, a, b);
        hipLaunchKernelGGL(dc_final2_kernel<T>, dim3(n, B), dim3(256), 0, s, a, b);
        GMC_HIP(hipGetLastError());
    }

    ps_dc.close();
    // ---------------- back transformation  Z <- H(0)...H(n-2) Z,  Z = Q
    ProfScope ps_bt(s, f64 ? "eig_backtr_f64" : "eig_backtr_f32", 0.0);
    const int nbb = p.nbb;
    T* Gm = p.Tf;
    T* Tm = p.Tf + (size_t)B * nbb * nbb;
    if (n > 1) {
        const size_t lds_t = ((size_t)nbb * (nbb + 1) / 2 + 2 * (size_t)nbb) * sizeof(T);
        if (!capturing(s))
            GMC_HIP(hipFuncSetAttribute((const void*)larft_kernel<T>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                        (int)lds_t));
        // H(0) ... H(nu-2) of order nu = p.n (a padded reduction's pad rows
        // and reflectors are zero and identities) on the first nu columns of
        // Z (ld n): V has ld nu, Z and the block products ld n
        const int nu = p.n;
        const int64_t nun = (int64_t)nu * nu;
        const int last = ((nu - 2) / nbb) * nbb;
        // Every block's T factor depends on its reflectors alone: when the
        // Gram matrices and factors of all blocks fit in the divide &
        // conquer's UB (free from here on), they are formed before the loop,
        // the factors in one launch over (problem, block), instead of one
        // sequential 16-workgroup launch per block
        const int nblk = last / nbb + 1;
        const size_t blk = (size_t)B * nbb * nbb;
        const bool hoist = 2 * (size_t)nblk * nbb * nbb <= nn;
        T* Gall = p.UB;
        T* Tall = p.UB + (size_t)nblk * blk;
        if (hoist) {
            for (int k = 0; k <= last; k += nbb) {
                const int kb = std::min(nbb, nu - 1 - k);
                const T* V = A + (k + 1) + (size_t)k * nu;
                GemmShape g1{kb, kb, nu - k - 1, true, false, (int64_t)nu, (int64_t)nu, (int64_t)nbb, nun, nun,
                             (int64_t)nbb * nbb, B, 0};
                GMC_TRY(gemm<T>(s, g1, T(1), V, V, T(0), Gall + (size_t)(k / nbb) * blk));
            }
            hipLaunchKernelGGL(larft_kernel<T>, dim3(B, nblk), dim3(256), lds_t, s, Gall, p.tau, Tall, n, nu, 0, nbb);
            GMC_HIP(hipGetLastError());
        }
        for (int k = last; k >= 0; k -= nbb) {
            const int kb = std::min(nbb, nu - 1 - k);
            const int m = nu - k - 1;
            const T* V = A + (k + 1) + (size_t)k * nu;
            T* Zs = p.Q + (k + 1);
            if (hoist) {
                Tm = Tall + (size_t)(k / nbb) * blk;
            } else {
                GemmShape g1{kb, kb, m, true, false, (int64_t)nu, (int64_t)nu, (int64_t)nbb, nun, nun,
                             (int64_t)nbb * nbb, B, 0};
                GMC_TRY(gemm<T>(s, g1, T(1), V, V, T(0), Gm));
                hipLaunchKernelGGL(larft_kernel<T>, dim3(B, 1), dim3(256), lds_t, s, Gm, p.tau, Tm, n, nu, k, nbb);
                GMC_HIP(hipGetLastError());
            }
            GemmShape g2{kb, nu, m, true, false, (int64_t)nu, (int64_t)n, (int64_t)nbb, nun, (int64_t)nn,
                         (int64_t)nn, B, 0};
            // own_gemm: the library's MFMA tile (no library workspace)
            if (p.own_gemm) GMC_TRY(gemm<T>(s, g2, T(1), V, Zs, T(0), p.QK));
            else GMC_TRY(gemm_big<T>(s, g2, T(1), V, Zs, T(0), p.QK));
            GemmShape g3{kb, nu, kb, false, false, (int64_t)nbb, (int64_t)nbb, (int64_t)nbb, (int64_t)nbb * nbb,
                         (int64_t)nn, (int64_t)nn, B, 0};
            GMC_TRY(gemm<T>(s, g3, T(1), Tm, p.QK, T(0), p.QN));
            GemmShape g4{m, nu, kb, false, false, (int64_t)nu, (int64_t)nbb, (int64_t)n, nun, (int64_t)nn,
                         (int64_t)nn, B, 0};
            if (p.own_gemm) GMC_TRY(gemm<T>(s, g4, T(-1), V, p.QN, T(1), Zs));
            else GMC_TRY(gemm_big<T>(s, g4, T(-1), V, p.QN, T(1), Zs));
        }
    }
    if (padded)
        hipLaunchKernelGGL(pad_out_kernel<T>, dim3(grid1((int64_t)p.n * p.n), B), dim3(256), 0, s, p.Q, p.wp, n, Au,
                           wu, p.n);
    else
        hipLaunchKernelGGL(copy_kernel<T>, dim3(grid1((int64_t)B * nn)), dim3(256), 0, s, p.Q, A, (int64_t)B * nn);
    GMC_HIP(hipGetLastError());
    return 0;
}

template <class T>
int eig_check_error(hipStream_t s, EigPlan<T>& p)
{
    int h = 0;
    GMC_HIP(hipMemcpyAsync(&h, p.err, sizeof(int), hipMemcpyDeviceToHost, s));
    GMC_HIP(hipStreamSynchronize(s));
    if (h != 0) p.clear_err = true;
    GMC_CHECK(h == 0, "syevd: a workgroup barrier timed out (co-residency lost)");
    return 0;
}

template struct EigPlan<double>;
template struct EigPlan<float>;
template int eig_plan<double>(int, int, EigPlan<double>*);
template int eig_plan<float>(int, int, EigPlan<float>*);
template int eig_plan_at<double>(int, int, EigPlan<double>*, void*, size_t, size_t*, hipStream_t);
template int eig_plan_at<float>(int, int, EigPlan<float>*, void*, size_t, size_t*, hipStream_t);
template int eig_bind<double>(EigPlan<double>*, void*);
template int eig_bind<float>(EigPlan<float>*, void*);
template int syevd_batched<double>(hipStream_t, EigPlan<double>&, double*, double*);
template int syevd_batched<float>(hipStream_t, EigPlan<float>&, float*, float*);
template int eig_check_error<double>(hipStream_t, EigPlan<double>&);
template int eig_check_error<float>(hipStream_t, EigPlan<float>&);

}  // namespace gmc
