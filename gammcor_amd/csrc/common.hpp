// Shared definitions for the gfx950 AC-CAS library.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstdio>
#include <string>
#include <initializer_list>

namespace gmc {

// Last error message of the calling thread (returned by gmc_last_error()).
void set_error(const char* fmt, ...);
const char* last_error();

struct Status {
    int code = 0;
};

#define GMC_HIP(call)                                                                     \
    do {                                                                                  \
        hipError_t e_ = (call);                                                           \
        if (e_ != hipSuccess) {                                                           \
            ::gmc::set_error("%s:%d %s -> %s", __FILE__, __LINE__, #call,                 \
                             hipGetErrorString(e_));                                      \
            return -1;                                                                    \
        }                                                                                 \
    } while (0)

#define GMC_CHECK(cond, ...)                                                              \
    do {                                                                                  \
        if (!(cond)) {                                                                    \
            ::gmc::set_error(__VA_ARGS__);                                                \
            return -2;                                                                    \
        }                                                                                 \
    } while (0)

#define GMC_TRY(expr)                                                                     \
    do {                                                                                  \
        int rc_ = (expr);                                                                 \
        if (rc_ != 0) return rc_;                                                         \
    } while (0)

// Device scratch owned by the library: grows on demand, freed at finalize.
// Counts the reallocations of every Workspace (and other frees of device
// memory a captured HIP graph could read): a graph captured at one count is
// replayed only while the count is unchanged.
uint64_t workspace_generation();
void bump_workspace_generation();

struct Workspace {
    void* ptr = nullptr;
    size_t bytes = 0;
    int reserve(size_t need);
    template <class T>
    T* as(size_t offset_bytes = 0) const { return reinterpret_cast<T*>(static_cast<char*>(ptr) + offset_bytes); }
};

inline size_t align_up(size_t x, size_t a = 256) { return (x + a - 1) / a * a; }

// Sub-allocator over a Workspace, used to carve per-call scratch.  Carving
// past `cap` is recorded and must be checked with fits() before any use.
struct Arena {
    char* base;
    size_t cap;
    size_t off = 0;
    template <class T>
    T* take(size_t count) {
        size_t o = align_up(off);
        off = o + count * sizeof(T);
        return reinterpret_cast<T*>(base + o);
    }
    bool fits() const { return off <= cap; }
};

// Bytes needed for a sequence of carves of the given element counts/sizes.
inline size_t carve_bytes(std::initializer_list<size_t> bytes)
{
    size_t off = 0;
    for (size_t b : bytes) off = align_up(off) + b;
    return align_up(off);
}

inline int64_t pair_idx(int64_t i, int64_t j) {
    int64_t a = i > j ? i : j, b = i > j ? j : i;
    return a * (a + 1) / 2 + b;
}

__host__ __device__ inline int64_t pidx(int64_t i, int64_t j) {
    int64_t a = i > j ? i : j, b = i > j ? j : i;
    return a * (a + 1) / 2 + b;
}

// Tile t of an enumeration of the lower triangle of a tile grid, row by row
// (tile (ta, tb) with tb <= ta is number ta (ta + 1) / 2 + tb).
__device__ inline void lower_tile(int t, int& ta, int& tb)
{
    int a = (int)((sqrt(8.0 * (double)t + 1.0) - 1.0) * 0.5);
    while ((a + 1) * (a + 2) / 2 <= t) a++;
    while (a * (a + 1) / 2 > t) a--;
    ta = a;
    tb = t - a * (a + 1) / 2;
}

}  // namespace gmc
