// common.hpp -- shared device/host helpers of the gfx950 phase-field kernels:
// stored-array layout, boundary index maps, error plumbing of the C ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <string>

#include "../../include/phix_amd.h"

namespace phx {

// ---------------------------------------------------------------- errors ----
void set_error(const std::string& msg);
int fail(int code, const std::string& msg);

#define PHX_HIP(call)                                                             \
    do {                                                                          \
        hipError_t e_ = (call);                                                   \
        if (e_ != hipSuccess)                                                     \
            return ::phx::fail(PHX_ERR_HIP, std::string(#call) + ": " +           \
                                                hipGetErrorString(e_) + " (" +    \
                                                __FILE__ + ":" +                  \
                                                std::to_string(__LINE__) + ")");  \
    } while (0)

#define PHX_LAUNCH_CHECK(what)                                                    \
    do {                                                                          \
        hipError_t e_ = hipGetLastError();                                        \
        if (e_ != hipSuccess)                                                     \
            return ::phx::fail(PHX_ERR_HIP, std::string(what) + ": " +            \
                                                hipGetErrorString(e_));           \
    } while (0)

inline hipStream_t as_stream(phx_stream s) { return reinterpret_cast<hipStream_t>(s); }

// ---------------------------------------------------------------- layout ----
// Stored layout of one field: storedDims = n + 2g on all three axes,
// index = (i+g) + sx*((j+g) + sy*(k+g)). Offsets are 64-bit so fields larger
// than 2^31 cells (possible in 288 GB of HBM) index correctly.
struct Layout {
    int n0, n1, n2;     // physical cells
    int g;              // ghost layers
    int sx, sy, sz;     // stored dims
    long long sxy;      // sx*sy
    int dim;

    __host__ __device__ inline long long at(int i, int j, int k) const {
        return (long long)(i + g) + (long long)sx * (j + g) + sxy * (long long)(k + g);
    }
    __host__ inline long long total() const { return sxy * (long long)sz; }
};

inline Layout make_layout(const phx_grid* G) {
    Layout L;
    L.n0 = G->n[0]; L.n1 = G->n[1]; L.n2 = G->n[2];
    L.g = G->ghost;
    L.sx = L.n0 + 2 * L.g; L.sy = L.n1 + 2 * L.g; L.sz = L.n2 + 2 * L.g;
    L.sxy = (long long)L.sx * L.sy;
    L.dim = G->dim;
    return L;
}

// Validates dim, n, ghost and spacing. Returns 0 or sets the error.
int check_grid(const phx_grid* G, const char* who);
// Additionally requires a BC on every face of the active axes, n >= 2 there,
// and matching periodic sides. Used by the fused kernels, which read the
// halo through index maps instead of stored ghosts.
int check_grid_fused(const phx_grid* G, const char* who);
int check_dtype(int dtype, const char* who);

// ------------------------------------------------------ boundary index maps --
// Per-axis boundary description as the fused kernels see it.
template <typename T>
struct AxisBC {
    int n;
    int klo, khi;   // phx_bc_kind of low/high face
    T vlo, vhi;     // FIXED values
};

// Index map of a halo position p (radius <= 2) onto a physical cell for the
// stencil INPUT state, or -1 if the position holds a FIXED constant.
//   PERIODIC: p mod n                    (ghost layer l -> n-l, n-1+l -> l-1)
//   NOFLUX  : mirror about the face       (-1 -> 0, -2 -> 1; n -> n-1, n+1 -> n-2)
//             -1/n are exactly the reference's zero-gradient ghosts; -2/n+1 are
//             only ever used to rebuild mu on the ghost layer, see mu_ghost_*.
//   FIXED   : constant
// Positions outside [-2, n+1] clamp into range (they only feed lanes whose
// results are discarded).
template <typename T>
__device__ __forceinline__ int map_in(int p, const AxisBC<T>& b) {
    if (p >= 0 && p < b.n) return p;
    if (p < 0) {
        if (p < -2) p = -2;
        if (b.klo == PHX_BC_PERIODIC) return p + b.n;
        if (b.klo == PHX_BC_FIXED) return -1;
        return -1 - p;  // NOFLUX mirror
    }
    if (p > b.n + 1) p = b.n + 1;
    if (b.khi == PHX_BC_PERIODIC) return p - b.n;
    if (b.khi == PHX_BC_FIXED) return -1;
    return 2 * b.n - 1 - p;
}

template <typename T>
__device__ __forceinline__ T fixed_value(int p, const AxisBC<T>& b) {
    return p < 0 ? b.vlo : b.vhi;
}

// ------------------------------------------------------------ misc helpers --
__host__ __device__ inline int cdiv(long long a, int b) { return (int)((a + b - 1) / b); }

// Benchmark launch timer (phx_ch_timing): while enabled, a launch bracketed by
// these two calls is timed with HIP events on its stream and tagged with a
// kind. Kinds: 0-3 the CH RK4 kernels, 4 the Allen-Cahn one-stage kernel,
// 5 / 6 its two-stage kernels (stages 1+2 / 3+4), 7 the ensemble one-stage
// kernel, 8 / 9 its two-stage kernels, 10 the GFA step kernel. Not for
// launches under graph capture.
void launch_timer_begin(hipStream_t st);
void launch_timer_end(int kind, hipStream_t st);

}  // namespace phx
