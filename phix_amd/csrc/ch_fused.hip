// ch_fused.hip -- fused Cahn-Hilliard kernels for gfx950.
//
// The reference evaluates dc/dt = M*lap(f'(c) - kappa*lap(c), bcs) as ~10
// launches per RHS: memset, a pw pass, a lap pass into a scratch mu, one BC
// launch per axis on mu, a lap pass into rhs (src/equation/Equation.cu:537-651),
// plus separate RK4 tmp/update passes (src/solver/Solver.cu:192-259). Here one
// kernel does a whole RK4 stage:
//
//   halo load of the stage state with the BCs as index maps   (HBM -> LDS)
//   mu = f'(u) - kappa*lap(u) on the tile plus a 1-cell ring   (LDS)
//   k  = M*lap(mu)                                              (registers)
//   stage epilogue: next stage state and the RK4 accumulator   (-> HBM)
//
// so per stage a cell is read once (its tile halo comes from L2) and the
// stage state, accumulator and c_n are the only HBM traffic.
//
// Boundary semantics reproduced exactly:
//   * the stage state's ghost layer is the BC fill of the stage state
//     (Solver.cu:206, BCs on phi_tmp), realised by map_in() on the halo load;
//   * mu's ghost layer is the BC fill of mu, NOT mu evaluated at a ghost cell
//     (Equation.cu:571-572 applies the same bcs to the materialised scratch).
//     PERIODIC: mu at ghost p equals mu at p mod n; computing it from the
//     wrapped halo gives the same operands in the same order, bit for bit.
//     NOFLUX: mu at ghost -1 equals mu(0); the one-stage kernels evaluate
//     the stencil at the LDS centre of the image cell (the same operands as
//     the image's own evaluation), the blocked kernels exchange the +/-
//     neighbours of the ghost axis of the mirrored halo -- bit for bit either
//     way.
//     FIXED: mu at the ghost is the BC constant.
//
// FP64 keeps the reference's operation order and no FMA contraction, so the
// FP64 path reproduces the FP64 oracle bit for bit. The FP64 RK4 accumulator
// stores k1+2k2+2k3 and finishes with c + dt/6*(acc+k4), the reference's
// expression (Solver.cu:61-71). FP32 uses the 16-pass form
// acc = c + dt/6*k1 + dt/3*(k2+k3) which does not re-read c_n in stage 4.
#pragma clang fp contract(off)

#include "common.hpp"

#include <algorithm>
#include <map>
#include <mutex>
#include <tuple>
#include <type_traits>
#include <vector>

namespace phx {

// ------------------------------------------------------------- parameters --
// Coefficients of the FP32 arithmetic form (below), rounded once from double
// on the host; unused by FP64.
template <typename T>
struct FastCoef {
    T icc;                // -2 (idx2 + idy2 + idz2)
    T mx, my, mz, mcc;    // M*idx2, M*idy2, M*idz2, -2 M (idx2 + idy2 + idz2)
    T ntrca, ncs;         // -2 rho ca, -(ca + cb)
};

template <typename T>
struct CHArgs {
    const T* u;     // stage state (stencil input)
    const T* c;     // c_n
    T* acc;         // RK4 accumulator (exact FP64 form); 3D FP64: b = u2 + 2 u3 - c_n
    T* out;         // next stage state, or c for stage 4
    Layout L;
    AxisBC<T> bx, by, bz;
    T idx2, idy2, idz2;
    T two_rho, ca, cb, nkappa, M;
    T a;            // stage shift dt/2, dt/2, dt
    T dt6, dt3;
    // 3D FP32: stage 4 reads the stage states u2, u3 instead of an accumulator
    const T* s2;
    const T* s3;
    T third;        // 1/3
    FastCoef<T> F;  // FP32 arithmetic form
    int stage;      // 1..4
    int zc;         // z planes per block (3D)
    int xcd_remap;  // 1: XCD-aware tile order (grid size a multiple of 8)
};

template <typename T>
__device__ __forceinline__ T dwell(T v, T two_rho, T ca, T cb) {
    // 2.0*rho*(c-ca)*(c-cb)*(2.0*c-ca-cb), evaluated left to right
    return two_rho * (v - ca) * (v - cb) * ((T)2 * v - ca - cb);
}

__device__ __forceinline__ float fma_t(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ __forceinline__ double fma_t(double a, double b, double c) { return __builtin_fma(a, b, c); }

// The arithmetic of one cell, shared by every kernel so that their results
// agree bit for bit. FP64 (EXACT) keeps the reference's operation order:
//   lap = (xp - 2c + xm)*idx2, then + (yp - 2c + ym)*idy2, then + z
//         (Equation.cu:164-180);
//   mu  = 2rho*(c-ca)*(c-cb)*(2c-ca-cb) + (-kappa)*lap
//         (Cahn-Hillard_double-well.cu:87-90, Equation.cu:578-591);
//   k   = M*lap(mu);  stage state c + a*k.
// FP32 is the bandwidth path and is held to a stated error against the FP64
// oracle instead; it uses explicit FMAs (contraction stays off file-wide):
//   lap = fma(zp+zm, idz2, fma(yp+ym, idy2, fma(xp+xm, idx2, c*icc)));
//   f'  = fma(2rho, c, -2rho*ca) * (c-cb) * fma(2, c, -(ca+cb)), exactly 0
//         at c = cb (and at c = ca whenever 2rho*ca is exact);
//   mu  = fma(-kappa, lap, f');
//   k   = the lap form with M folded into the coefficients;
//   stage state fma(a, k, c).
// The neighbour pairs are sums, so exchanging the +/- neighbours of an axis
// (NOFLUX images in the blocked kernels) leaves the FP32 result unchanged too.
template <typename T, bool EXACT, int DIM, class Args>
__device__ __forceinline__ T mu_at(const Args& A, T c, T xp, T xm, T yp, T ym, T zp, T zm) {
    if constexpr (EXACT) {
        T val = (xp - (T)2 * c + xm) * A.idx2;
        if constexpr (DIM >= 2) val += (yp - (T)2 * c + ym) * A.idy2;
        if constexpr (DIM >= 3) val += (zp - (T)2 * c + zm) * A.idz2;
        return dwell(c, A.two_rho, A.ca, A.cb) + A.nkappa * val;
    } else {
        T val = fma_t(xp + xm, A.idx2, c * A.F.icc);
        if constexpr (DIM >= 2) val = fma_t(yp + ym, A.idy2, val);
        if constexpr (DIM >= 3) val = fma_t(zp + zm, A.idz2, val);
        const T fp = fma_t(A.two_rho, c, A.F.ntrca) * (c - A.cb) * fma_t((T)2, c, A.F.ncs);
        return fma_t(A.nkappa, val, fp);
    }
}

template <typename T, bool EXACT, int DIM, class Args>
__device__ __forceinline__ T k_at(const Args& A, T m, T xp, T xm, T yp, T ym, T zp, T zm) {
    if constexpr (EXACT) {
        T val = (xp - (T)2 * m + xm) * A.idx2;
        if constexpr (DIM >= 2) val += (yp - (T)2 * m + ym) * A.idy2;
        if constexpr (DIM >= 3) val += (zp - (T)2 * m + zm) * A.idz2;
        return A.M * val;
    } else {
        T val = fma_t(xp + xm, A.F.mx, m * A.F.mcc);
        if constexpr (DIM >= 2) val = fma_t(yp + ym, A.F.my, val);
        if constexpr (DIM >= 3) val = fma_t(zp + zm, A.F.mz, val);
        return val;
    }
}

// c + a*k
template <typename T, bool EXACT>
__device__ __forceinline__ T axpy(T a, T k, T c) {
    if constexpr (EXACT) return c + a * k;
    else return fma_t(a, k, c);
}

// FP32 RK4 update of the 3D kernels from the stage states u2 = c + dt/2 k1,
// u3 = c + dt/2 k2, u4 = c + dt k3 and k4:
//   c_{n+1} = c + dt/6 k4 + ((u2 - c) + 2 (u3 - c) + (u4 - c)) / 3,
// the classical RK4 sum with the increments recovered from the states. No
// accumulator field: a step reads and writes 13 fields instead of 16
// (stage 1: c -> u2; 2: u2, c -> u3; 3: u3, c -> u4; 4: u4, c, u2, u3 -> c).
template <typename T>
__device__ __forceinline__ T rk4_fin(T dt6, T third, T c, T u2, T u3, T u4, T k) {
    const T d = (u2 - c) + fma_t((T)2, u3 - c, u4 - c);
    return c + fma_t(dt6, k, third * d);
}

// FP64 form of the 3D kernels (12 field passes per step): the same sum
// regrouped as c_{n+1} = u4 + (b - 2 u4) / 3 + dt/6 k4 with b = u2 + 2 u3 - c,
// which stage 2 forms (it holds u2, c and u3 at the cell) and writes to a
// field of its own, so stage 4 reads b instead of c, u2 and u3
// (stage 1: c -> u2; 2: u2, c -> u3, b; 3: u3, c -> u4; 4: u4, b -> c).
// b is rounded once at the scale of c (u2 - c is exact whenever the two lie
// within a factor 2, as consecutive RK4 states do); b - 2 u4 is then exact
// for the same reason and of increment size, so 1/3, itself rounded, scales
// only an increment and adds no drift.
// One more rounding at the scale of c per step than rk4_fin (unbiased, ~1e-16
// in FP64), which is why FP32 keeps rk4_fin.
template <typename T>
__device__ __forceinline__ T rk4_b(T u2, T u3, T c) {
    return fma_t((T)2, u3, u2 - c);
}
template <typename T>
__device__ __forceinline__ T rk4_fin_b(T dt6, T third, T b, T u4, T k) {
    return u4 + fma_t(dt6, k, third * fma_t((T)-2, u4, b));
}
// The 3D kernels use the b form for FP64 outside the exact (oracle-order) mode.
template <typename T, bool EXACT>
constexpr bool kBForm = sizeof(T) == 8 && !EXACT;

// The six FIXED constants (x lo, x hi, y lo, y hi, z lo, z hi) staged in LDS
// by each block: reading them there in divergent ghost code keeps the
// compiler from turning a data-dependent choice between kernel arguments into
// a vector memory load, whose wait would also drain the prefetched planes.
template <typename T, typename Args>
__device__ __forceinline__ void stage_bc_values(T* bcv, const Args& A) {
    if (threadIdx.x == 0) {
        bcv[0] = A.bx.vlo;
        bcv[1] = A.bx.vhi;
        bcv[2] = A.by.vlo;
        bcv[3] = A.by.vhi;
        bcv[4] = A.bz.vlo;
        bcv[5] = A.bz.vhi;
    }
}

// A workgroup-uniform copy of an axis's BC, read once into scalar registers.
// Per-plane code then chooses between its fields with scalar selects; left
// to itself the compiler may choose between kernel-argument addresses and
// issue a vector memory load for a uniform value, whose wait drains the
// prefetched planes.
__device__ __forceinline__ int uni(int v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ __forceinline__ float uni(float v) {
    return __builtin_bit_cast(float, __builtin_amdgcn_readfirstlane(__builtin_bit_cast(int, v)));
}
__device__ __forceinline__ double uni(double v) {
    const long long b = __builtin_bit_cast(long long, v);
    const int lo = __builtin_amdgcn_readfirstlane((int)(b & 0xffffffffLL));
    const int hi = __builtin_amdgcn_readfirstlane((int)(b >> 32));
    return __builtin_bit_cast(double, ((long long)hi << 32) | (unsigned int)lo);
}
template <typename T>
__device__ __forceinline__ AxisBC<T> uni_axis(const AxisBC<T>& b) {
    AxisBC<T> u;
    u.n = uni(b.n);
    u.klo = uni(b.klo);
    u.khi = uni(b.khi);
    u.vlo = uni(b.vlo);
    u.vhi = uni(b.vhi);
    return u;
}

// In-plane ghost rule of a position (y decides before x):
// FIX -> the FIXED constant of axis FY ? y : x, side HI ? high : low;
// SWX/SWY -> exchange the +/- neighbours of that axis (NOFLUX image).
enum : int { GC_FIX = 1, GC_SWX = 2, GC_SWY = 4, GC_FY = 8, GC_HI = 16 };

template <typename T>
__device__ __forceinline__ T ghost_value_xy(int code, const T* bcv) {
    return bcv[((code & GC_FY) ? 2 : 0) + ((code & GC_HI) ? 1 : 0)];
}

// Stage epilogue for one physical output cell; cn, an: c_n and the
// accumulator at the cell (loaded beforehand, used by stages 2-4). Stage 0 is
// an explicit Euler step, u + a k with a = dt (no accumulator).
template <typename T, bool EXACT>
__device__ __forceinline__ void epilogue(const CHArgs<T>& A, long long gi, T uc, T k, T cn, T an) {
    const int s = A.stage;
    if (s == 0) {
        A.out[gi] = EXACT ? uc + A.a * k : axpy<T, false>(A.a, k, uc);
        return;
    }
    if (EXACT) {
        if (s == 1) {
            A.acc[gi] = k;
            A.out[gi] = uc + A.a * k;
        } else if (s < 4) {
            A.acc[gi] = an + (T)2 * k;
            A.out[gi] = cn + A.a * k;
        } else {
            A.out[gi] = cn + A.dt6 * (an + k);
        }
    } else {
        if (s == 1) {
            A.acc[gi] = axpy<T, false>(A.dt6, k, uc);
            A.out[gi] = axpy<T, false>(A.a, k, uc);
        } else if (s < 4) {
            A.acc[gi] = axpy<T, false>(A.dt3, k, an);
            A.out[gi] = axpy<T, false>(A.a, k, cn);
        } else {
            A.out[gi] = axpy<T, false>(A.dt6, k, an);
        }
    }
}

// --------------------------------------------------------- 2D (and 1D) ----
// One block = TX x TY output tile. LDS holds the (TX+4)x(TY+4) halo of the
// stage state and the (TX+2)x(TY+2) mu ring.
template <typename T, int DIM, int TX, int TY, bool EXACT>
__global__ __launch_bounds__(256) void k_ch_stage_2d(CHArgs<T> A) {
    constexpr int UX = TX + 4, UY = TY + 4, MX = TX + 2, MY = TY + 2;
    __shared__ T U[UY][UX];
    __shared__ T MU[MY][MX];
    __shared__ T BCV[6];
    const int tid = threadIdx.x;
    int bx = blockIdx.x, by = blockIdx.y;
    if (A.xcd_remap) {   // XCD-aware tile order (row bands per XCD), as in k_ch_stage_3d
        const int gx = gridDim.x, nb = gx * gridDim.y;
        const int h = bx + gx * by;
        const int lb = (h & 7) * (nb >> 3) + (h >> 3);
        bx = lb % gx;
        by = lb / gx;
    }
    const int x0 = bx * TX, y0 = by * TY;
    const int kz = 0;
    stage_bc_values(BCV, A);

    // Halo load: every slot's offset (BC maps applied) first, then all loads
    // back to back with no branch between them, so their latencies overlap;
    // FIXED and surplus slots read a valid cell and are replaced at the store.
    constexpr int NS = (UX * UY + 255) / 256;
    const long long zb = (long long)A.L.g * A.L.sxy;
    const int cell0 = A.L.g * A.L.sx + A.L.g;
    int off[NS];
    unsigned fm = 0;
    T fv[NS];
#pragma unroll
    for (int i = 0; i < NS; ++i) {
        const int t = min(tid + 256 * i, UX * UY - 1);
        const int ly = t / UX, lx = t - ly * UX;
        const int px = x0 - 2 + lx, py = DIM >= 2 ? y0 - 2 + ly : 0;
        const int qy = DIM >= 2 ? map_in(py, A.by) : 0, qx = map_in(px, A.bx);
        off[i] = cell0;
        fv[i] = (T)0;
        if (qy < 0) {
            fm |= 1u << i;
            fv[i] = fixed_value(py, A.by);
        } else if (qx < 0) {
            fm |= 1u << i;
            fv[i] = fixed_value(px, A.bx);
        } else {
            off[i] = (qy + A.L.g) * A.L.sx + (qx + A.L.g);
        }
    }
    T r[NS];
#pragma unroll
    for (int i = 0; i < NS; ++i) r[i] = A.u[zb + off[i]];
    // epilogue operands of the tile cells, in flight with the halo; an
    // operand the stage does not need is read at element 0 of the stage state
    // by every lane (one cache line)
    constexpr int NO = (TX * TY) / 256;
    static_assert((TX * TY) % 256 == 0, "tile must be a multiple of 256 cells");
    const int st = A.stage;
    const bool need_c = st == 2 || st == 3 || (EXACT && st == 4), need_a = st >= 2;
    const T* csrc = need_c ? A.c : A.u;
    const T* asrc = need_a ? A.acc : A.u;
    const long long cmask = need_c ? -1LL : 0LL, amask = need_a ? -1LL : 0LL;
    T pc[NO], pa[NO];
#pragma unroll
    for (int j = 0; j < NO; ++j) {
        const int t = tid + 256 * j;
        const int oy = t / TX, ox = t - oy * TX;
        const long long gi = A.L.at(min(x0 + ox, A.L.n0 - 1), DIM >= 2 ? min(y0 + oy, A.L.n1 - 1) : 0, kz);
        pc[j] = csrc[gi & cmask];
        pa[j] = asrc[gi & amask];
    }
#pragma unroll
    for (int i = 0; i < NS; ++i) {
        const int t = tid + 256 * i;
        if (t < UX * UY) (&U[0][0])[t] = ((fm >> i) & 1u) ? fv[i] : r[i];
    }
    __syncthreads();
    // mu at a ghost position is the BC fill of mu: evaluated at the LDS
    // centre of its BC image (mirror for NOFLUX, its own wrapped position for
    // PERIODIC), which reproduces the image's own evaluation bit for bit, or
    // the FIXED constant (y decides before x).
    for (int t = tid; t < MX * MY; t += 256) {
        const int my = t / MX, mx = t - my * MX;
        const int px = x0 - 1 + mx, py = DIM >= 2 ? y0 - 1 + my : 0;
        int ix = px, iy = py, code = 0;
        if (DIM >= 2 && (py < 0 || py >= A.L.n1)) {
            const int k = py < 0 ? A.by.klo : A.by.khi;
            if (k == PHX_BC_FIXED) code = GC_FIX | GC_FY | (py < 0 ? 0 : GC_HI);
            else if (k == PHX_BC_NOFLUX) iy = py < 0 ? -1 - py : 2 * A.L.n1 - 1 - py;
        }
        if (!code && (px < 0 || px >= A.L.n0)) {
            const int k = px < 0 ? A.bx.klo : A.bx.khi;
            if (k == PHX_BC_FIXED) code = GC_FIX | (px < 0 ? 0 : GC_HI);
            else if (k == PHX_BC_NOFLUX) ix = px < 0 ? -1 - px : 2 * A.L.n0 - 1 - px;
        }
        ix = min(max(ix, x0 - 1), x0 + TX);   // only partial tiles' unused cells clamp
        const int lyc = DIM >= 2 ? min(max(iy, y0 - 1), y0 + TY) - y0 + 2 : 2;
        const int lxc = ix - x0 + 2;
        const T c = U[lyc][lxc];
        T m = DIM >= 2 ? mu_at<T, EXACT, 2>(A, c, U[lyc][lxc + 1], U[lyc][lxc - 1], U[lyc + 1][lxc], U[lyc - 1][lxc],
                                             (T)0, (T)0)
                       : mu_at<T, EXACT, 1>(A, c, U[lyc][lxc + 1], U[lyc][lxc - 1], (T)0, (T)0, (T)0, (T)0);
        if (code & GC_FIX) m = ghost_value_xy(code, BCV);
        MU[my][mx] = m;
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < NO; ++j) {
        const int t = tid + 256 * j;
        int oy = t / TX, ox = t - oy * TX;
        int x = x0 + ox, y = y0 + oy;
        if (x >= A.L.n0 || y >= A.L.n1) continue;
        T m = MU[oy + 1][ox + 1];
        T k = DIM >= 2 ? k_at<T, EXACT, 2>(A, m, MU[oy + 1][ox + 2], MU[oy + 1][ox], MU[oy + 2][ox + 1], MU[oy][ox + 1],
                                            (T)0, (T)0)
                       : k_at<T, EXACT, 1>(A, m, MU[oy + 1][ox + 2], MU[oy + 1][ox], (T)0, (T)0, (T)0, (T)0);
        epilogue<T, EXACT>(A, A.L.at(x, y, kz), U[oy + 2][ox + 2], k, pc[j], pa[j]);
    }
}

// ------------------------------------------------------------------- 3D ----
// One block = TX x TY columns of a z-chunk (the grid's z extent splits the
// planes into balanced chunks, about A.zc planes each), streaming along z
// with a ring of 4 halo planes of the stage state and 3 mu planes in LDS.
// Software pipelined: while plane o is computed, the halo plane o+3 and the
// next output plane's c/acc values are already in flight to registers, so the
// HBM latency overlaps the LDS work instead of sitting behind a barrier.
// Per-thread load offsets (with the x/y BC maps folded in) are computed once.
//
// Ring reuse, with barriers A and B of every iteration o:
//   halo plane o+2 is stored (before A) into the slot of plane o-2, whose last
//   readers (mu(o-1) and output o-2) ran before iteration o-1's barriers;
//   mu(o+1) is written (between A and B) into the slot of mu(o-2), last read
//   by output o-1 before barrier A of iteration o. mu of a NOFLUX ghost plane
//   o+1 = n_z reads the image plane o and its neighbours o-1..o+1, all held by
//   the ring in iteration o (plane o-1's slot is rewritten in iteration o+1).
template <typename T, int TX, int TY, bool EXACT>
__global__ __launch_bounds__(256) void k_ch_stage_3d(CHArgs<T> A) {
    constexpr int UX = TX + 4, UY = TY + 4, MX = TX + 2, MY = TY + 2;
    constexpr int NU = (UX * UY + 255) / 256;
    constexpr int NM = (MX * MY + 255) / 256;
    constexpr int NO = (TX * TY) / 256;
    static_assert((TX * TY) % 256 == 0, "tile must be a multiple of 256 cells");
    __shared__ T U[4][UY * UX];
    __shared__ T MU[3][MY * MX];
    __shared__ T BCV[6];
    const int tid = threadIdx.x;
    stage_bc_values(BCV, A);
    // XCD-aware tile order: workgroups are dispatched round-robin over the 8
    // XCDs by linear id, so hand each XCD a contiguous range of tiles (whole
    // xy slabs of consecutive z-chunks) -- tiles sharing halos then share an L2.
    int bx = blockIdx.x, by = blockIdx.y, bz = blockIdx.z;
    if (A.xcd_remap) {
        const int gx = gridDim.x, gy = gridDim.y;
        const int nb = gx * gy * gridDim.z;
        const int h = bx + gx * (by + gy * bz);
        const int lb = (h & 7) * (nb >> 3) + (h >> 3);
        bx = lb % gx;
        by = (lb / gx) % gy;
        bz = lb / (gx * gy);
    }
    const int x0 = bx * TX, y0 = by * TY;
    // balanced z chunks: gridDim.z chunks of floor or ceil(n2 / gridDim.z) planes
    const int z0 = (int)(((long long)bz * A.L.n2) / gridDim.z);
    const int z1 = (int)(((long long)(bz + 1) * A.L.n2) / gridDim.z);
    const int g = A.L.g;
    const long long sxy = A.L.sxy;

    // Halo slots: in-plane offset with the x/y maps applied. Every slot loads
    // unconditionally from a valid cell (FIXED and surplus slots read the
    // plane's first cell), so no load sits in a divergent branch and the
    // prefetched plane stays in flight across the pipeline; the FIXED
    // constants replace the loaded values when the plane is stored.
    int hoff[NU];
    T hfix[NU];
    unsigned hfm = 0;   // bit i: slot i holds a FIXED constant
#pragma unroll
    for (int i = 0; i < NU; ++i) {
        const int t = tid + 256 * i;
        hoff[i] = g * A.L.sx + g;
        hfix[i] = (T)0;
        if (t < UX * UY) {
            const int ly = t / UX, lx = t - ly * UX;
            const int px = x0 - 2 + lx, py = y0 - 2 + ly;
            const int qy = map_in(py, A.by), qx = map_in(px, A.bx);
            if (qy < 0) {
                hfm |= 1u << i;
                hfix[i] = fixed_value(py, A.by);
            } else if (qx < 0) {
                hfm |= 1u << i;
                hfix[i] = fixed_value(px, A.bx);
            } else {
                hoff[i] = (qy + g) * A.L.sx + (qx + g);
            }
        }
    }
    // output slots (cells beyond the domain read a valid cell, never store)
    int ooff[NO];
    bool oval[NO];
#pragma unroll
    for (int j = 0; j < NO; ++j) {
        const int t = tid + 256 * j;
        const int oy = t / TX, ox = t - oy * TX;
        const int x = x0 + ox, y = y0 + oy;
        oval[j] = x < A.L.n0 && y < A.L.n1;
        ooff[j] = oval[j] ? (y + g) * A.L.sx + (x + g) : g * A.L.sx + g;
    }

    // a plane in flight: its registers, and its FIXED z constant if any
    struct Plane {
        T r[NU];
        bool fix;
        T val;
    };
    auto load_plane = [&](int pz, Plane& P) {
        const int qz = map_in(pz, A.bz);
        P.fix = qz < 0;
        P.val = P.fix ? fixed_value(pz, A.bz) : (T)0;
        const T* base = A.u + (long long)((P.fix ? 0 : qz) + g) * sxy;
#pragma unroll
        for (int i = 0; i < NU; ++i) P.r[i] = base[hoff[i]];
    };
    auto store_plane = [&](int pz, const Plane& P) {
        T* dst = U[(pz + 4) & 3];
#pragma unroll
        for (int i = 0; i < NU; ++i)
            if (tid + 256 * i < UX * UY) dst[tid + 256 * i] = P.fix ? P.val : (((hfm >> i) & 1u) ? hfix[i] : P.r[i]);
    };
    // mu slots. mu at a ghost position is the BC fill of mu: the value at the
    // mirror image for NOFLUX, the value computed from the wrapped halo (its
    // own position) for PERIODIC, the constant for FIXED. Each slot therefore
    // evaluates the stencil at the LDS centre of its image position (x and y
    // images precomputed here, z by taking the image plane's ring slots),
    // which reproduces the image's own evaluation bit for bit; a FIXED slot
    // takes its constant. The plane loop has no per-point BC logic.
    int mc0[NM], mcode[NM];
#pragma unroll
    for (int i = 0; i < NM; ++i) {
        const int t = min(tid + 256 * i, MX * MY - 1);
        const int my = t / MX, mx = t - my * MX;
        const int px = x0 - 1 + mx, py = y0 - 1 + my;
        int ix = px, iy = py, code = 0;
        if (py < 0 || py >= A.L.n1) {   // y decides before x
            const int k = py < 0 ? A.by.klo : A.by.khi;
            if (k == PHX_BC_FIXED) code = GC_FIX | GC_FY | (py < 0 ? 0 : GC_HI);
            else if (k == PHX_BC_NOFLUX) iy = py < 0 ? -1 - py : 2 * A.L.n1 - 1 - py;
        }
        if (!code && (px < 0 || px >= A.L.n0)) {
            const int k = px < 0 ? A.bx.klo : A.bx.khi;
            if (k == PHX_BC_FIXED) code = GC_FIX | (px < 0 ? 0 : GC_HI);
            else if (k == PHX_BC_NOFLUX) ix = px < 0 ? -1 - px : 2 * A.L.n0 - 1 - px;
        }
        // positions past the first ghost layer (partial tiles) feed no
        // output; keep their stencil inside the halo
        ix = min(max(ix, x0 - 1), x0 + TX);
        iy = min(max(iy, y0 - 1), y0 + TY);
        mc0[i] = (iy - y0 + 2) * UX + (ix - x0 + 2);
        mcode[i] = code;
    }
    auto compute_mu = [&](int m) {
        // z: NOFLUX ghost planes evaluate at the image plane, FIXED ones are
        // the constant, PERIODIC ones use their own (wrapped) planes
        int mz = m;
        bool zfix = false;
        if (m < 0 || m >= A.L.n2) {
            const int k = m < 0 ? A.bz.klo : A.bz.khi;
            if (k == PHX_BC_FIXED) zfix = true;
            else if (k == PHX_BC_NOFLUX) mz = m < 0 ? -1 - m : 2 * A.L.n2 - 1 - m;
        }
        const T zv = BCV[m < 0 ? 4 : 5];
        const T* ul = U[(mz + 3) & 3];
        const T* um = U[(mz + 4) & 3];
        const T* uh = U[(mz + 5) & 3];
        T* dst = MU[(m + 3) % 3];
#pragma unroll
        for (int i = 0; i < NM; ++i) {
            const int t = tid + 256 * i;
            if (t < MX * MY) {
                const int c0 = mc0[i];
                const T c = um[c0];
                T v = mu_at<T, EXACT, 3>(A, c, um[c0 + 1], um[c0 - 1], um[c0 + UX], um[c0 - UX], uh[c0], ul[c0]);
                if (mcode[i] & GC_FIX) v = ghost_value_xy(mcode[i], BCV);
                dst[t] = zfix ? zv : v;
            }
        }
    };
    // epilogue operands of one output plane
    const int st = A.stage;
    // Both epilogue operands are loaded for every stage so that each
    // iteration issues the same loads (the compiler then waits for exactly the
    // ones it uses); an operand the stage does not need is read from the
    // first element of the stage state for every lane (one cache line).
    // Exact FP64 form: c_n (stages 2-4) and the accumulator (2-4); FP64 b
    // form: c_n (2-3) and b (4); FP32: c_n (2-4) and the stage states u2, u3 (4)
    constexpr bool BF = kBForm<T, EXACT>;
    const bool need_c = EXACT ? (st == 2 || st == 3 || st == 4) : (BF ? (st == 2 || st == 3) : st >= 2);
    const bool need_a = EXACT ? st >= 2 : st == 4;
    const bool need_b = !EXACT && !BF && st == 4;
    const T* csrc = need_c ? A.c : A.u;
    const T* asrc = need_a ? (EXACT || BF ? A.acc : A.s2) : A.u;
    const T* bsrc = need_b ? A.s3 : A.u;
    const long long cmask = need_c ? -1LL : 0LL, amask = need_a ? -1LL : 0LL, bmask = need_b ? -1LL : 0LL;
    T pc[NO], pa[NO], pq[NO], nc[NO], na[NO], nq[NO];
    auto load_epi = [&](int o, T (&rc)[NO], T (&ra)[NO], T (&rb)[NO]) {
        const long long pb = (long long)(o + g) * sxy;
#pragma unroll
        for (int j = 0; j < NO; ++j) {
            rc[j] = csrc[(pb + ooff[j]) & cmask];
            ra[j] = asrc[(pb + ooff[j]) & amask];
            if constexpr (!EXACT && !BF) rb[j] = bsrc[(pb + ooff[j]) & bmask];
        }
    };

    // ---- prologue: halo planes z0-2 .. z0+1, mu planes z0-1, z0; the five
    // planes are loaded in two batches (each batch's loads in flight together),
    // then halo planes z0+2 and z0+3 go in flight for the loop
    Plane pa0, pa1, p2;
    load_plane(z0 - 2, pa0);
    load_plane(z0 - 1, pa1);
    load_plane(z0, p2);
    store_plane(z0 - 2, pa0);
    store_plane(z0 - 1, pa1);
    store_plane(z0, p2);
    load_plane(z0 + 1, p2);
    load_plane(z0 + 2, pa0);
    load_epi(z0, pc, pa, pq);
    load_plane(min(z0 + 3, z1 + 1), pa1);
    store_plane(z0 + 1, p2);
    __syncthreads();
    compute_mu(z0 - 1);
    compute_mu(z0);
    __syncthreads();

    // Iteration o stores halo plane o+2 from the register set P and refills P
    // with plane o+4: each plane is in flight for two iterations. The loop runs
    // pairs of iterations on the two register sets (registers cannot be
    // indexed at run time), with one single iteration for an odd count.
    auto iteration = [&](int o, Plane& P) {
        store_plane(o + 2, P);
        __syncthreads();   // A
        // unconditional prefetch (clamped planes): the same loads every iteration
        load_plane(min(o + 4, z1 + 1), P);
        load_epi(min(o + 1, z1 - 1), nc, na, nq);
        compute_mu(o + 1);
        __syncthreads();   // B
        const T* mlo = MU[(o + 2) % 3];
        const T* mmid = MU[(o + 3) % 3];
        const T* mhi = MU[(o + 4) % 3];
        const T* uo = U[(o + 4) & 3];
        const long long pb = (long long)(o + g) * sxy;
#pragma unroll
        for (int j = 0; j < NO; ++j) {
            if (!oval[j]) continue;
            const int t = tid + 256 * j;
            const int oy = t / TX, ox = t - oy * TX;
            const int m0 = (oy + 1) * MX + ox + 1;
            const T mc = mmid[m0];
            const T k = k_at<T, EXACT, 3>(A, mc, mmid[m0 + 1], mmid[m0 - 1], mmid[m0 + MX], mmid[m0 - MX], mhi[m0],
                                          mlo[m0]);
            const long long gi = pb + ooff[j];
            const T uc = uo[(oy + 2) * UX + ox + 2];
            if (EXACT) {
                if (st == 0) {   // an explicit Euler step (multi-step Euler)
                    A.out[gi] = uc + A.a * k;
                } else if (st == 1) {
                    A.acc[gi] = k;
                    A.out[gi] = uc + A.a * k;
                } else if (st < 4) {
                    A.acc[gi] = pa[j] + (T)2 * k;
                    A.out[gi] = pc[j] + A.a * k;
                } else {
                    A.out[gi] = pc[j] + A.dt6 * (pa[j] + k);
                }
            } else {
                if (st == 1) {
                    A.out[gi] = axpy<T, false>(A.a, k, uc);
                } else if (st < 4) {
                    const T v = axpy<T, false>(A.a, k, pc[j]);
                    A.out[gi] = v;
                    if (BF && st == 2) A.acc[gi] = rk4_b<T>(uc, v, pc[j]);
                } else {
                    A.out[gi] = BF ? rk4_fin_b<T>(A.dt6, A.third, pa[j], uc, k)
                                   : rk4_fin<T>(A.dt6, A.third, pc[j], pa[j], pq[j], uc, k);
                }
            }
        }
#pragma unroll
        for (int j = 0; j < NO; ++j) {
            pc[j] = nc[j];
            pa[j] = na[j];
            pq[j] = nq[j];
        }
    };
    int o = z0;
    for (; o + 1 < z1; o += 2) {
        iteration(o, pa0);
        iteration(o + 1, pa1);
    }
    if (o < z1) iteration(o, pa0);
}

// ------------------------------------------ 3D, z-neighbours in registers --
// One-stage kernel in the 2.5D form. A workgroup of NT threads owns TX x TY
// columns of a z-chunk and streams along z like k_ch_stage_3d, but each
// thread keeps the z-neighbours its cells need in registers: the stage state
// at planes o, o+1, o+2 and mu at planes o-1, o, o+1 of its own columns. LDS
// holds one plane of the stage state (2-cell ring) and one plane of mu
// (1-cell ring), from which only in-plane neighbours are read. That is
// ~10 KB per 1024 cells instead of ~36 KB, so tiles 128-256 cells wide keep
// their occupancy; their rows are long contiguous HBM bursts, which stream
// faster on MI355X than 64-cell row pieces. Every value is formed from the
// same operands by the same operations as in k_ch_stage_3d, so the results
// are equal bit for bit.
//
// Slots of a thread (index q = tid + NT * i within each class):
//   interior: q < TX*TY              output cells (mu kept for its z-stencil)
//   ring:     q < 2*(TX+2) + 2*TY    the 1-cell ring of the mu plane
//   outer:    q < 2*(TX+4) + 2*(TY+2) the 2-cell ring of the state plane
// Interior and ring slots (the mu slots) keep the state at three planes in
// registers; outer slots only feed the state plane in LDS. Ghost mu slots
// follow the rules of k_ch_stage_3d: FIXED takes the constant, NOFLUX and
// PERIODIC evaluate the stencil at the LDS centre of their BC image, whose
// z-neighbours equal the slot's own (the column loads apply the x/y maps).
//
// Iteration o (output plane o), barriers B1, B2:
//   state plane o+1 -> LDS                                          | B1
//   mu(o+1) on every mu slot; output plane o from LDS mu(o) in-plane and the
//   registers mu(o-1), mu(o+1)                                      | B2
//   mu(o+1) -> LDS
// The state plane written in iteration o was last read by the mu pass of
// iteration o-1 (before its B2); mu(o+1) is written after B2 of iteration o,
// which follows the last reads of mu(o), and is read after B1 of o+1.

// FULL: every tile lies inside the domain (n0 % TX == 0, n1 % TY == 0). A
// thread's tile cells are then never ghosts and sit RS = NT / TX rows apart,
// so their LDS indices and global offsets are one base plus constants
// (immediate offsets) instead of per-slot tables: ~25 fewer VGPRs.
// SC: 1, 2 (stages 2 and 3, which share their epilogue form), 3 (stage 2
// of the FP64 b form, which also writes b) and 4 fix the stage class at
// compile time, so the stage tests and the unused epilogue loads fold away
// (0 would read it from A.stage).
template <typename T, int TX, int TY, int NT, bool EXACT, bool FULL = false, int SC = 0>
__global__ __attribute__((amdgpu_flat_work_group_size(1, NT))) void k_ch_stage_3dz(CHArgs<T> A) {
    constexpr int UX = TX + 4, UY = TY + 4, MX = TX + 2, MY = TY + 2;
    constexpr int NI = TX * TY, NRING = 2 * MX + 2 * TY, NOUT = 2 * UX + 2 * MY;
    static_assert(NI % NT == 0, "tile must be a multiple of the workgroup size");
    constexpr int NO = NI / NT;
    constexpr int NR = (NRING + NT - 1) / NT;
    constexpr int NQ = (NOUT + NT - 1) / NT;
    constexpr int NS = NO + NR;
    __shared__ T SU[UY * UX];
    __shared__ T SM[MY * MX];
    __shared__ T BCV[6];
    const int tid = threadIdx.x;
    stage_bc_values(BCV, A);
    int bx = blockIdx.x, by = blockIdx.y, bz = blockIdx.z;
    if (A.xcd_remap) {
        const int gx = gridDim.x, gy = gridDim.y;
        const int nb = gx * gy * gridDim.z;
        const int h = bx + gx * (by + gy * bz);
        const int lb = (h & 7) * (nb >> 3) + (h >> 3);
        bx = lb % gx;
        by = (lb / gx) % gy;
        bz = lb / (gx * gy);
    }
    const int x0 = bx * TX, y0 = by * TY;
    const int z0 = (int)(((long long)bz * A.L.n2) / gridDim.z);
    const int z1 = (int)(((long long)(bz + 1) * A.L.n2) / gridDim.z);
    const int g = A.L.g;
    const long long sxy = A.L.sxy;
    const int cell0 = g * A.L.sx + g;

    // source offset of a state position (x/y maps applied; FIXED -> flag)
    auto src_of = [&](int px, int py, int& off, bool& fix, T& fv) {
        const int qy = map_in(py, A.by), qx = map_in(px, A.bx);
        off = cell0;
        fix = false;
        fv = (T)0;
        if (qy < 0) {
            fix = true;
            fv = fixed_value(py, A.by);
        } else if (qx < 0) {
            fix = true;
            fv = fixed_value(px, A.bx);
        } else {
            off = (qy + g) * A.L.sx + (qx + g);
        }
    };
    // mu slots: interior then ring. Column offsets are unsigned byte
    // offsets within a plane: a uniform plane base plus a 32-bit offset
    // addresses without per-lane 64-bit arithmetic.
    static_assert(NT % TX == 0, "a slot step must be whole rows");
    constexpr int RS = NT / TX;   // rows between a thread's consecutive tile cells
    const int tox = tid % TX, toy = tid / TX;
    const int sl0 = (toy + 2) * UX + tox + 2, mi0 = (toy + 1) * MX + tox + 1;
    const unsigned off0 = (unsigned)((y0 + toy + g) * A.L.sx + (x0 + tox + g)) * (unsigned)sizeof(T);
    const long long rowb = (long long)RS * A.L.sx * (long long)sizeof(T);
    unsigned soff[NS];
    int sl[NS], mi[NS], mc[NS], mcode[NS];
    T sfv[NS];
    unsigned sfm = 0, sact = 0;
#pragma unroll
    for (int s = 0; s < NS; ++s) {
        if (FULL && s < NO) continue;   // affine: SLOT_* below
        int mx, my;
        bool act = true;
        if (s < NO) {
            const int q = tid + NT * s;
            my = q / TX + 1;
            mx = q - (my - 1) * TX + 1;
        } else {
            int q = tid + NT * (s - NO);
            act = q < NRING;
            q = min(q, NRING - 1);
            if (q < MX) { mx = q; my = 0; }
            else if (q < 2 * MX) { mx = q - MX; my = MY - 1; }
            else if (q < 2 * MX + TY) { mx = 0; my = 1 + q - 2 * MX; }
            else { mx = MX - 1; my = 1 + q - 2 * MX - TY; }
        }
        const int px = x0 - 1 + mx, py = y0 - 1 + my;
        bool fix;
        int off;
        src_of(px, py, off, fix, sfv[s]);
        soff[s] = (unsigned)off * (unsigned)sizeof(T);
        if (fix) sfm |= 1u << s;
        if (act) sact |= 1u << s;
        sl[s] = (my + 1) * UX + (mx + 1);
        mi[s] = my * MX + mx;
        // mu image centre: y decides before x (as in k_ch_stage_3d)
        int ix = px, iy = py, code = 0;
        if (py < 0 || py >= A.L.n1) {
            const int k = py < 0 ? A.by.klo : A.by.khi;
            if (k == PHX_BC_FIXED) code = GC_FIX | GC_FY | (py < 0 ? 0 : GC_HI);
            else if (k == PHX_BC_NOFLUX) iy = py < 0 ? -1 - py : 2 * A.L.n1 - 1 - py;
        }
        if (!code && (px < 0 || px >= A.L.n0)) {
            const int k = px < 0 ? A.bx.klo : A.bx.khi;
            if (k == PHX_BC_FIXED) code = GC_FIX | (px < 0 ? 0 : GC_HI);
            else if (k == PHX_BC_NOFLUX) ix = px < 0 ? -1 - px : 2 * A.L.n0 - 1 - px;
        }
        // positions past the first ghost layer (partial tiles) feed no
        // output; keep their stencil inside the state plane
        ix = min(max(ix, x0 - 1), x0 + TX);
        iy = min(max(iy, y0 - 1), y0 + TY);
        mc[s] = (iy - y0 + 2) * UX + (ix - x0 + 2);
        mcode[s] = code;
    }
    // per-slot properties (s is a compile-time constant at every use)
    auto s_full = [](int s) { return FULL && s < NO; };
    auto SL = [&](int s) { return s_full(s) ? sl0 + s * RS * UX : sl[s]; };
    auto MI = [&](int s) { return s_full(s) ? mi0 + s * RS * MX : mi[s]; };
    auto MC = [&](int s) { return s_full(s) ? sl0 + s * RS * UX : mc[s]; };
    auto MCODE = [&](int s) { return s_full(s) ? 0 : mcode[s]; };
    auto FIXS = [&](int s) { return !s_full(s) && ((sfm >> s) & 1u); };
    auto ACTS = [&](int s) { return s_full(s) || ((sact >> s) & 1u); };
    // outer slots
    unsigned qoff[NQ];
    int ql[NQ];
    T qfv[NQ];
    unsigned qfm = 0, qact = 0;
#pragma unroll
    for (int k = 0; k < NQ; ++k) {
        int q = tid + NT * k;
        if (q < NOUT) qact |= 1u << k;
        q = min(q, NOUT - 1);
        int ux, uy;
        if (q < UX) { ux = q; uy = 0; }
        else if (q < 2 * UX) { ux = q - UX; uy = UY - 1; }
        else if (q < 2 * UX + MY) { ux = 0; uy = 1 + q - 2 * UX; }
        else { ux = UX - 1; uy = 1 + q - 2 * UX - MY; }
        bool fix;
        int off;
        src_of(x0 - 2 + ux, y0 - 2 + uy, off, fix, qfv[k]);
        qoff[k] = (unsigned)off * (unsigned)sizeof(T);
        if (fix) qfm |= 1u << k;
        ql[k] = uy * UX + ux;
    }
    // output cells (cells beyond the domain read a valid cell, never store)
    unsigned ooff[NO];
    bool oval[NO];
#pragma unroll
    for (int j = 0; j < NO; ++j) {
        if (FULL) {
            oval[j] = true;
            continue;
        }
        const int q = tid + NT * j;
        const int oy = q / TX, ox = q - oy * TX;
        const int x = x0 + ox, y = y0 + oy;
        oval[j] = x < A.L.n0 && y < A.L.n1;
        ooff[j] = (unsigned)(oval[j] ? (y + g) * A.L.sx + (x + g) : cell0) * (unsigned)sizeof(T);
    }

    // column loads of plane pz: raw values plus the plane's FIXED z state
    // The empty asm re-defines the 32-bit offset at each use: the compiler
    // would otherwise hoist its zero extension out of the plane loop and add
    // 64-bit VGPR addresses per load instead of using the scalar-base form.
    auto at = [](const T* base, unsigned off) {
        asm volatile("" : "+v"(off));
        return *(const T*)((const char*)base + off);
    };
    auto at_nt = [](const T* base, unsigned off) {   // a last use: no cache allocation
        asm volatile("" : "+v"(off));
        return __builtin_nontemporal_load((const T*)((const char*)base + off));
    };
    auto zplane = [&](int pz, bool& fix, T& val) -> const T* {
        const int qz = map_in(pz, A.bz);
        fix = qz < 0;
        val = fix ? fixed_value(pz, A.bz) : (T)0;
        return A.u + (long long)((fix ? 0 : qz) + g) * sxy;
    };
    auto load_cols = [&](int pz, T (&r)[NS], bool& fix, T& val) {
        const T* base = zplane(pz, fix, val);
#pragma unroll
        for (int s = 0; s < NS; ++s) {
            if (s_full(s)) r[s] = at((const T*)((const char*)base + s * rowb), off0);
            else r[s] = at(base, soff[s]);
        }
    };
    auto load_outer = [&](int pz, T (&r)[NQ], bool& fix, T& val) {
        const T* base = zplane(pz, fix, val);
#pragma unroll
        for (int k = 0; k < NQ; ++k) r[k] = at(base, qoff[k]);
    };
    auto resolve_cols = [&](const T (&r)[NS], bool fix, T val, T (&u)[NS]) {
#pragma unroll
        for (int s = 0; s < NS; ++s) u[s] = fix ? val : (FIXS(s) ? sfv[s] : r[s]);
    };
    auto store_state = [&](const T (&u)[NS], const T (&r)[NQ], bool qfix, T qval) {
#pragma unroll
        for (int s = 0; s < NS; ++s)
            if (ACTS(s)) SU[SL(s)] = u[s];
#pragma unroll
        for (int k = 0; k < NQ; ++k)
            if ((qact >> k) & 1u) SU[ql[k]] = qfix ? qval : (((qfm >> k) & 1u) ? qfv[k] : r[k]);
    };
    // z rule of mu plane m: FIXED ghost -> constant; NOFLUX ghost -> the
    // image plane's mu (copied by the caller); else evaluate
    auto zrule = [&](int m, bool& zfix, bool& znf, T& zv) {
        zfix = znf = false;
        zv = (T)0;
        if (m < 0 || m >= A.L.n2) {
            const int k = m < 0 ? A.bz.klo : A.bz.khi;
            if (k == PHX_BC_FIXED) {
                zfix = true;
                zv = BCV[m < 0 ? 4 : 5];
            } else if (k == PHX_BC_NOFLUX) {
                znf = true;
            }
        }
    };
    auto mu_slot = [&](int s, T zp, T zm) {
        const int c0 = MC(s);
        T v = mu_at<T, EXACT, 3>(A, SU[c0], SU[c0 + 1], SU[c0 - 1], SU[c0 + UX], SU[c0 - UX], zp, zm);
        if (MCODE(s) & GC_FIX) v = ghost_value_xy(MCODE(s), BCV);
        return v;
    };
    // epilogue operands (all loaded for every stage: the same loads each
    // iteration; an operand the stage does not need is read from the first
    // element of the stage state for every lane: one cache line per wave).
    // Exact FP64 form: c_n (stages 2-4) and the accumulator (2-4); FP64 b
    // form: c_n (2-3) and b (4), its stage 2 (class 3) also writing b to
    // A.acc; FP32: c_n (2-4) and the stage states u2, u3 (4).
    constexpr bool BF = kBForm<T, EXACT>;
    const int st = SC == 3 ? 2 : (SC ? SC : A.stage);
    const bool wb = BF && (SC ? SC == 3 : A.stage == 2);
    const bool need_c = EXACT ? (st == 2 || st == 3 || st == 4) : (BF ? (st == 2 || st == 3) : st >= 2);
    const bool need_a = EXACT ? st >= 2 : st == 4;
    const bool need_b = !EXACT && !BF && st == 4;
    const T* csrc = need_c ? A.c : A.u;
    const T* asrc = need_a ? (EXACT || BF ? A.acc : A.s2) : A.u;
    const T* bsrc = need_b ? A.s3 : A.u;
    const long long cpm = need_c ? -1LL : 0LL, apm = need_a ? -1LL : 0LL, bpm = need_b ? -1LL : 0LL;
    const unsigned com = need_c ? ~0u : 0u, aom = need_a ? ~0u : 0u, bom = need_b ? ~0u : 0u;
    const long long rowe = (long long)RS * A.L.sx;   // elements between a thread's tile cells (FULL)
    auto load_epi = [&](int o, T (&rc)[NO], T (&ra)[NO], T (&rb)[NO]) {
        const long long pb = (long long)(o + g) * sxy;
#pragma unroll
        for (int j = 0; j < NO; ++j) {
            const long long pj = FULL ? pb + j * rowe : pb;
            const unsigned oj = FULL ? off0 : ooff[j];
            // b form: the last uses of c_n (stage 3, class 2) and of b (stage 4)
            // allocate no cache lines, leaving the MALL to the next stage's state
            if constexpr (BF && SC == 2) rc[j] = at_nt(csrc + (pj & cpm), oj & com);
            else rc[j] = at(csrc + (pj & cpm), oj & com);
            if constexpr (BF && SC == 4) ra[j] = at_nt(asrc + (pj & apm), oj & aom);
            else ra[j] = at(asrc + (pj & apm), oj & aom);
            if constexpr (!EXACT && !BF) rb[j] = at(bsrc + (pj & bpm), oj & bom);
        }
    };

    T u0[NS], u1[NS], u2[NS];
    T mum[NO], mu0[NO], mup[NS];
    // Loads in flight. DEPTH 2 (FP64, and FP32 tiles of at most two cells
    // per thread): two sets, one per plane parity; iteration o consumes its
    // set (state plane o+2 at the mu slots, outer plane o+1, epilogue
    // operands of plane o) and refills it for iteration o+2, so every load
    // has two iterations to land and no register copy waits for one.
    // DEPTH 1 (FP32 tiles of four cells per thread): one set refilled for
    // the next iteration, the epilogue operands staged in the second set's
    // and copied at the iteration's end -- fewer VGPRs, which these need
    // for occupancy.
    constexpr int DEPTH = (sizeof(T) == 8 || NO <= 2) ? 2 : 1;
    struct Pending {
        T upf[NS], qv[NQ], ec[NO], ea[NO], eb[NO];
        bool ufix, qfix;
        T uval, qval;
    } PA, PB;
    // ---- prologue: state planes z0-2 .. z0+1 at the mu slots, z0-1 and z0
    // on the outer ring; mu(z0-1) on the interior, mu(z0) everywhere
    {
        T r0[NS], r1[NS], r2[NS], r3[NS], o0[NQ], o1[NQ];
        bool f0, f1, f2, f3, g0, g1;
        T v0, v1, v2, v3, w0, w1;
        load_cols(z0 - 2, r0, f0, v0);
        load_cols(z0 - 1, r1, f1, v1);
        load_cols(z0, r2, f2, v2);
        load_cols(z0 + 1, r3, f3, v3);
        load_outer(z0 - 1, o0, g0, w0);
        load_outer(z0, o1, g1, w1);
        load_epi(z0, PA.ec, PA.ea, PA.eb);
        T a0[NS], a1[NS];
        resolve_cols(r0, f0, v0, a0);
        resolve_cols(r1, f1, v1, a1);
        resolve_cols(r2, f2, v2, u0);
        resolve_cols(r3, f3, v3, u1);
        store_state(a1, o0, g0, w0);
        __syncthreads();
        bool zfix, znf;
        T zv;
        zrule(z0 - 1, zfix, znf, zv);
#pragma unroll
        for (int j = 0; j < NO; ++j) mum[j] = zfix ? zv : mu_slot(j, u0[j], a0[j]);
        __syncthreads();
        store_state(u0, o1, g1, w1);
        load_cols(min(z0 + 2, z1 + 1), PA.upf, PA.ufix, PA.uval);
        load_outer(min(z0 + 1, z1), PA.qv, PA.qfix, PA.qval);
        if constexpr (DEPTH == 2) {
            load_cols(min(z0 + 3, z1 + 1), PB.upf, PB.ufix, PB.uval);
            load_outer(min(z0 + 2, z1), PB.qv, PB.qfix, PB.qval);
            load_epi(min(z0 + 1, z1 - 1), PB.ec, PB.ea, PB.eb);
        }
        __syncthreads();
#pragma unroll
        for (int s = 0; s < NS; ++s) {
            const T v = mu_slot(s, u1[s], a1[s]);
            if (ACTS(s)) SM[MI(s)] = v;
            if (s < NO) mu0[s] = v;
        }
        // NOFLUX: mu of the ghost plane z0-1 = -1 is mu of plane 0
        if (znf) {
#pragma unroll
            for (int j = 0; j < NO; ++j) mum[j] = mu0[j];
        }
        __syncthreads();
    }

    auto iteration = [&](int o, Pending& P, Pending& N) {
        resolve_cols(P.upf, P.ufix, P.uval, u2);
        store_state(u1, P.qv, P.qfix, P.qval);
        // unconditional prefetch (clamped planes): the same loads every iteration
        load_cols(min(o + 2 + DEPTH, z1 + 1), P.upf, P.ufix, P.uval);
        load_outer(min(o + 1 + DEPTH, z1), P.qv, P.qfix, P.qval);
        if constexpr (DEPTH == 1) load_epi(min(o + 1, z1 - 1), N.ec, N.ea, N.eb);
        __syncthreads();   // B1
        bool zfix, znf;
        T zv;
        zrule(o + 1, zfix, znf, zv);
#pragma unroll
        for (int s = 0; s < NS; ++s) {
            T v = mu_slot(s, u2[s], u0[s]);
            if (s < NO && znf) v = mu0[s];
            mup[s] = zfix ? zv : v;
        }
        const long long pb = (long long)(o + g) * sxy;
        T* const accp = A.acc + pb;
        T* const outp = A.out + pb;
#pragma unroll
        for (int j = 0; j < NO; ++j) {
            if (!oval[j]) continue;
            const int m0 = MI(j);
            const T k = k_at<T, EXACT, 3>(A, SM[m0], SM[m0 + 1], SM[m0 - 1], SM[m0 + MX], SM[m0 - MX], mup[j], mum[j]);
            unsigned ob = FULL ? off0 : ooff[j];
            asm volatile("" : "+v"(ob));
            T& acc = *(T*)((char*)(FULL ? accp + j * rowe : accp) + ob);
            T& out = *(T*)((char*)(FULL ? outp + j * rowe : outp) + ob);
            const T uc = u0[j];
            if (EXACT) {
                if (st == 0) {   // an explicit Euler step (multi-step Euler)
                    out = uc + A.a * k;
                } else if (st == 1) {
                    acc = k;
                    out = uc + A.a * k;
                } else if (st < 4) {
                    acc = P.ea[j] + (T)2 * k;
                    out = P.ec[j] + A.a * k;
                } else {
                    out = P.ec[j] + A.dt6 * (P.ea[j] + k);
                }
            } else {
                if (st == 1) {
                    out = axpy<T, false>(A.a, k, uc);
                } else if (st < 4) {
                    const T v = axpy<T, false>(A.a, k, P.ec[j]);
                    out = v;
                    // b is read two launches later: streamed past the caches, so
                    // that u3 stays in the MALL for stage 3
                    if (wb) __builtin_nontemporal_store(rk4_b<T>(uc, v, P.ec[j]), &acc);
                } else {
                    out = BF ? rk4_fin_b<T>(A.dt6, A.third, P.ea[j], uc, k)
                             : rk4_fin<T>(A.dt6, A.third, P.ec[j], P.ea[j], P.eb[j], uc, k);
                }
            }
        }
        if constexpr (DEPTH == 2) load_epi(min(o + 2, z1 - 1), P.ec, P.ea, P.eb);
        __syncthreads();   // B2
#pragma unroll
        for (int s = 0; s < NS; ++s)
            if (ACTS(s)) SM[MI(s)] = mup[s];
#pragma unroll
        for (int s = 0; s < NS; ++s) {
            u0[s] = u1[s];
            u1[s] = u2[s];
        }
#pragma unroll
        for (int j = 0; j < NO; ++j) {
            mum[j] = mu0[j];
            mu0[j] = mup[j];
            if constexpr (DEPTH == 1) {
                P.ec[j] = N.ec[j];
                P.ea[j] = N.ea[j];
                P.eb[j] = N.eb[j];
            }
        }
    };
    int o = z0;
    if constexpr (DEPTH == 2) {
        for (; o + 1 < z1; o += 2) {
            iteration(o, PA, PA);
            iteration(o + 1, PB, PB);
        }
        if (o < z1) iteration(o, PA, PA);
    } else {
        for (; o < z1; ++o) iteration(o, PA, PB);
    }
}

// ------------------------------------------------- 3D, two stages fused --
// Temporal blocking: one launch evaluates two consecutive RK4 stages, (1,2) or
// (3,4), so a step costs 7 field passes over HBM instead of 16 (17 in FP64):
//   (1,2): read c_n;                write acc and the stage-3 state;
//   (3,4): read stage-3 state, c_n, acc;  write c_{n+1} (to a buffer other than
//          c_n: the (3,4) launch reads c_n in a halo around its tile).
// Per z plane the block computes, in LDS, five levels over shrinking regions
// (TX+8)x(TY+8) -> ... -> TX x TY:
//   UA = first stage's input state, MA = its mu, UB = second stage's input
//   state (c_n + a1*k_first), MB = its mu, and the output cells.
// The stage-internal ghost values are the same as in the one-stage kernels: a
// level's value at a ghost position is its value at the BC image position
// (mirror for NOFLUX with the +/- neighbours of the ghost axis exchanged,
// wrap for PERIODIC) or the FIXED constant, decided z, then y, then x. The
// image must lie in the domain for ghost depths up to 4, hence n >= 4 per axis.
// Every value is formed by the operations, in the order, of the one-stage
// kernels, so results match them bit for bit in FP32 and FP64.
//
// Pipeline, iteration o (output plane o), barriers B1..B4:
//   P1 store UA(o+4) [regs loaded one iteration earlier]         | B1
//   P2 MA(o+3) from UA(o+2..o+4)                                  | B2
//   P3 UB(o+2) from MA(o+1..o+3); keep k and c of the tile cells  | B3
//   P4 MB(o+1) from UB(o..o+2)                                    | B4
//   P5 output plane o from MB(o-1..o+1)
// Every level keeps 3 planes. The slot written in a phase held the plane
// whose last reader ran before an earlier barrier: UA(o+1) last read in P3 of
// iteration o-1, MA(o) in P3(o-1), UB(o-1) in P4(o-1), MB(o-2) and the tile
// ring's plane o-1 in P5(o-1), which B1(o) separates from P3/P4(o).
template <typename T>
struct PairArgs {
    const T* u;    // first stage's input state: c_n for (1,2), the stage-3 state for (3,4)
    const T* c;    // c_n (read by (3,4))
    T* acc;        // RK4 accumulator (exact FP64 form), b = u2 + 2 u3 - c_n
                   // (FP64) or the stage-2 state u2 (FP32): written by (1,2),
                   // read by (3,4)
    T* out;        // (1,2): stage-3 state; (3,4): c_{n+1}
    Layout L;
    AxisBC<T> bx, by, bz;
    T idx2, idy2, idz2;
    T two_rho, ca, cb, nkappa, M;
    T a1, a2;      // stage shifts: (dt/2, dt/2) for (1,2); (dt, unused) for (3,4)
    T dt6, dt3;
    T third;       // 1/3 (FP32 RK4 update)
    FastCoef<T> F; // FP32 arithmetic form
    int zc;
    int xcd_remap;
    long long bstride = 0;   // 2D whole step: elements between the fields of a batch (blockIdx.z)
};

// BC image of position p for ghost depths up to D (needs n >= D); -1: FIXED.
// Positions beyond depth D only occur in partial tiles and feed no output;
// they are clamped to keep every load in bounds.
template <typename T, int D = 4>
__device__ __forceinline__ int map_deep(int p, const AxisBC<T>& b) {
    p = p < -D ? -D : (p > b.n + D - 1 ? b.n + D - 1 : p);
    if (p >= 0 && p < b.n) return p;
    if (p < 0) {
        if (b.klo == PHX_BC_PERIODIC) return p + b.n;
        if (b.klo == PHX_BC_FIXED) return -1;
        return -1 - p;
    }
    if (b.khi == PHX_BC_PERIODIC) return p - b.n;
    if (b.khi == PHX_BC_FIXED) return -1;
    return 2 * b.n - 1 - p;
}

template <typename T>
__device__ __forceinline__ int ghost_code_xy(int px, int py, const PairArgs<T>& A) {
    int code = 0;
    if (py < 0 || py >= A.L.n1) {
        const int k = py < 0 ? A.by.klo : A.by.khi;
        if (k == PHX_BC_FIXED) return GC_FIX | GC_FY | (py < 0 ? 0 : GC_HI);
        if (k == PHX_BC_NOFLUX) code |= GC_SWY;
    }
    if (px < 0 || px >= A.L.n0) {
        const int k = px < 0 ? A.bx.klo : A.bx.khi;
        if (k == PHX_BC_FIXED) return GC_FIX | (px < 0 ? 0 : GC_HI);
        if (k == PHX_BC_NOFLUX) code |= GC_SWX;
    }
    return code;
}


// Seven-point star around LDS index i of a level with row width W, with the
// neighbours of NOFLUX ghost axes exchanged: s = {c, xp, xm, yp, ym, zp, zm}.
template <typename T, int W, bool GEN>
__device__ __forceinline__ void star(const T* lo, const T* mid, const T* hi, int i, int code, bool zsw, T (&s)[7]) {
    s[0] = mid[i];
    s[1] = mid[i + 1];
    s[2] = mid[i - 1];
    s[3] = mid[i + W];
    s[4] = mid[i - W];
    s[5] = hi[i];
    s[6] = lo[i];
    if (GEN) {
        if (code & GC_SWX) { const T t = s[1]; s[1] = s[2]; s[2] = t; }
        if (code & GC_SWY) { const T t = s[3]; s[3] = s[4]; s[4] = t; }
        if (zsw) { const T t = s[5]; s[5] = s[6]; s[6] = t; }
    }
}

template <typename T, bool EXACT>
__device__ __forceinline__ T mu_star(const PairArgs<T>& A, const T (&s)[7]) {
    return mu_at<T, EXACT, 3>(A, s[0], s[1], s[2], s[3], s[4], s[5], s[6]);
}
template <typename T, bool EXACT>
__device__ __forceinline__ T k_star(const PairArgs<T>& A, const T (&s)[7]) {
    return k_at<T, EXACT, 3>(A, s[0], s[1], s[2], s[3], s[4], s[5], s[6]);
}

template <typename T, int TX, int TY, bool EXACT, bool SECOND>
__global__ __launch_bounds__(256) void k_ch_pair_3d(PairArgs<T> A) {
    constexpr int W4 = TX + 8, N4 = W4 * (TY + 8);
    constexpr int W3 = TX + 6, N3 = W3 * (TY + 6);
    constexpr int W2 = TX + 4, N2 = W2 * (TY + 4);
    constexpr int W1 = TX + 2, N1 = W1 * (TY + 2);
    constexpr int N0 = TX * TY;
    constexpr int S4 = (N4 + 255) / 256, S3 = (N3 + 255) / 256, S2 = (N2 + 255) / 256, S1 = (N1 + 255) / 256;
    constexpr int S0 = N0 / 256;
    static_assert(N0 % 256 == 0, "tile must be a multiple of 256 cells");
    constexpr bool NEED_CR = true;   // c_n of the tile cells, read by both epilogues
    constexpr bool BF = kBForm<T, EXACT>;   // FP64: (1,2) writes b to acc, (3,4) finishes from b
    __shared__ T UA[3][N4];
    __shared__ T MA[3][N3];
    __shared__ T UB[3][N2];
    __shared__ T MB[3][N1];
    __shared__ T KR[3][N0];
    __shared__ T CR[NEED_CR ? 3 : 1][NEED_CR ? N0 : 1];
    __shared__ T BCV[6];
    const int tid = threadIdx.x;
    stage_bc_values(BCV, A);
    const AxisBC<T> bzu = uni_axis(A.bz);
    const int n2 = uni(A.L.n2);
    int bx = blockIdx.x, by = blockIdx.y, bz = blockIdx.z;
    if (A.xcd_remap) {
        const int gx = gridDim.x, gy = gridDim.y;
        const int nb = gx * gy * gridDim.z;
        const int h = bx + gx * (by + gy * bz);
        const int lb = (h & 7) * (nb >> 3) + (h >> 3);
        bx = lb % gx;
        by = (lb / gx) % gy;
        bz = lb / (gx * gy);
    }
    const int x0 = bx * TX, y0 = by * TY;
    // balanced z chunks: gridDim.z chunks of floor or ceil(n2 / gridDim.z) planes
    const int z0 = (int)(((long long)bz * n2) / gridDim.z);
    const int z1 = (int)(((long long)(bz + 1) * n2) / gridDim.z);
    const int g = A.L.g;
    const long long sxy = A.L.sxy;
    const bool xy_in = x0 >= 4 && x0 + TX + 4 <= A.L.n0 && y0 >= 4 && y0 + TY + 4 <= A.L.n1;
    auto rs = [](int p) { return (p + 12) % 3; };   // ring slot, planes >= -4

    // UA halo slots: in-plane offset with the x/y images. Every slot loads
    // unconditionally from a valid cell (FIXED and surplus slots read the
    // plane's first cell) so loads never sit in divergent branches and stay in
    // flight across the pipeline; FIXED constants replace them at the store.
    const int cell0 = g * A.L.sx + g;
    int hoff[S4];
    T hfix[S4];
    unsigned hfm = 0;   // bit i: slot i holds a FIXED constant
#pragma unroll
    for (int i = 0; i < S4; ++i) {
        const int t = tid + 256 * i;
        hoff[i] = cell0;
        hfix[i] = (T)0;
        if (t < N4) {
            const int ly = t / W4, lx = t - ly * W4;
            const int px = x0 - 4 + lx, py = y0 - 4 + ly;
            const int qy = map_deep(py, A.by), qx = map_deep(px, A.bx);
            if (qy < 0) {
                hfm |= 1u << i;
                hfix[i] = fixed_value(py, A.by);
            } else if (qx < 0) {
                hfm |= 1u << i;
                hfix[i] = fixed_value(px, A.bx);
            } else {
                hoff[i] = (qy + g) * A.L.sx + (qx + g);
            }
        }
    }
    // in-plane ghost codes of the level slots (all zero for interior tiles)
    int code3[S3], code2[S2], code1[S1];
#pragma unroll
    for (int i = 0; i < S3; ++i) {
        const int t = tid + 256 * i, ly = t / W3, lx = t - ly * W3;
        code3[i] = xy_in ? 0 : ghost_code_xy(x0 - 3 + lx, y0 - 3 + ly, A);
    }
#pragma unroll
    for (int i = 0; i < S2; ++i) {
        const int t = tid + 256 * i, ly = t / W2, lx = t - ly * W2;
        code2[i] = xy_in ? 0 : ghost_code_xy(x0 - 2 + lx, y0 - 2 + ly, A);
    }
#pragma unroll
    for (int i = 0; i < S1; ++i) {
        const int t = tid + 256 * i, ly = t / W1, lx = t - ly * W1;
        code1[i] = xy_in ? 0 : ghost_code_xy(x0 - 1 + lx, y0 - 1 + ly, A);
    }
    // c_n at the UB slots ((3,4) only): in-plane offsets with the images
    int coff[SECOND ? S2 : 1];
    T creg[SECOND ? S2 : 1];
    if (SECOND) {
#pragma unroll
        for (int i = 0; i < S2; ++i) {
            const int t = tid + 256 * i, ly = t / W2, lx = t - ly * W2;
            const int qy = map_deep(y0 - 2 + ly, A.by), qx = map_deep(x0 - 2 + lx, A.bx);
            coff[i] = (t < N2 && qy >= 0 && qx >= 0) ? (qy + g) * A.L.sx + (qx + g) : cell0;
            creg[i] = (T)0;
        }
    }
    // output slots
    int ooff[S0];
    bool oval[S0];
    T areg[SECOND ? S0 : 1];
    T breg[(SECOND && !EXACT && !BF) ? S0 : 1];   // FP32 (3,4): the stage-3 state at the tile cells
#pragma unroll
    for (int j = 0; j < S0; ++j) {
        const int t = tid + 256 * j;
        const int oy = t / TX, ox = t - oy * TX;
        const int x = x0 + ox, y = y0 + oy;
        oval[j] = x < A.L.n0 && y < A.L.n1;
        ooff[j] = oval[j] ? (y + g) * A.L.sx + (x + g) : cell0;
    }

    // a UA plane in flight: its registers, and its FIXED z constant if any
    struct Plane {
        T r[S4];
        bool fix;
        T val;
    };
    auto load_ua = [&](int pz, Plane& P) {
        const int qz = map_deep(pz, bzu);
        P.fix = qz < 0;
        P.val = P.fix ? fixed_value(pz, bzu) : (T)0;
        const T* base = A.u + (long long)((P.fix ? 0 : qz) + g) * sxy;
#pragma unroll
        for (int i = 0; i < S4; ++i) P.r[i] = base[hoff[i]];
    };
    auto store_ua = [&](int pz, const Plane& P) {
        T* dst = UA[rs(pz)];
#pragma unroll
        for (int i = 0; i < S4; ++i)
            if (tid + 256 * i < N4) dst[tid + 256 * i] = P.fix ? P.val : (((hfm >> i) & 1u) ? hfix[i] : P.r[i]);
    };
    // c_n of a plane at the UB slots; on a FIXED z plane UB is the constant
    // and the values are unused
    auto load_c = [&](int pz) {
        const int qz = map_deep(pz, bzu);
        const T* base = A.c + (long long)((qz < 0 ? 0 : qz) + g) * sxy;
#pragma unroll
        for (int i = 0; i < (SECOND ? S2 : 1); ++i) creg[i] = base[coff[i]];
    };
    auto load_acc = [&](int o) {
        const long long pb = (long long)(o + g) * sxy;
#pragma unroll
        for (int j = 0; j < (SECOND ? S0 : 1); ++j) areg[j] = A.acc[pb + ooff[j]];
        if constexpr (SECOND && !EXACT && !BF) {
#pragma unroll
            for (int j = 0; j < S0; ++j) breg[j] = A.u[pb + ooff[j]];
        }
    };
    // ghost rule of plane pz along z (decides before y and x)
    auto zrule = [&](int pz, bool& zfix, bool& zsw, T& zv) {
        zfix = false;
        zsw = false;
        zv = (T)0;
        if (pz < 0 || pz >= n2) {
            const int k = pz < 0 ? bzu.klo : bzu.khi;
            if (k == PHX_BC_FIXED) {
                zfix = true;
                zv = fixed_value(pz, bzu);
            } else if (k == PHX_BC_NOFLUX) {
                zsw = true;
            }
        }
    };

    // P2: mu of the first stage on plane m
    auto eval_ma = [&](int m) {
        const T* lo = UA[rs(m - 1)];
        const T* mid = UA[rs(m)];
        const T* hi = UA[rs(m + 1)];
        T* dst = MA[rs(m)];
        bool zfix, zsw;
        T zv;
        zrule(m, zfix, zsw, zv);
        auto body = [&](auto gen) {
            constexpr bool GEN = decltype(gen)::value;
#pragma unroll
            for (int i = 0; i < S3; ++i) {
                const int t = tid + 256 * i;
                if (t < N3) {
                    const int ly = t / W3, lx = t - ly * W3;
                    const int ci = (ly + 1) * W4 + lx + 1;
                    T v;
                    if (GEN && zfix) v = zv;
                    else if (GEN && (code3[i] & GC_FIX)) v = ghost_value_xy(code3[i], BCV);
                    else {
                        T s7[7];
                        star<T, W4, GEN>(lo, mid, hi, ci, code3[i], zsw, s7);
                        v = mu_star<T, EXACT>(A, s7);
                    }
                    dst[t] = v;
                }
            }
        };
        if (xy_in && !zfix && !zsw) body(std::false_type{});
        else body(std::true_type{});
    };
    // P3: second stage's input state on plane b, and k / c_n of the tile cells
    auto eval_ub = [&](int b) {
        const T* lo = MA[rs(b - 1)];
        const T* mid = MA[rs(b)];
        const T* hi = MA[rs(b + 1)];
        const T* uc = UA[rs(b)];
        T* dst = UB[rs(b)];
        T* kr = KR[rs(b)];
        T* cr = CR[NEED_CR ? rs(b) : 0];
        bool zfix, zsw;
        T zv;
        zrule(b, zfix, zsw, zv);
        auto body = [&](auto gen) {
            constexpr bool GEN = decltype(gen)::value;
#pragma unroll
            for (int i = 0; i < S2; ++i) {
                const int t = tid + 256 * i;
                if (t < N2) {
                    const int ly = t / W2, lx = t - ly * W2;
                    const int mi = (ly + 1) * W3 + lx + 1;
                    T cc;
                    if constexpr (SECOND) cc = creg[i];
                    else cc = uc[(ly + 2) * W4 + lx + 2];
                    T v, k = (T)0;
                    if (GEN && zfix) v = zv;
                    else if (GEN && (code2[i] & GC_FIX)) v = ghost_value_xy(code2[i], BCV);
                    else {
                        T s7[7];
                        star<T, W3, GEN>(lo, mid, hi, mi, code2[i], zsw, s7);
                        k = k_star<T, EXACT>(A, s7);
                        v = axpy<T, EXACT>(A.a1, k, cc);
                    }
                    dst[t] = v;
                    const int ox = lx - 2, oy = ly - 2;
                    if (ox >= 0 && ox < TX && oy >= 0 && oy < TY) {
                        // FP64 keeps k; FP32 the first stage's state (u2 or u4)
                        kr[oy * TX + ox] = EXACT ? k : v;
                        if constexpr (NEED_CR) cr[oy * TX + ox] = cc;
                    }
                }
            }
        };
        if (xy_in && !zfix && !zsw) body(std::false_type{});
        else body(std::true_type{});
    };
    // P4: mu of the second stage on plane m
    auto eval_mb = [&](int m) {
        const T* lo = UB[rs(m - 1)];
        const T* mid = UB[rs(m)];
        const T* hi = UB[rs(m + 1)];
        T* dst = MB[rs(m)];
        bool zfix, zsw;
        T zv;
        zrule(m, zfix, zsw, zv);
        auto body = [&](auto gen) {
            constexpr bool GEN = decltype(gen)::value;
#pragma unroll
            for (int i = 0; i < S1; ++i) {
                const int t = tid + 256 * i;
                if (t < N1) {
                    const int ly = t / W1, lx = t - ly * W1;
                    const int ci = (ly + 1) * W2 + lx + 1;
                    T v;
                    if (GEN && zfix) v = zv;
                    else if (GEN && (code1[i] & GC_FIX)) v = ghost_value_xy(code1[i], BCV);
                    else {
                        T s7[7];
                        star<T, W2, GEN>(lo, mid, hi, ci, code1[i], zsw, s7);
                        v = mu_star<T, EXACT>(A, s7);
                    }
                    dst[t] = v;
                }
            }
        };
        if (xy_in && !zfix && !zsw) body(std::false_type{});
        else body(std::true_type{});
    };
    // P5: second stage's k on output plane o and the stage epilogues
    auto out_plane = [&](int o) {
        const T* lo = MB[rs(o - 1)];
        const T* mid = MB[rs(o)];
        const T* hi = MB[rs(o + 1)];
        const T* kr = KR[rs(o)];
        const T* cr = CR[NEED_CR ? rs(o) : 0];
        const long long pb = (long long)(o + g) * sxy;
#pragma unroll
        for (int j = 0; j < S0; ++j) {
            if (!oval[j]) continue;
            const int t = tid + 256 * j;
            const int oy = t / TX, ox = t - oy * TX;
            const int mi = (oy + 1) * W1 + ox + 1;
            T s7[7];
            star<T, W1, false>(lo, mid, hi, mi, 0, false, s7);
            const T k = k_star<T, EXACT>(A, s7);
            const long long gi = pb + ooff[j];
            if constexpr (!SECOND) {
                // stage 2: state c + dt/2*k2; acc = k1 + 2*k2 (exact form), b (FP64
                // b form) or u2 (FP32)
                const T v = axpy<T, EXACT>(A.a2, k, cr[t]);
                if (EXACT) A.acc[gi] = kr[t] + (T)2 * k;
                else if (BF) A.acc[gi] = rk4_b<T>(kr[t], v, cr[t]);
                else A.acc[gi] = kr[t];
                A.out[gi] = v;
            } else {
                // stage 3 accumulates k3 (exact form) or forms u4, stage 4 finishes the step
                if (EXACT) {
                    const T ac = areg[j] + (T)2 * kr[t];
                    A.out[gi] = cr[t] + A.dt6 * (ac + k);
                } else if (BF) {
                    A.out[gi] = rk4_fin_b<T>(A.dt6, A.third, areg[j], kr[t], k);
                } else {
                    A.out[gi] = rk4_fin<T>(A.dt6, A.third, cr[t], areg[j], breg[j], kr[t], k);
                }
            }
        }
    };

    // prologue: UA planes z0-4, z0-3 stored, z0-2 in flight
    Plane pre;
    Plane p1;
    load_ua(z0 - 4, pre);
    load_ua(z0 - 3, p1);
    store_ua(z0 - 4, pre);
    store_ua(z0 - 3, p1);
    load_ua(z0 - 2, pre);
    if constexpr (SECOND) {
        // same load sequence as an iteration's (values unused: the loop
        // reloads both before their first use), so the loop header is
        // entered with one wait state
        load_c(z0 - 3);
        load_acc(z0);
    }
    // Every iteration issues the same loads (planes clamped to the chunk's
    // range), so the compiler waits for exactly the ones it uses and the
    // younger prefetches stay in flight: UA(o+5) after B1, c_n(o+3) after P3,
    // acc(o+1) after P5, each consumed one iteration later.
    for (int o = z0 - 6; o < z1; ++o) {
        store_ua(o + 4, pre);
        __syncthreads();   // B1
        load_ua(min(o + 5, z1 + 3), pre);
        if (o + 3 >= z0 - 3) eval_ma(o + 3);
        __syncthreads();   // B2
        if (o + 2 >= z0 - 2) eval_ub(o + 2);
        if constexpr (SECOND) load_c(min(o + 3, z1 + 1));
        __syncthreads();   // B3
        if (o + 1 >= z0 - 1) eval_mb(o + 1);
        __syncthreads();   // B4
        if (o >= z0) out_plane(o);
        if constexpr (SECOND) load_acc(max(z0, min(o + 1, z1 - 1)));
    }
}

// The eight levels of the 2D whole-step kernel after c_n is in SC. Every
// level lives in one frame: position (x, y) of the block's tile plus its
// 8-cell ring sits at index (y - y0 + 8) * P + (x - x0 + 8) in SC, SA, SB
// and SACC alike (P = TX + 16). A level covers whole frame rows -- ring R
// spans rows 8-R .. 8+TY+R-1, all P columns -- cut into runs of V = 16 /
// sizeof(T) consecutive positions; run q of a level starts at F0 + V q and
// thread tid takes runs tid + NT i. A run's operands are three 16-byte LDS
// reads (the run, the rows above and below) and the two scalars left and
// right of it; its results are one 16-byte store per array. The arithmetic
// per position is the one-stage kernels' (mu_at, k_at, axpy), so the levels
// return their bits. The frame columns outside the ring (at most 16 of P) are
// evaluated too: they read written, finite entries of the previous level
// (its rows cover one more row on each side, which also holds the scalars
// beyond a level's first and last run) and feed only other out-of-ring
// entries, never a ring or tile cell. The last partial pass reads a valid run
// and skips the store. GEN: the block needs the ghost rules of NOFLUX /
// FIXED faces (a code per position).
template <typename T, int TX, int TY, int NT, bool EXACT, bool GEN>
__device__ __forceinline__ void step2d_levels(const PairArgs<T>& A, const T* SC, T* SA, T* SB, T* SACC,
                                              const T* BCV, int x0, int y0, long long zb) {
    constexpr int P = TX + 16;
    constexpr int V = 16 / (int)sizeof(T);
    typedef T VT __attribute__((ext_vector_type(V)));
    static_assert(P % V == 0, "frame rows hold whole runs");
    static_assert(NT % TX == 0, "tile rows must divide the workgroup");
    const int tid = threadIdx.x;
    // the run at f and its neighbours: xp/xm along x, yp/ym along y, with the
    // ghost swaps and the FIXED code of each position
    auto operands = [&](const T* src, int f, T (&c)[V], T (&xp)[V], T (&xm)[V], T (&yp)[V], T (&ym)[V],
                        int (&code)[V]) {
        const VT cv = *reinterpret_cast<const VT*>(&src[f]);
        const VT uv = *reinterpret_cast<const VT*>(&src[f + P]);
        const VT dv = *reinterpret_cast<const VT*>(&src[f - P]);
        const T l = src[f - 1], r = src[f + V];
#pragma unroll
        for (int q = 0; q < V; ++q) {
            c[q] = cv[q];
            xp[q] = q + 1 < V ? cv[q + 1] : r;
            xm[q] = q > 0 ? cv[q - 1] : l;
            yp[q] = uv[q];
            ym[q] = dv[q];
            code[q] = GEN ? ghost_code_xy(x0 - 8 + (f + q) % P, y0 - 8 + (f + q) / P, A) : 0;
            if (GEN) {
                if (code[q] & GC_SWX) { const T t = xp[q]; xp[q] = xm[q]; xm[q] = t; }
                if (code[q] & GC_SWY) { const T t = yp[q]; yp[q] = ym[q]; ym[q] = t; }
            }
        }
    };
    // mu on ring R from the state level on ring R+1
    auto eval_mu = [&](auto rr, const T* src) {
        constexpr int R = decltype(rr)::value;
        constexpr int F0 = (8 - R) * P, NQ = (TY + 2 * R) * P / V;
        constexpr int NIT = (NQ + NT - 1) / NT;
#pragma unroll
        for (int i = 0; i < NIT; ++i) {
            const int q = tid + NT * i;
            const int f = F0 + V * (NQ % NT != 0 && i == NIT - 1 ? min(q, NQ - 1) : q);
            T c[V], xp[V], xm[V], yp[V], ym[V];
            int code[V];
            operands(src, f, c, xp, xm, yp, ym, code);
            VT o;
#pragma unroll
            for (int k = 0; k < V; ++k) {
                T v = mu_at<T, EXACT, 2>(A, c[k], xp[k], xm[k], yp[k], ym[k], (T)0, (T)0);
                if (GEN && (code[k] & GC_FIX)) v = ghost_value_xy(code[k], BCV);
                o[k] = v;
            }
            if (NQ % NT == 0 || i < NIT - 1 || q < NQ) *reinterpret_cast<VT*>(&SA[f]) = o;
        }
    };
    // stage state on ring R from the mu level on ring R+1: c_n + a*k, and the
    // accumulator update of stage S (kept on every position of the ring: only
    // the tile cells' entries are read)
    auto eval_state = [&](auto rr, auto ss, T a) {
        constexpr int R = decltype(rr)::value;
        constexpr int S = decltype(ss)::value;
        constexpr int F0 = (8 - R) * P, NQ = (TY + 2 * R) * P / V;
        constexpr int NIT = (NQ + NT - 1) / NT;
#pragma unroll
        for (int i = 0; i < NIT; ++i) {
            const int q = tid + NT * i;
            const int f = F0 + V * (NQ % NT != 0 && i == NIT - 1 ? min(q, NQ - 1) : q);
            T m[V], xp[V], xm[V], yp[V], ym[V];
            int code[V];
            operands(SA, f, m, xp, xm, yp, ym, code);
            const VT cc = *reinterpret_cast<const VT*>(&SC[f]);
            VT ac;
            if (S != 1) ac = *reinterpret_cast<const VT*>(&SACC[f]);
            VT vo, no;
#pragma unroll
            for (int k = 0; k < V; ++k) {
                T kk = k_at<T, EXACT, 2>(A, m[k], xp[k], xm[k], yp[k], ym[k], (T)0, (T)0);
                T v = axpy<T, EXACT>(a, kk, cc[k]);
                if (GEN && (code[k] & GC_FIX)) {
                    v = ghost_value_xy(code[k], BCV);
                    kk = (T)0;
                }
                T nac;
                if (S == 1) nac = EXACT ? kk : axpy<T, EXACT>(A.dt6, kk, cc[k]);
                else nac = EXACT ? ac[k] + (T)2 * kk : axpy<T, EXACT>(A.dt3, kk, ac[k]);
                vo[k] = v;
                no[k] = nac;
            }
            if (NQ % NT == 0 || i < NIT - 1 || q < NQ) {
                *reinterpret_cast<VT*>(&SB[f]) = vo;
                *reinterpret_cast<VT*>(&SACC[f]) = no;
            }
        }
    };
    using I1 = std::integral_constant<int, 1>;
    using I2 = std::integral_constant<int, 2>;
    using I3 = std::integral_constant<int, 3>;
    const T half = A.a1, full = A.a2;
    eval_mu(std::integral_constant<int, 7>{}, SC);
    __syncthreads();
    eval_state(std::integral_constant<int, 6>{}, I1{}, half);
    __syncthreads();
    eval_mu(std::integral_constant<int, 5>{}, SB);
    __syncthreads();
    eval_state(std::integral_constant<int, 4>{}, I2{}, half);
    __syncthreads();
    eval_mu(I3{}, SB);
    __syncthreads();
    eval_state(I2{}, I3{}, full);
    __syncthreads();
    eval_mu(I1{}, SB);
    __syncthreads();
    // stage 4 on the tile cells: thread tid owns column tid % TX of rows
    // tid / TX + (NT / TX) j
    constexpr int RS = NT / TX, NIT = TY / RS;
    static_assert(TY % RS == 0, "tile rows must be a multiple of NT / TX");
    const int ox = tid % TX, oy0 = tid / TX;
    const int fb = (8 + oy0) * P + 8 + ox;
    const int x = x0 + ox;
    const long long gb = zb + (long long)(y0 + oy0 + A.L.g) * A.L.sx + (x + A.L.g);
#pragma unroll
    for (int j = 0; j < NIT; ++j) {
        const int f = fb + j * RS * P;
        const T m = SA[f];
        const T k = k_at<T, EXACT, 2>(A, m, SA[f + 1], SA[f - 1], SA[f + P], SA[f - P], (T)0, (T)0);
        const T r = EXACT ? SC[f] + A.dt6 * (SACC[f] + k) : axpy<T, EXACT>(A.dt6, k, SACC[f]);
        if (x < A.L.n0 && y0 + oy0 + j * RS < A.L.n1) A.out[gb + (long long)j * RS * A.L.sx] = r;
    }
}

// Four explicit Euler steps of a 2D field in the whole-step frame: c in SC
// (ring 8), the mu levels in SA. Step s evaluates mu on ring 7-2s from the
// state on ring 8-2s, then the state c + dt*k on ring 6-2s in place (a run
// reads its own centre and SA's neighbours only); the last update writes the
// tile cells to global memory. Runs, frame and ghost rules (GEN: NOFLUX
// images by exchanged neighbours, FIXED constants) as in step2d_levels; the
// operations are the one-stage kernel's Euler step in the reference's order
// (mu_at, k_at, c + dt*k, exact form), so one launch gives the bits of four
// one-stage Euler launches.
template <typename T, int TX, int TY, int NT, bool GEN>
__device__ __forceinline__ void euler2d_levels(const PairArgs<T>& A, T* SC, T* SA, const T* BCV, int x0, int y0,
                                               long long zb) {
    constexpr int P = TX + 16;
    constexpr int V = 16 / (int)sizeof(T);
    typedef T VT __attribute__((ext_vector_type(V)));
    static_assert(P % V == 0, "frame rows hold whole runs");
    static_assert(NT % TX == 0, "tile rows must divide the workgroup");
    const int tid = threadIdx.x;
    auto operands = [&](const T* src, int f, T (&c)[V], T (&xp)[V], T (&xm)[V], T (&yp)[V], T (&ym)[V],
                        int (&code)[V]) {
        const VT cv = *reinterpret_cast<const VT*>(&src[f]);
        const VT uv = *reinterpret_cast<const VT*>(&src[f + P]);
        const VT dv = *reinterpret_cast<const VT*>(&src[f - P]);
        const T l = src[f - 1], r = src[f + V];
#pragma unroll
        for (int q = 0; q < V; ++q) {
            c[q] = cv[q];
            xp[q] = q + 1 < V ? cv[q + 1] : r;
            xm[q] = q > 0 ? cv[q - 1] : l;
            yp[q] = uv[q];
            ym[q] = dv[q];
            code[q] = GEN ? ghost_code_xy(x0 - 8 + (f + q) % P, y0 - 8 + (f + q) / P, A) : 0;
            if (GEN) {
                if (code[q] & GC_SWX) { const T t = xp[q]; xp[q] = xm[q]; xm[q] = t; }
                if (code[q] & GC_SWY) { const T t = yp[q]; yp[q] = ym[q]; ym[q] = t; }
            }
        }
    };
    auto eval_mu = [&](auto rr) {
        constexpr int R = decltype(rr)::value;
        constexpr int F0 = (8 - R) * P, NQ = (TY + 2 * R) * P / V;
        constexpr int NIT = (NQ + NT - 1) / NT;
#pragma unroll
        for (int i = 0; i < NIT; ++i) {
            const int q = tid + NT * i;
            const int f = F0 + V * (NQ % NT != 0 && i == NIT - 1 ? min(q, NQ - 1) : q);
            T c[V], xp[V], xm[V], yp[V], ym[V];
            int code[V];
            operands(SC, f, c, xp, xm, yp, ym, code);
            VT o;
#pragma unroll
            for (int k = 0; k < V; ++k) {
                T v = mu_at<T, true, 2>(A, c[k], xp[k], xm[k], yp[k], ym[k], (T)0, (T)0);
                if (GEN && (code[k] & GC_FIX)) v = ghost_value_xy(code[k], BCV);
                o[k] = v;
            }
            if (NQ % NT == 0 || i < NIT - 1 || q < NQ) *reinterpret_cast<VT*>(&SA[f]) = o;
        }
    };
    auto eval_update = [&](auto rr) {
        constexpr int R = decltype(rr)::value;
        constexpr int F0 = (8 - R) * P, NQ = (TY + 2 * R) * P / V;
        constexpr int NIT = (NQ + NT - 1) / NT;
#pragma unroll
        for (int i = 0; i < NIT; ++i) {
            const int q = tid + NT * i;
            const int f = F0 + V * (NQ % NT != 0 && i == NIT - 1 ? min(q, NQ - 1) : q);
            T m[V], xp[V], xm[V], yp[V], ym[V];
            int code[V];
            operands(SA, f, m, xp, xm, yp, ym, code);
            const VT cc = *reinterpret_cast<const VT*>(&SC[f]);
            VT vo;
#pragma unroll
            for (int k = 0; k < V; ++k) {
                T v = cc[k] + A.a1 * k_at<T, true, 2>(A, m[k], xp[k], xm[k], yp[k], ym[k], (T)0, (T)0);
                if (GEN && (code[k] & GC_FIX)) v = ghost_value_xy(code[k], BCV);
                vo[k] = v;
            }
            if (NQ % NT == 0 || i < NIT - 1 || q < NQ) *reinterpret_cast<VT*>(&SC[f]) = vo;
        }
    };
    eval_mu(std::integral_constant<int, 7>{});
    __syncthreads();
    eval_update(std::integral_constant<int, 6>{});
    __syncthreads();
    eval_mu(std::integral_constant<int, 5>{});
    __syncthreads();
    eval_update(std::integral_constant<int, 4>{});
    __syncthreads();
    eval_mu(std::integral_constant<int, 3>{});
    __syncthreads();
    eval_update(std::integral_constant<int, 2>{});
    __syncthreads();
    eval_mu(std::integral_constant<int, 1>{});
    __syncthreads();
    // the fourth update on the tile cells: thread tid owns column tid % TX of
    // rows tid / TX + (NT / TX) j
    constexpr int RS = NT / TX, NIT = TY / RS;
    static_assert(TY % RS == 0, "tile rows must be a multiple of NT / TX");
    const int ox = tid % TX, oy0 = tid / TX;
    const int fb = (8 + oy0) * P + 8 + ox;
    const int x = x0 + ox;
    const long long gb = zb + (long long)(y0 + oy0 + A.L.g) * A.L.sx + (x + A.L.g);
#pragma unroll
    for (int j = 0; j < NIT; ++j) {
        const int f = fb + j * RS * P;
        const T k = k_at<T, true, 2>(A, SA[f], SA[f + 1], SA[f - 1], SA[f + P], SA[f - P], (T)0, (T)0);
        const T r = SC[f] + A.a1 * k;
        if (x < A.L.n0 && y0 + oy0 + j * RS < A.L.n1) A.out[gb + (long long)j * RS * A.L.sx] = r;
    }
}

// ------------------------------------------------ 2D, whole step fused --
// All four RK4 stages of a 2D step in one launch: the block loads c_n on its
// tile plus an 8-cell ring (HBM -> LDS once) and evaluates, in LDS, the eight
// levels mu1 (ring 7), state 2 (ring 6), mu2 (5), state 3 (4), mu3 (3),
// state 4 (2), mu4 (1) and the output; the RK4 accumulator of the tile cells
// stays in LDS. A step costs 2 field passes over HBM (read c_n, write c_{n+1})
// and one launch instead of 16 passes and four launches. c_{n+1} goes to a
// buffer other than c_n (blocks read c_n in a ring around their tile). The
// ghost rules and operation order are those of the one-stage kernels (see the
// two-stage 3D kernel above), for ghost depths up to 8, hence n >= 8 per axis;
// results equal the one-stage kernels bit for bit.
//
// LDS: SC = c_n (ring 8, kept for the stage-state centres), SA = mu levels,
// SB = stage states 2..4, SACC = accumulator. A level reads the other buffer
// and is separated by a barrier from the level that last read the buffer it
// writes.
// EULER: four explicit Euler steps instead (see euler2d_levels; c_n and mu
// only, SB and SACC unused).
template <typename T, int TX, int TY, int NT, bool EXACT, bool EULER = false>
__global__ __launch_bounds__(NT) void k_ch_step_2d(PairArgs<T> A) {
    constexpr int WC = TX + 16, NC = WC * (TY + 16);
    // every level in the same frame as c_n (see step2d_levels); 16-byte
    // aligned for the levels' vector accesses
    __shared__ __attribute__((aligned(16))) T SC[NC];
    __shared__ __attribute__((aligned(16))) T SA[NC];
    __shared__ __attribute__((aligned(16))) T SB[EULER ? 4 : NC];
    __shared__ __attribute__((aligned(16))) T SACC[EULER ? 4 : NC];
    __shared__ T BCV[6];
    const int tid = threadIdx.x;
    stage_bc_values(BCV, A);
    int bx = blockIdx.x, by = blockIdx.y;
    if (A.xcd_remap) {
        const int gx = gridDim.x, nb = gx * gridDim.y;
        const int h = bx + gx * by;
        const int lb = (h & 7) * (nb >> 3) + (h >> 3);
        bx = lb % gx;
        by = lb / gx;
    }
    const int x0 = bx * TX, y0 = by * TY;
    const int g = A.L.g;
    const long long zb = (long long)g * A.L.sxy + (long long)blockIdx.z * A.bstride;
    const bool xy_in = x0 >= 8 && x0 + TX + 8 <= A.L.n0 && y0 >= 8 && y0 + TY + 8 <= A.L.n1;

    // c_n on the tile and an 8-cell ring, x/y BC images applied: offsets
    // first, then all loads back to back (no branch between them, so their
    // latencies overlap); FIXED and surplus slots read a valid cell and are
    // replaced at the store. Blocks whose ring lies inside the domain read
    // at affine offsets (no BC maps).
    constexpr int NSC = (NC + NT - 1) / NT;
    {
        const int cell0 = g * A.L.sx + g;
        int off[NSC];
        unsigned long long fm = 0;
        if (xy_in) {
            const int base = (y0 - 8 + g) * A.L.sx + (x0 - 8 + g);
#pragma unroll
            for (int i = 0; i < NSC; ++i) {
                const int t = min(tid + NT * i, NC - 1);
                const int ly = t / WC, lx = t - ly * WC;
                off[i] = base + ly * A.L.sx + lx;
            }
        } else {
#pragma unroll
            for (int i = 0; i < NSC; ++i) {
                const int t = min(tid + NT * i, NC - 1);
                const int ly = t / WC, lx = t - ly * WC;
                const int px = x0 - 8 + lx, py = y0 - 8 + ly;
                const int qy = map_deep<T, 8>(py, A.by), qx = map_deep<T, 8>(px, A.bx);
                off[i] = cell0;
                if (qy < 0 || qx < 0) fm |= 1ull << i;
                else off[i] = (qy + g) * A.L.sx + (qx + g);
            }
        }
        T r[NSC];
#pragma unroll
        for (int i = 0; i < NSC; ++i) r[i] = A.u[zb + off[i]];
#pragma unroll
        for (int i = 0; i < NSC; ++i) {
            const int t = tid + NT * i;
            if (NC % NT == 0 || t < NC) {
                T v = r[i];
                if ((fm >> i) & 1ull) {
                    const int ly = t / WC, lx = t - ly * WC;
                    const int px = x0 - 8 + lx, py = y0 - 8 + ly;
                    // once per block, before the barrier that publishes BCV
                    v = map_deep<T, 8>(py, A.by) < 0 ? fixed_value(py, A.by) : fixed_value(px, A.bx);
                }
                SC[t] = v;
            }
        }
    }
    __syncthreads();

    // Blocks whose 8-cell ring crosses a NOFLUX or FIXED face need the
    // per-position ghost rules; every other block (all of them on a periodic
    // grid: the wrapped load already holds the images) runs the same levels
    // with those rules compiled out.
    const bool per = A.bx.klo == PHX_BC_PERIODIC && A.bx.khi == PHX_BC_PERIODIC &&
                     A.by.klo == PHX_BC_PERIODIC && A.by.khi == PHX_BC_PERIODIC;
    if constexpr (EULER) {
        if (xy_in || per) euler2d_levels<T, TX, TY, NT, false>(A, SC, SA, BCV, x0, y0, zb);
        else euler2d_levels<T, TX, TY, NT, true>(A, SC, SA, BCV, x0, y0, zb);
    } else {
        if (xy_in || per) step2d_levels<T, TX, TY, NT, EXACT, false>(A, SC, SA, SB, SACC, BCV, x0, y0, zb);
        else step2d_levels<T, TX, TY, NT, EXACT, true>(A, SC, SA, SB, SACC, BCV, x0, y0, zb);
    }
}

// ---------------------------------------------------------- Euler pieces --
// Multi-step Euler for the CH app (Solver.cu:317-336):
//   K_mu : mu on every stored cell of the active region (physical: from c with
//          BC-mapped neighbours; ghost: the BC fill of mu, i.e. mu at the
//          mapped physical cell or the FIXED constant) and the BC fill of c's
//          ghosts. Reads only physical cells of c, writes only ghosts of c.
//   K_c  : c += dt*(M*lap(mu)) on physical cells, mu read with its ghosts.
template <typename T>
struct EulerArgs {
    T* c;
    T* mu;
    Layout L;
    AxisBC<T> b[3];
    T idx2, idy2, idz2;
    T two_rho, ca, cb, nkappa, M, dt;
};

// Value of c at physical (i,j,k) +- one step along `axis`, the neighbour
// rule of a single BC-filled ghost layer.
template <typename T>
__device__ __forceinline__ T c_nb(const EulerArgs<T>& A, int q[3], int axis, int step) {
    int r[3] = {q[0], q[1], q[2]};
    int p = q[axis] + step;
    const AxisBC<T>& b = A.b[axis];
    if (p < 0 || p >= b.n) {
        int kind = p < 0 ? b.klo : b.khi;
        if (kind == PHX_BC_FIXED) return p < 0 ? b.vlo : b.vhi;
        if (kind == PHX_BC_PERIODIC) p = p < 0 ? p + b.n : p - b.n;
        else p = p < 0 ? 0 : b.n - 1;
    }
    r[axis] = p;
    return A.c[A.L.at(r[0], r[1], r[2])];
}

template <typename T>
__device__ __forceinline__ T mu_phys(const EulerArgs<T>& A, int q[3]) {
    T c = A.c[A.L.at(q[0], q[1], q[2])];
    T val = (c_nb(A, q, 0, +1) - (T)2 * c + c_nb(A, q, 0, -1)) * A.idx2;
    if (A.L.dim >= 2) val += (c_nb(A, q, 1, +1) - (T)2 * c + c_nb(A, q, 1, -1)) * A.idy2;
    if (A.L.dim >= 3) val += (c_nb(A, q, 2, +1) - (T)2 * c + c_nb(A, q, 2, -1)) * A.idz2;
    return dwell(c, A.two_rho, A.ca, A.cb) + A.nkappa * val;
}

template <typename T>
__global__ __launch_bounds__(256) void k_euler_mu(EulerArgs<T> A) {
    // thread per stored cell of the active region (inactive axes pinned to 0)
    int si = blockIdx.x * 64 + threadIdx.x;
    int sj = blockIdx.y * 4 + threadIdx.y;
    int sk = blockIdx.z;
    const int dim = A.L.dim, g = A.L.g;
    if (si >= A.L.sx) return;
    if (dim >= 2 && sj >= A.L.sy) return;
    if (dim < 2 && sj > 0) return;
    int q[3] = {si - g, dim >= 2 ? sj - g : 0, dim >= 3 ? sk - g : 0};
    const long long s = A.L.at(q[0], q[1], q[2]);
    bool ghost = false;
    for (int a = 0; a < dim; ++a) ghost |= (q[a] < 0 || q[a] >= A.b[a].n);
    if (!ghost) {
        A.mu[s] = mu_phys(A, q);
        return;
    }
    // BC fill source of this ghost cell: axes in reverse fill order
    for (int a = dim - 1; a >= 0; --a) {
        const AxisBC<T>& b = A.b[a];
        int p = q[a];
        if (p >= 0 && p < b.n) continue;
        int kind = p < 0 ? b.klo : b.khi;
        if (kind == PHX_BC_FIXED) {
            T v = p < 0 ? b.vlo : b.vhi;
            A.mu[s] = v;
            A.c[s] = v;
            return;
        }
        if (kind == PHX_BC_PERIODIC) q[a] = p < 0 ? p + b.n : p - b.n;
        else q[a] = p < 0 ? 0 : b.n - 1;
    }
    A.mu[s] = mu_phys(A, q);
    A.c[s] = A.c[A.L.at(q[0], q[1], q[2])];
}

template <typename T>
__global__ __launch_bounds__(256) void k_euler_c(EulerArgs<T> A) {
    int i = blockIdx.x * 64 + threadIdx.x;
    int j = blockIdx.y * 4 + threadIdx.y;
    int k = blockIdx.z;
    if (i >= A.L.n0 || j >= A.L.n1) return;
    const long long s = A.L.at(i, j, k);
    const T* mu = A.mu;
    T m = mu[s];
    T val = (mu[s + 1] - (T)2 * m + mu[s - 1]) * A.idx2;
    if (A.L.dim >= 2) val += (mu[s + A.L.sx] - (T)2 * m + mu[s - A.L.sx]) * A.idy2;
    if (A.L.dim >= 3) val += (mu[s + A.L.sxy] - (T)2 * m + mu[s - A.L.sxy]) * A.idz2;
    A.c[s] = A.c[s] + A.dt * (A.M * val);
}

// ------------------------------------------------------------------ host --
template <typename T>
static AxisBC<T> axis_bc(const phx_grid* G, int a) {
    AxisBC<T> b;
    b.n = G->n[a];
    b.klo = a < G->dim ? G->bc[a][0] : PHX_BC_PERIODIC;
    b.khi = a < G->dim ? G->bc[a][1] : PHX_BC_PERIODIC;
    b.vlo = (T)G->bc_value[a][0];
    b.vhi = (T)G->bc_value[a][1];
    return b;
}

template <typename T>
static void fill_common(const phx_grid* G, const phx_ch_params* p, T& idx2, T& idy2, T& idz2, T& two_rho, T& ca,
                        T& cb, T& nk, T& M) {
    idx2 = (T)(1.0 / (G->d[0] * G->d[0]));
    idy2 = G->dim >= 2 ? (T)(1.0 / (G->d[1] * G->d[1])) : (T)0;
    idz2 = G->dim >= 3 ? (T)(1.0 / (G->d[2] * G->d[2])) : (T)0;
    two_rho = (T)(2.0 * p->rho);
    ca = (T)p->ca;
    cb = (T)p->cb;
    nk = (T)(-p->kappa);
    M = (T)p->M;
}

// FP32 arithmetic-form coefficients, from the same double values as the
// FP64 path's, each rounded once.
template <typename T>
static FastCoef<T> fast_coef(const phx_grid* G, const phx_ch_params* p) {
    const double ix = 1.0 / (G->d[0] * G->d[0]);
    const double iy = G->dim >= 2 ? 1.0 / (G->d[1] * G->d[1]) : 0.0;
    const double iz = G->dim >= 3 ? 1.0 / (G->d[2] * G->d[2]) : 0.0;
    const double sum = ix + iy + iz;
    FastCoef<T> F;
    F.icc = (T)(-2.0 * sum);
    F.mx = (T)(p->M * ix);
    F.my = (T)(p->M * iy);
    F.mz = (T)(p->M * iz);
    F.mcc = (T)(-2.0 * p->M * sum);
    F.ntrca = (T)(-2.0 * p->rho * p->ca);
    F.ncs = (T)(-(p->ca + p->cb));
    return F;
}

static int g_xcd_remap = 1;
static int g_euler_fused = 1;
// FP64 RK4 arithmetic form. 0 (default): the FP32 form evaluated in double --
// explicit FMAs, M folded into lap(mu)'s coefficients and, in 3D, the step
// finished from the stage states and b (rk4_b, rk4_fin_b: 12 field passes per
// step instead of 17);
// within a few ulp of the oracle per step. 1: the reference's operation order
// without contraction and its k1+2k2+2k3+k4 accumulator, bit-equal to the
// oracle. Multi-step Euler always uses the reference's order.
static int g_fp64_exact = 0;

// Multi-step Euler, one launch per step: the one-stage stencil kernel in its
// exact form (32x8 tiles in 1D/2D, the z-register kernel's 128x16 columns in
// 3D) evaluates mu on a ring around each tile (the ghost mu follows the BC
// fill of mu, as k_euler_mu's ghost cells do) and
// writes c + dt * (M lap mu), the operations of k_euler_mu / k_euler_c in the
// same order, so the results are the same bits. The state alternates
// between c and mu (free until the last step); every step but the last runs
// this way, the last one with the two kernels on c, which leaves c's ghost
// layer and mu as the two-kernel path does.
static int g_euler4_tile = 0;   // four-step Euler tile (phx_ch_euler_set_quad); -1: off
static bool euler4_ok(const phx_grid* G);
static bool zreg_ok(const phx_grid* G, int dtype);
template <typename T>
static void euler4_run(const phx_grid* G, T*& X, T*& Y, const phx_ch_params* p, int count, hipStream_t st);

// Blocks of four of those steps run as one launch of the whole-step
// kernel's Euler form when the grid allows it (2D, n >= 8; same bits), the
// rest one launch per step.
template <typename T>
static int euler_fused_steps_t(const phx_grid* G, T* c, T* mu, const phx_ch_params* p, int nsteps, hipStream_t st) {
    CHArgs<T> A;
    A.L = make_layout(G);
    A.bx = axis_bc<T>(G, 0);
    A.by = axis_bc<T>(G, 1);
    A.bz = axis_bc<T>(G, 2);
    fill_common<T>(G, p, A.idx2, A.idy2, A.idz2, A.two_rho, A.ca, A.cb, A.nkappa, A.M);
    A.F = fast_coef<T>(G, p);
    A.a = (T)p->dt;
    A.dt6 = A.dt3 = (T)0;
    A.third = (T)0;
    A.s2 = A.s3 = nullptr;
    A.stage = 0;
    // 3D: the z-register one-stage kernel, 128x16 columns per 512 threads, with
    // as many z-chunks (of at least 4 planes) as fill one round of resident
    // workgroups on the device -- a partial second round would cost a whole
    // round; the LDS-ring kernel's 64x8 columns over chunks of 16 planes where
    // a plane is too large for 32-bit offsets
    const bool zr = G->dim == 3 && zreg_ok(G, sizeof(T) == 8 ? PHX_F64 : PHX_F32);
    const bool zfull = A.L.n0 % 128 == 0 && A.L.n1 % 16 == 0;
    const auto kz = zfull ? &k_ch_stage_3dz<T, 128, 16, 512, true, true, 0>
                          : &k_ch_stage_3dz<T, 128, 16, 512, true, false, 0>;
    int ezc = 1;
    if (zr) {
        int dev = 0, ncu = 0, per_cu = 0;
        PHX_HIP(hipGetDevice(&dev));
        PHX_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, reinterpret_cast<const void*>(kz), 512, 0) !=
            hipSuccess) {
            (void)hipGetLastError();
            per_cu = 1;
        }
        const int ncol = cdiv(A.L.n0, 128) * cdiv(A.L.n1, 16);
        ezc = std::max(1, std::min(std::max(1, A.L.n2 / 4), ncu * std::max(per_cu, 1) / ncol));
    }
    const dim3 grid = zr ? dim3(cdiv(A.L.n0, 128), cdiv(A.L.n1, 16), ezc)
                    : G->dim == 3 ? dim3(cdiv(A.L.n0, 64), cdiv(A.L.n1, 8), cdiv(A.L.n2, 16))
                                  : dim3(cdiv(A.L.n0, 32), G->dim >= 2 ? cdiv(A.L.n1, 8) : 1, 1);
    A.zc = G->dim == 3 ? 16 : 1;
    A.xcd_remap = g_xcd_remap && (grid.x * grid.y * grid.z) % 8 == 0;
    const int fused = nsteps - 1;
    const int quads = g_euler4_tile >= 0 && euler4_ok(G) ? fused / 4 : 0;
    const int singles = fused - 4 * quads;
    T* X = c;
    T* Y = mu;
    if ((quads + singles) % 2) {   // odd launch count: start from a copy in mu so the state ends in c
        PHX_HIP(hipMemcpyAsync(mu, c, (size_t)A.L.total() * sizeof(T), hipMemcpyDeviceToDevice, st));
        X = mu;
        Y = c;
    }
    if (quads) euler4_run<T>(G, X, Y, p, quads, st);
    for (int s = 0; s < singles; ++s) {
        A.u = X;
        A.c = X;
        A.acc = Y;
        A.out = Y;
        if (zr) hipLaunchKernelGGL(kz, grid, dim3(512), 0, st, A);
        else if (G->dim == 3) hipLaunchKernelGGL((k_ch_stage_3d<T, 64, 8, true>), grid, dim3(256), 0, st, A);
        else if (G->dim == 2) hipLaunchKernelGGL((k_ch_stage_2d<T, 2, 32, 8, true>), grid, dim3(256), 0, st, A);
        else hipLaunchKernelGGL((k_ch_stage_2d<T, 1, 32, 8, true>), grid, dim3(256), 0, st, A);
        std::swap(X, Y);
    }
    return PHX_OK;
}

template <typename T>
static int euler_steps_t(const phx_grid* G, T* c, T* mu, const phx_ch_params* p, int nsteps, hipStream_t st) {
    if (g_euler_fused && nsteps >= 2) {
        int rc = euler_fused_steps_t<T>(G, c, mu, p, nsteps, st);
        if (rc) return rc;
        nsteps = 1;
    }
    EulerArgs<T> A;
    A.c = c;
    A.mu = mu;
    A.L = make_layout(G);
    for (int a = 0; a < 3; ++a) A.b[a] = axis_bc<T>(G, a);
    fill_common<T>(G, p, A.idx2, A.idy2, A.idz2, A.two_rho, A.ca, A.cb, A.nkappa, A.M);
    A.dt = (T)p->dt;
    const Layout& L = A.L;
    dim3 blk(64, 4);
    dim3 gmu(cdiv(L.sx, 64), L.dim >= 2 ? cdiv(L.sy, 4) : 1, L.dim >= 3 ? L.sz : 1);
    dim3 gc(cdiv(L.n0, 64), cdiv(L.n1, 4), L.n2);
    for (int s = 0; s < nsteps; ++s) {
        hipLaunchKernelGGL(k_euler_mu<T>, gmu, blk, 0, st, A);
        hipLaunchKernelGGL(k_euler_c<T>, gc, blk, 0, st, A);
    }
    PHX_LAUNCH_CHECK("phx_ch_euler_steps");
    return PHX_OK;
}

int bc_fill_impl(const phx_grid* G, int dtype, void* data, hipStream_t st);
int bc_fill_batch_impl(const phx_grid* G, int dtype, void* data, int nbatch, hipStream_t st);

// dst <- src on the physical cells of nbatch 2D fields back to back (blockIdx.y)
template <typename T>
__global__ __launch_bounds__(256) void k_copy_phys2d(T* __restrict__ dst, const T* __restrict__ src, Layout L,
                                                     long long bstride) {
    const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
    if (i >= (long long)L.n0 * L.n1) return;
    const int y = (int)(i / L.n0), x = (int)(i - (long long)y * L.n0);
    const long long o = (long long)blockIdx.y * bstride + L.at(x, y, 0);
    dst[o] = src[o];
}

// Tile shapes (autotuned later; these are the defaults).
constexpr int T2X = 64, T2Y = 16;
// 3D tile shapes the autotuner chooses from (TX x TY columns, 256 threads).
struct Tile3 {
    int tx, ty;
};
static const Tile3 kTiles3[] = {{64, 16}, {64, 8}, {128, 8}, {32, 16}, {64, 4}};
constexpr int kNumTiles3 = sizeof(kTiles3) / sizeof(kTiles3[0]);

// Two-stage kernel tiles; launch tile indices kNumTiles3.. select them.
static const Tile3 kTilesP[] = {{64, 8}, {32, 16}, {32, 8}, {64, 4}};
constexpr int kNumTilesP = sizeof(kTilesP) / sizeof(kTilesP[0]);

// One-stage kernel with z-neighbours in registers (TX x TY columns, NT
// threads); launch tile indices kFirstZ.. select them.
struct TileZ {
    int tx, ty, nt;
};
static const TileZ kTilesZ[] = {{64, 16, 256}, {128, 8, 256},  {128, 16, 512},  {256, 8, 512}, {256, 16, 1024},
                                {256, 4, 256},  {256, 8, 1024}, {128, 16, 1024}, {128, 8, 512}};
constexpr int kNumTilesZ = sizeof(kTilesZ) / sizeof(kTilesZ[0]);
constexpr int kFirstZ = kNumTiles3 + kNumTilesP;

// Columns of a 3D launch tile (any kind) and its workgroup size.
static Tile3 shape3(int tile) {
    if (tile < kNumTiles3) return kTiles3[tile];
    if (tile < kFirstZ) return kTilesP[tile - kNumTiles3];
    return Tile3{kTilesZ[tile - kFirstZ].tx, kTilesZ[tile - kFirstZ].ty};
}
static int threads3(int tile) { return tile >= kFirstZ ? kTilesZ[tile - kFirstZ].nt : 256; }
static bool is_pair_tile(int tile) { return tile >= kNumTiles3 && tile < kFirstZ; }
// The z-register kernel addresses a plane with 32-bit byte offsets.
static bool zreg_ok(const phx_grid* G, int dtype) {
    const Layout L = make_layout(G);
    return L.sxy * (dtype == PHX_F64 ? 8LL : 4LL) < (1LL << 32);
}

// z-register launch: the FULL form when every tile lies inside the domain
template <typename T, int TX, int TY, int NT, bool EXACT, bool FULL>
static void launch_zs(dim3 grid, const CHArgs<T>& A, hipStream_t st) {
    switch (A.stage) {
        case 1: hipLaunchKernelGGL((k_ch_stage_3dz<T, TX, TY, NT, EXACT, FULL, 1>), grid, dim3(NT), 0, st, A); break;
        case 4: hipLaunchKernelGGL((k_ch_stage_3dz<T, TX, TY, NT, EXACT, FULL, 4>), grid, dim3(NT), 0, st, A); break;
        case 2:
            if constexpr (kBForm<T, EXACT>) {
                hipLaunchKernelGGL((k_ch_stage_3dz<T, TX, TY, NT, EXACT, FULL, 3>), grid, dim3(NT), 0, st, A);
                break;
            }
            [[fallthrough]];
        default: hipLaunchKernelGGL((k_ch_stage_3dz<T, TX, TY, NT, EXACT, FULL, 2>), grid, dim3(NT), 0, st, A);
    }
}

template <typename T, int TX, int TY, int NT, bool EXACT>
static void launch_z(dim3 grid, const CHArgs<T>& A, hipStream_t st) {
    if (A.L.n0 % TX == 0 && A.L.n1 % TY == 0) launch_zs<T, TX, TY, NT, EXACT, true>(grid, A, st);
    else launch_zs<T, TX, TY, NT, EXACT, false>(grid, A, st);
}

template <typename T, bool EXACT>
static void launch3(int tile, dim3 grid, const CHArgs<T>& A, hipStream_t st) {
    switch (tile) {
        case 0: hipLaunchKernelGGL((k_ch_stage_3d<T, 64, 16, EXACT>), grid, dim3(256), 0, st, A); break;
        case 1: hipLaunchKernelGGL((k_ch_stage_3d<T, 64, 8, EXACT>), grid, dim3(256), 0, st, A); break;
        case 2: hipLaunchKernelGGL((k_ch_stage_3d<T, 128, 8, EXACT>), grid, dim3(256), 0, st, A); break;
        case 3: hipLaunchKernelGGL((k_ch_stage_3d<T, 32, 16, EXACT>), grid, dim3(256), 0, st, A); break;
        case 4: hipLaunchKernelGGL((k_ch_stage_3d<T, 64, 4, EXACT>), grid, dim3(256), 0, st, A); break;
        case kFirstZ + 0: launch_z<T, 64, 16, 256, EXACT>(grid, A, st); break;
        case kFirstZ + 1: launch_z<T, 128, 8, 256, EXACT>(grid, A, st); break;
        case kFirstZ + 2: launch_z<T, 128, 16, 512, EXACT>(grid, A, st); break;
        case kFirstZ + 3: launch_z<T, 256, 8, 512, EXACT>(grid, A, st); break;
        case kFirstZ + 4:
            launch_z<T, 256, 16, 1024, EXACT>(grid, A, st);
            break;
        case kFirstZ + 5: launch_z<T, 256, 4, 256, EXACT>(grid, A, st); break;
        case kFirstZ + 6:
            launch_z<T, 256, 8, 1024, EXACT>(grid, A, st);
            break;
        case kFirstZ + 7:
            launch_z<T, 128, 16, 1024, EXACT>(grid, A, st);
            break;
        default: launch_z<T, 128, 8, 512, EXACT>(grid, A, st); break;
    }
}

template <typename T, bool EXACT, bool SECOND>
static void launch_pair(int tile, dim3 grid, const PairArgs<T>& P, hipStream_t st) {
    switch (tile) {
        case 0: hipLaunchKernelGGL((k_ch_pair_3d<T, 64, 8, EXACT, SECOND>), grid, dim3(256), 0, st, P); break;
        case 1: hipLaunchKernelGGL((k_ch_pair_3d<T, 32, 16, EXACT, SECOND>), grid, dim3(256), 0, st, P); break;
        case 2: hipLaunchKernelGGL((k_ch_pair_3d<T, 32, 8, EXACT, SECOND>), grid, dim3(256), 0, st, P); break;
        default: hipLaunchKernelGGL((k_ch_pair_3d<T, 64, 4, EXACT, SECOND>), grid, dim3(256), 0, st, P); break;
    }
}

// The two-stage kernel needs every axis of a 3D grid to hold its depth-4 BC
// images inside the domain.
static bool pair_ok(const phx_grid* G) { return G->dim == 3 && G->n[0] >= 4 && G->n[1] >= 4 && G->n[2] >= 4; }

// 2D whole-step kernel tiles; launch tile indices 1.. select them in 2D
// (0 is the one-stage kernel).
static const TileZ kTilesS2[] = {{64, 16, 256}, {64, 32, 256}, {128, 16, 256}, {32, 32, 256},
                                 {32, 32, 512}, {64, 32, 512}, {128, 32, 1024}, {32, 64, 512}};
constexpr int kNumTilesS2 = sizeof(kTilesS2) / sizeof(kTilesS2[0]);

template <typename T, bool EXACT>
static void launch_step2d(int tile, dim3 grid, const PairArgs<T>& P, hipStream_t st) {
    switch (tile) {
        case 0: hipLaunchKernelGGL((k_ch_step_2d<T, 64, 16, 256, EXACT>), grid, dim3(256), 0, st, P); break;
        case 1: hipLaunchKernelGGL((k_ch_step_2d<T, 64, 32, 256, EXACT>), grid, dim3(256), 0, st, P); break;
        case 2: hipLaunchKernelGGL((k_ch_step_2d<T, 128, 16, 256, EXACT>), grid, dim3(256), 0, st, P); break;
        case 3: hipLaunchKernelGGL((k_ch_step_2d<T, 32, 32, 256, EXACT>), grid, dim3(256), 0, st, P); break;
        case 4: hipLaunchKernelGGL((k_ch_step_2d<T, 32, 32, 512, EXACT>), grid, dim3(512), 0, st, P); break;
        case 5: hipLaunchKernelGGL((k_ch_step_2d<T, 64, 32, 512, EXACT>), grid, dim3(512), 0, st, P); break;
        case 6:
            // FP32 only: 128x32 in FP64 would need 221 KB of LDS (step2d_tile maps it to 5)
            if constexpr (sizeof(T) == 4)
                hipLaunchKernelGGL((k_ch_step_2d<T, 128, 32, 1024, EXACT>), grid, dim3(1024), 0, st, P);
            break;
        default: hipLaunchKernelGGL((k_ch_step_2d<T, 32, 64, 512, EXACT>), grid, dim3(512), 0, st, P); break;
    }
}

// Whole-step tile actually launched for tile index stile: 128x32 (6) needs
// more LDS than a CU has in FP64 and runs as 64x32 (5) there.
static int step2d_tile(int stile, size_t esz) { return stile == 6 && esz == 8 ? 5 : stile; }

// The 2D whole-step kernel needs depth-8 BC images inside the domain.
static bool step2d_ok(const phx_grid* G) { return G->dim == 2 && G->n[0] >= 8 && G->n[1] >= 8; }

// Four-step Euler launches (the whole-step kernel's EULER form): tiles
// 0: 32x16/512 (fastest at 128^2), 1: 32x32/512, 2: 64x16/256, 3: 32x16/256;
// -1 disables them.
static const TileZ kTilesE4[] = {{32, 16, 512}, {32, 32, 512}, {64, 16, 256}, {32, 16, 256}};
constexpr int kNumTilesE4 = sizeof(kTilesE4) / sizeof(kTilesE4[0]);

template <typename T>
static void launch_euler4(int tile, dim3 grid, const PairArgs<T>& P, hipStream_t st) {
    switch (tile) {
        case 1: hipLaunchKernelGGL((k_ch_step_2d<T, 32, 32, 512, true, true>), grid, dim3(512), 0, st, P); break;
        case 2: hipLaunchKernelGGL((k_ch_step_2d<T, 64, 16, 256, true, true>), grid, dim3(256), 0, st, P); break;
        case 3: hipLaunchKernelGGL((k_ch_step_2d<T, 32, 16, 256, true, true>), grid, dim3(256), 0, st, P); break;
        default: hipLaunchKernelGGL((k_ch_step_2d<T, 32, 16, 512, true, true>), grid, dim3(512), 0, st, P); break;
    }
}

// The four-step Euler form: 2D with depth-8 BC images inside the domain.
static bool euler4_ok(const phx_grid* G) { return step2d_ok(G); }

// `count` launches of four Euler steps each, X -> Y, swapping after each.
template <typename T>
static void euler4_run(const phx_grid* G, T*& X, T*& Y, const phx_ch_params* p, int count, hipStream_t st) {
    PairArgs<T> P;
    P.L = make_layout(G);
    P.bx = axis_bc<T>(G, 0);
    P.by = axis_bc<T>(G, 1);
    P.bz = axis_bc<T>(G, 2);
    fill_common<T>(G, p, P.idx2, P.idy2, P.idz2, P.two_rho, P.ca, P.cb, P.nkappa, P.M);
    P.F = fast_coef<T>(G, p);
    P.a1 = (T)p->dt;
    P.a2 = (T)0;
    P.dt6 = P.dt3 = P.third = (T)0;
    P.c = nullptr;
    P.acc = nullptr;
    P.zc = 1;
    P.bstride = P.L.total();
    const int tile = std::min(g_euler4_tile, kNumTilesE4 - 1);
    const TileZ tl = kTilesE4[tile];
    const dim3 grid(cdiv(P.L.n0, tl.tx), cdiv(P.L.n1, tl.ty), 1);
    P.xcd_remap = g_xcd_remap && (grid.x * grid.y) % 8 == 0;
    for (int i = 0; i < count; ++i) {
        P.u = X;
        P.out = Y;
        launch_euler4<T>(tile, grid, P, st);
        std::swap(X, Y);
    }
}

// Optional per-launch timing for benchmarks: while enabled, every launch of
// the fused RK4 kernels is bracketed by HIP events on its stream, tagged with
// its kind (0: one-stage kernel, 1: two-stage (1,2), 2: two-stage (3,4),
// 3: 2D whole step; 4-9: the Allen-Cahn and ensemble kernels, see
// common.hpp).
// Not thread safe; meant for one benchmarking thread.
struct LaunchTimer {
    int on = 0;
    std::vector<hipEvent_t> ev;
    std::vector<int> kind;
    size_t used = 0;
    int pending = 0;   // begin() recorded an event that end() must close
};
static LaunchTimer g_lt;

// Set while an entry point that must not allocate runs on this thread: its
// launches are never timed (the timer creates events on demand).
static thread_local int t_timer_off = 0;

// A launch is timed only when the timer is on, this thread allows it and the
// stream is not being captured into a graph (an event recorded into a graph
// would time capture, not replay); end() follows what begin() decided.
void launch_timer_begin(hipStream_t st) {
    g_lt.pending = 0;
    if (!g_lt.on || t_timer_off) return;
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(st, &cs) != hipSuccess || cs != hipStreamCaptureStatusNone) {
        (void)hipGetLastError();
        return;
    }
    if (2 * g_lt.used + 2 > g_lt.ev.size()) {
        for (int i = 0; i < 2; ++i) {
            hipEvent_t e;
            if (hipEventCreate(&e) != hipSuccess) {
                g_lt.on = 0;
                return;
            }
            g_lt.ev.push_back(e);
        }
        g_lt.kind.push_back(0);
    }
    (void)hipEventRecord(g_lt.ev[2 * g_lt.used], st);
    g_lt.pending = 1;
}

void launch_timer_end(int kind, hipStream_t st) {
    if (!g_lt.pending) return;
    g_lt.pending = 0;
    (void)hipEventRecord(g_lt.ev[2 * g_lt.used + 1], st);
    g_lt.kind[g_lt.used] = kind;
    ++g_lt.used;
}

static inline void lt_begin(hipStream_t st) { launch_timer_begin(st); }
static inline void lt_end(int kind, hipStream_t st) { launch_timer_end(kind, st); }


// Resident workgroups per CU of the 3D kernel of launch tile `tile` (one-stage
// 0-4; two-stage 5-8, sized by its (3,4) kernel; z-register 9-17, sized by
// its stage 2-3 kernel for tiles that do not cover the domain exactly), from
// the occupancy calculator; 0 if unknown.
template <typename T, bool EXACT>
static int blocks_per_cu(int tile) {
    const void* f = nullptr;
    switch (tile) {
        case 0: f = reinterpret_cast<const void*>(&k_ch_stage_3d<T, 64, 16, EXACT>); break;
        case 1: f = reinterpret_cast<const void*>(&k_ch_stage_3d<T, 64, 8, EXACT>); break;
        case 2: f = reinterpret_cast<const void*>(&k_ch_stage_3d<T, 128, 8, EXACT>); break;
        case 3: f = reinterpret_cast<const void*>(&k_ch_stage_3d<T, 32, 16, EXACT>); break;
        case 4: f = reinterpret_cast<const void*>(&k_ch_stage_3d<T, 64, 4, EXACT>); break;
        case 5: f = reinterpret_cast<const void*>(&k_ch_pair_3d<T, 64, 8, EXACT, true>); break;
        case 6: f = reinterpret_cast<const void*>(&k_ch_pair_3d<T, 32, 16, EXACT, true>); break;
        case 7: f = reinterpret_cast<const void*>(&k_ch_pair_3d<T, 32, 8, EXACT, true>); break;
        case 8: f = reinterpret_cast<const void*>(&k_ch_pair_3d<T, 64, 4, EXACT, true>); break;
        case kFirstZ + 0: f = reinterpret_cast<const void*>(&k_ch_stage_3dz<T, 64, 16, 256, EXACT, false, 2>); break;
        case kFirstZ + 1: f = reinterpret_cast<const void*>(&k_ch_stage_3dz<T, 128, 8, 256, EXACT, false, 2>); break;
        case kFirstZ + 2: f = reinterpret_cast<const void*>(&k_ch_stage_3dz<T, 128, 16, 512, EXACT, false, 2>); break;
        case kFirstZ + 3: f = reinterpret_cast<const void*>(&k_ch_stage_3dz<T, 256, 8, 512, EXACT, false, 2>); break;
        case kFirstZ + 4: f = reinterpret_cast<const void*>(&k_ch_stage_3dz<T, 256, 16, 1024, EXACT, false, 2>); break;
        case kFirstZ + 5: f = reinterpret_cast<const void*>(&k_ch_stage_3dz<T, 256, 4, 256, EXACT, false, 2>); break;
        case kFirstZ + 6: f = reinterpret_cast<const void*>(&k_ch_stage_3dz<T, 256, 8, 1024, EXACT, false, 2>); break;
        case kFirstZ + 7: f = reinterpret_cast<const void*>(&k_ch_stage_3dz<T, 128, 16, 1024, EXACT, false, 2>); break;
        case kFirstZ + 8: f = reinterpret_cast<const void*>(&k_ch_stage_3dz<T, 128, 8, 512, EXACT, false, 2>); break;
        default: return 0;
    }
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, f, threads3(tile), 0) != hipSuccess) {
        (void)hipGetLastError();
        return 0;
    }
    return nb;
}

// One-stage RK4 steps. In 3D each stage class -- 0: stage 1, 1: stages 2
// and 3, 2: stage 4 -- may run its own one-stage launch (tile, z-chunk):
// every one-stage kernel reads and writes the same stage buffers and returns
// the same bits, so any mix gives the results of a uniform launch.
template <typename T, bool EXACT>
static int rk4_steps_t(const phx_grid* G, T* c, T* work, const phx_ch_params* p, int nsteps, hipStream_t st,
                       const int (&zc)[3], const int (&tile)[3], int skip_fill = 0) {
    CHArgs<T> A;
    A.L = make_layout(G);
    const long long N = A.L.total();
    T* w0 = work;
    T* w1 = work + N;
    T* acc = work + 2 * N;
    A.bx = axis_bc<T>(G, 0);
    A.by = axis_bc<T>(G, 1);
    A.bz = axis_bc<T>(G, 2);
    fill_common<T>(G, p, A.idx2, A.idy2, A.idz2, A.two_rho, A.ca, A.cb, A.nkappa, A.M);
    A.F = fast_coef<T>(G, p);
    const double dt = p->dt, dt6 = dt / 6.0;
    A.dt6 = (T)dt6;
    A.dt3 = (T)(2.0 * dt6);
    A.c = c;
    A.acc = acc;
    A.s2 = acc;
    A.s3 = w0;
    A.third = (T)(1.0 / 3.0);
    A.zc = zc[1];
    A.xcd_remap = 0;
    const T shift[4] = {(T)(dt * 0.5), (T)(dt * 0.5), (T)dt, (T)0};
    // exact form and 2D: stage states in w0, w1, w0 and the accumulator in acc;
    // 3D FP32: u2, u3, u4 in acc, w0, w1 (stage 4 reads u2 and u3);
    // 3D FP64: u2, u3, u4 in acc, w0, acc and b in w1 (stage 4 reads b)
    const bool states = G->dim == 3 && !EXACT;
    const bool bform = states && kBForm<T, EXACT>;
    const T* uin[4] = {c, states ? acc : w0, states ? w0 : w1, bform ? acc : (states ? w1 : w0)};
    T* uout[4] = {states ? acc : w0, states ? w0 : w1, bform ? acc : (states ? w1 : w0), c};
    if (bform) A.acc = w1;
    dim3 grid[3];
    int remap[3];
    for (int k = 0; k < 3; ++k) {
        if (G->dim == 3) {
            const Tile3 tl = shape3(tile[k]);
            grid[k] = dim3(cdiv(A.L.n0, tl.tx), cdiv(A.L.n1, tl.ty), cdiv(A.L.n2, zc[k]));
            remap[k] = g_xcd_remap && (grid[k].x * grid[k].y * grid[k].z) % 8 == 0;
        } else {
            grid[k] = dim3(cdiv(A.L.n0, T2X), cdiv(A.L.n1, T2Y), 1);
            remap[k] = g_xcd_remap && (grid[k].x * grid[k].y) % 8 == 0;
        }
    }
    for (int s = 0; s < nsteps; ++s) {
        if (s == nsteps - 1 && !skip_fill) {
            int rc = bc_fill_impl(G, sizeof(T) == 8 ? PHX_F64 : PHX_F32, c, st);
            if (rc) return rc;
        }
        for (int stg = 0; stg < 4; ++stg) {
            const int k = stg == 0 ? 0 : (stg == 3 ? 2 : 1);
            const dim3 gr = grid[k];
            A.stage = stg + 1;
            A.u = uin[stg];
            A.out = uout[stg];
            A.a = shift[stg];
            A.zc = zc[k];
            A.xcd_remap = remap[k];
            lt_begin(st);
            if (G->dim == 3) launch3<T, EXACT>(tile[k], gr, A, st);
            else if (G->dim == 2) hipLaunchKernelGGL((k_ch_stage_2d<T, 2, T2X, T2Y, EXACT>), gr, dim3(256), 0, st, A);
            else hipLaunchKernelGGL((k_ch_stage_2d<T, 1, T2X, T2Y, EXACT>), gr, dim3(256), 0, st, A);
            lt_end(0, st);
        }
    }
    PHX_LAUNCH_CHECK("phx_ch_rk4_steps");
    return PHX_OK;
}

// The same with one launch for every stage.
template <typename T, bool EXACT>
static int rk4_steps_t(const phx_grid* G, T* c, T* work, const phx_ch_params* p, int nsteps, hipStream_t st,
                       int zc, int tile, int skip_fill = 0) {
    const int z3[3] = {zc, zc, zc}, t3[3] = {tile, tile, tile};
    return rk4_steps_t<T, EXACT>(G, c, work, p, nsteps, st, z3, t3, skip_fill);
}

// RK4 steps with the two-stage kernel (3D, tile index ptile into kTilesP).
// The (3,4) launch cannot update c_n in place, so the state alternates between
// c and the work buffer w1; the last step runs (1,2) and then stages 3 and 4
// with the one-stage kernel, which ends in c. When the steps before the last
// are odd in number, the first step runs with the one-stage kernel so that the
// alternation ends in c. As in the one-stage path the ghost layer of c is the
// BC fill of the state before the last step.
template <typename T, bool EXACT>
static int rk4_pair_steps_t(const phx_grid* G, T* c, T* work, const phx_ch_params* p, int nsteps, hipStream_t st,
                            int zc, int ptile, int pure_pairs) {
    const Layout L = make_layout(G);
    const long long N = L.total();
    T* w0 = work;
    T* w1 = work + N;
    T* acc = work + 2 * N;
    PairArgs<T> P;
    P.L = L;
    P.bx = axis_bc<T>(G, 0);
    P.by = axis_bc<T>(G, 1);
    P.bz = axis_bc<T>(G, 2);
    fill_common<T>(G, p, P.idx2, P.idy2, P.idz2, P.two_rho, P.ca, P.cb, P.nkappa, P.M);
    P.F = fast_coef<T>(G, p);
    const double dt = p->dt, dt6 = dt / 6.0;
    P.dt6 = (T)dt6;
    P.dt3 = (T)(2.0 * dt6);
    P.third = (T)(1.0 / 3.0);
    P.acc = acc;
    P.zc = zc;
    const Tile3 tl = kTilesP[ptile];
    const dim3 grid(cdiv(L.n0, tl.tx), cdiv(L.n1, tl.ty), cdiv(L.n2, zc));
    P.xcd_remap = g_xcd_remap && (grid.x * grid.y * grid.z) % 8 == 0;
    auto pair12 = [&](const T* x) {
        P.u = x;
        P.c = x;
        P.out = w0;
        P.a1 = (T)(dt * 0.5);
        P.a2 = (T)(dt * 0.5);
        lt_begin(st);
        launch_pair<T, EXACT, false>(ptile, grid, P, st);
        lt_end(1, st);
    };
    auto pair34 = [&](const T* x, T* y) {
        P.u = w0;
        P.c = x;
        P.out = y;
        P.a1 = (T)dt;
        P.a2 = (T)0;
        lt_begin(st);
        launch_pair<T, EXACT, true>(ptile, grid, P, st);
        lt_end(2, st);
    };
    if (pure_pairs) {
        // timing helper: nsteps (even) two-stage steps, state ends in c
        T* x = c;
        T* y = w1;
        for (int s = 0; s < nsteps; ++s) {
            pair12(x);
            pair34(x, y);
            std::swap(x, y);
        }
        PHX_LAUNCH_CHECK("phx_ch_rk4_steps");
        return PHX_OK;
    }
    int rest = nsteps - 1;
    if (rest % 2) {
        // one one-stage step, in place
        int rc = rk4_steps_t<T, EXACT>(G, c, work, p, 1, st, 16, sizeof(T) == 8 ? 1 : 0, /*skip_fill=*/1);
        if (rc) return rc;
        --rest;
    }
    T* x = c;
    T* y = w1;
    for (int s = 0; s < rest; ++s) {
        pair12(x);
        pair34(x, y);
        std::swap(x, y);
    }
    int rc = bc_fill_impl(G, sizeof(T) == 8 ? PHX_F64 : PHX_F32, c, st);
    if (rc) return rc;
    pair12(c);
    // stages 3 and 4 with the one-stage kernel: w0 -> w1 -> c
    CHArgs<T> A;
    A.L = L;
    A.bx = P.bx;
    A.by = P.by;
    A.bz = P.bz;
    fill_common<T>(G, p, A.idx2, A.idy2, A.idz2, A.two_rho, A.ca, A.cb, A.nkappa, A.M);
    A.F = fast_coef<T>(G, p);
    A.dt6 = P.dt6;
    A.dt3 = P.dt3;
    A.c = c;
    A.acc = acc;
    A.s2 = acc;   // FP32: u2 and u3 from the (1,2) launch
    A.s3 = w0;
    A.third = P.third;
    A.zc = 16;
    const int stile = sizeof(T) == 8 ? 1 : 0;
    const dim3 sgrid(cdiv(L.n0, kTiles3[stile].tx), cdiv(L.n1, kTiles3[stile].ty), cdiv(L.n2, 16));
    A.xcd_remap = g_xcd_remap && (sgrid.x * sgrid.y * sgrid.z) % 8 == 0;
    A.stage = 3;
    A.u = w0;
    A.out = w1;
    A.a = (T)dt;
    lt_begin(st);
    launch3<T, EXACT>(stile, sgrid, A, st);
    lt_end(0, st);
    A.stage = 4;
    A.u = w1;
    A.out = c;
    A.a = (T)0;
    lt_begin(st);
    launch3<T, EXACT>(stile, sgrid, A, st);
    lt_end(0, st);
    PHX_LAUNCH_CHECK("phx_ch_rk4_steps");
    return PHX_OK;
}

// RK4 steps with the 2D whole-step kernel (tile index stile into kTilesS2).
// The state alternates between two buffers X and Y, chosen so that the last
// step reads the caller's c (X = c for an odd step count; for an even one c
// is first copied to w1 and X = w1, Y = c). Before the last step c's ghost
// layer gets the BC fill of that state; the last step writes w0 and its
// physical cells are copied into c. c thus ends with the state and the ghost
// layer the one-stage path leaves, and every step runs the whole-step kernel.
// nbatch > 1: nbatch fields of this grid back to back in c (and each work
// buffer), all advanced by every launch (blockIdx.z = field).
template <typename T, bool EXACT>
static int rk4_step2d_steps_t(const phx_grid* G, T* c, T* work, const phx_ch_params* p, int nsteps, hipStream_t st,
                              int stile, int pure, int nbatch = 1) {
    const Layout L = make_layout(G);
    const long long N = L.total();
    const long long NB = N * nbatch;
    T* w0 = work;
    T* w1 = work + NB;
    PairArgs<T> P;
    P.L = L;
    P.bx = axis_bc<T>(G, 0);
    P.by = axis_bc<T>(G, 1);
    P.bz = axis_bc<T>(G, 2);
    fill_common<T>(G, p, P.idx2, P.idy2, P.idz2, P.two_rho, P.ca, P.cb, P.nkappa, P.M);
    P.F = fast_coef<T>(G, p);
    const double dt = p->dt, dt6 = dt / 6.0;
    P.dt6 = (T)dt6;
    P.dt3 = (T)(2.0 * dt6);
    P.third = (T)(1.0 / 3.0);
    P.a1 = (T)(dt * 0.5);
    P.a2 = (T)dt;
    P.c = nullptr;
    P.acc = nullptr;
    P.zc = 1;
    stile = step2d_tile(stile, sizeof(T));
    const TileZ tl = kTilesS2[stile];
    const dim3 grid(cdiv(L.n0, tl.tx), cdiv(L.n1, tl.ty), nbatch);
    P.xcd_remap = g_xcd_remap && (grid.x * grid.y) % 8 == 0;
    P.bstride = N;
    auto step = [&](const T* x, T* y) {
        P.u = x;
        P.out = y;
        lt_begin(st);
        launch_step2d<T, EXACT>(stile, grid, P, st);
        lt_end(3, st);
    };
    if (pure) {
        // timing helper: nsteps (even) whole steps, the state ends in c
        T* x = c;
        T* y = w1;
        for (int s = 0; s < nsteps; ++s) {
            step(x, y);
            std::swap(x, y);
        }
        PHX_LAUNCH_CHECK("phx_ch_rk4_steps");
        return PHX_OK;
    }
    const bool even = nsteps % 2 == 0;
    T* X = even ? w1 : c;
    T* Y = even ? c : w1;
    if (even) PHX_HIP(hipMemcpyAsync(w1, c, (size_t)NB * sizeof(T), hipMemcpyDeviceToDevice, st));
    for (int s = 0; s < nsteps - 1; ++s) step(s % 2 == 0 ? X : Y, s % 2 == 0 ? Y : X);
    int rc = bc_fill_batch_impl(G, sizeof(T) == 8 ? PHX_F64 : PHX_F32, c, nbatch, st);
    if (rc) return rc;
    step(c, w0);
    if (nbatch == 1) {
        const size_t pitch = (size_t)L.sx * sizeof(T);
        const long long o0 = L.at(0, 0, 0);
        PHX_HIP(hipMemcpy2DAsync(c + o0, pitch, w0 + o0, pitch, (size_t)L.n0 * sizeof(T), L.n1,
                                 hipMemcpyDeviceToDevice, st));
    } else {
        const dim3 cg((unsigned)cdiv((long long)L.n0 * L.n1, 256), (unsigned)nbatch);
        hipLaunchKernelGGL(k_copy_phys2d<T>, cg, dim3(256), 0, st, c, (const T*)w0, L, N);
    }
    PHX_LAUNCH_CHECK("phx_ch_rk4_steps");
    return PHX_OK;
}

static int check_ch_params(const phx_ch_params* p, const char* who) {
    if (!p) return fail(PHX_ERR_INVALID, std::string(who) + ": params is NULL");
    if (!(p->dt > 0.0)) return fail(PHX_ERR_INVALID, std::string(who) + ": dt must be > 0");
    return PHX_OK;
}

}  // namespace phx

using namespace phx;

extern "C" {

int phx_ch_euler_set_fused(int on) {
    g_euler_fused = on ? 1 : 0;
    return PHX_OK;
}

int phx_ch_euler_set_quad(int tile) {
    if (tile < -1 || tile >= kNumTilesE4) return fail(PHX_ERR_INVALID, "phx_ch_euler_set_quad: tile must be -1..3");
    g_euler4_tile = tile;
    return PHX_OK;
}

int phx_ch_euler_steps(const phx_grid* G, int dtype, void* c, void* mu, const phx_ch_params* p, int nsteps,
                       phx_stream s) {
    int rc = check_grid_fused(G, "phx_ch_euler_steps");
    if (rc) return rc;
    if ((rc = check_dtype(dtype, "phx_ch_euler_steps"))) return rc;
    if ((rc = check_ch_params(p, "phx_ch_euler_steps"))) return rc;
    if (!c || !mu) return fail(PHX_ERR_INVALID, "phx_ch_euler_steps: NULL field");
    if (c == mu) return fail(PHX_ERR_INVALID, "phx_ch_euler_steps: c and mu must be distinct buffers");
    if (nsteps < 0) return fail(PHX_ERR_INVALID, "phx_ch_euler_steps: nsteps < 0");
    if (nsteps == 0) return PHX_OK;
    if (G->dim >= 3 && G->n[2] + 2 * G->ghost > 65535)
        return fail(PHX_ERR_UNSUPPORTED, "phx_ch_euler_steps: nz too large");
    return dtype == PHX_F64 ? euler_steps_t<double>(G, (double*)c, (double*)mu, p, nsteps, as_stream(s))
                            : euler_steps_t<float>(G, (float*)c, (float*)mu, p, nsteps, as_stream(s));
}

size_t phx_ch_rk4_workspace_bytes(const phx_grid* G, int dtype) {
    if (check_grid(G, "phx_ch_rk4_workspace_bytes") || check_dtype(dtype, "phx_ch_rk4_workspace_bytes")) return 0;
    return 3 * (size_t)make_layout(G).total() * (dtype == PHX_F64 ? 8 : 4);
}

// Launch of the RK4 kernels per (dim, shape, dtype): tile index (3D: one-stage
// 0-4, two-stage 5-8, one-stage with z-neighbours in registers 9-17; 2D:
// one-stage 0, whole-step 1-4) and, in 3D, the z
// planes streamed per workgroup. Set by phx_ch_rk4_set_tile/_set_zchunk or the
// autotuner; defaults below.
struct Launch3 {
    int zc, tile;          // stages 2 and 3; the schedule (two-stage, whole step, 2D) follows this tile
    int zc1, tile1;        // stage 1 (3D one-stage schedules only)
    int zc4, tile4;        // stage 4 (3D one-stage schedules only)
};
static Launch3 uniform_launch(int zc, int tile) { return Launch3{zc, tile, zc, tile, zc, tile}; }
// one-stage 3D tiles (LDS ring or z-register): the ones a stage class may mix
static bool one_stage_tile(int tile) { return tile < kNumTiles3 || tile >= kFirstZ; }
static std::mutex g_zc_mu;
static std::map<std::tuple<int, int, int, int, int>, Launch3> g_zc;

static Launch3 launch_for(const phx_grid* G, int dtype) {
    std::lock_guard<std::mutex> lk(g_zc_mu);
    auto it = g_zc.find(std::make_tuple(G->dim, G->n[0], G->n[1], G->n[2], dtype));
    if (it != g_zc.end()) return it->second;
    // defaults: the launches measured fastest at 1024^2 and 256^3 on MI355X
    // (2D: whole-step 128x32 per 1024 threads in FP32 on grids of 256^2 and
    // more, 32x32 per 512 threads otherwise and in FP64; 3D: z-register
    // kernel per stage class, below); every launch returns the same bits
    if (G->dim == 2) {
        if (!step2d_ok(G)) return uniform_launch(1, 0);
        return uniform_launch(1, dtype == PHX_F32 && G->n[0] >= 256 && G->n[1] >= 256 ? 7 : 5);
    }
    if (G->dim != 3) return uniform_launch(1, 0);
    Launch3 l;
    if (dtype == PHX_F32) {
        // FP32, per stage class (256^3 measurement): stages 1-3 128x16/512 at
        // 24 z-chunks, stage 4 256x8/512 at 16
        l.tile1 = l.tile = kFirstZ + 2;
        l.zc1 = l.zc = (int)cdiv(G->n[2], 24);
        l.tile4 = kFirstZ + 3;
        l.zc4 = (int)cdiv(G->n[2], 16);
        return l;
    }
    // FP64, per stage class (256^3 measurement, loads two planes ahead):
    // stage 1 128x16/512 at 8 chunks, stages 2-3 128x8/512 at 4 chunks,
    // stage 4 128x8/512 at 8 chunks
    l.tile1 = kFirstZ + 2;
    l.zc1 = (int)cdiv(G->n[2], 8);
    l.tile = kFirstZ + 8;
    l.zc = (int)cdiv(G->n[2], 4);
    l.tile4 = kFirstZ + 8;
    l.zc4 = (int)cdiv(G->n[2], 8);
    return l;
}

static void set_launch(const phx_grid* G, int dtype, Launch3 l) {
    std::lock_guard<std::mutex> lk(g_zc_mu);
    g_zc[std::make_tuple(G->dim, G->n[0], G->n[1], G->n[2], dtype)] = l;
}

int phx_ch_set_xcd_remap(int on) {
    g_xcd_remap = on ? 1 : 0;
    return PHX_OK;
}

int phx_ch_rk4_set_zchunk(const phx_grid* G, int dtype, int zc) {
    int rc = check_grid(G, "phx_ch_rk4_set_zchunk");
    if (rc) return rc;
    if (zc < 1 || zc > 4096) return fail(PHX_ERR_INVALID, "phx_ch_rk4_set_zchunk: zc must be in [1, 4096]");
    Launch3 l = launch_for(G, dtype);
    l.zc = l.zc1 = l.zc4 = zc;
    set_launch(G, dtype, l);
    return PHX_OK;
}

int phx_ch_rk4_get_launch(const phx_grid* G, int dtype, int* zc, int* tile) {
    int rc = check_grid(G, "phx_ch_rk4_get_launch");
    if (rc) return rc;
    Launch3 l = launch_for(G, dtype);
    if (zc) *zc = l.zc;
    if (tile) *tile = l.tile;
    return PHX_OK;
}

int phx_ch_rk4_set_tile(const phx_grid* G, int dtype, int tile) {
    int rc = check_grid(G, "phx_ch_rk4_set_tile");
    if (rc) return rc;
    const int ntiles = G->dim == 3 ? kFirstZ + kNumTilesZ : (G->dim == 2 ? 1 + kNumTilesS2 : 1);
    if (tile < 0 || tile >= ntiles) return fail(PHX_ERR_INVALID, "phx_ch_rk4_set_tile: tile index out of range");
    if (G->dim == 3 && is_pair_tile(tile) && !pair_ok(G))
        return fail(PHX_ERR_UNSUPPORTED, "phx_ch_rk4_set_tile: two-stage tiles need n >= 4 on every axis");
    if (G->dim == 3 && tile >= kFirstZ && !zreg_ok(G, dtype))
        return fail(PHX_ERR_UNSUPPORTED, "phx_ch_rk4_set_tile: z-register tiles need a stored plane below 4 GiB");
    if (G->dim == 2 && tile >= 1 && !step2d_ok(G))
        return fail(PHX_ERR_UNSUPPORTED, "phx_ch_rk4_set_tile: whole-step 2D tiles need n >= 8 on both axes");
    Launch3 l = launch_for(G, dtype);
    l.tile = l.tile1 = l.tile4 = tile;
    set_launch(G, dtype, l);
    return PHX_OK;
}

}  // extern "C"

// The schedule of one FP64 form (X64: exact) and of FP32.
template <bool X64>
static int run_rk4_t(const phx_grid* G, int dtype, void* c, void* work, const phx_ch_params* p, int nsteps,
                     hipStream_t st, Launch3 l, int pure) {
    if (G->dim == 2 && l.tile >= 1 && step2d_ok(G))
        return dtype == PHX_F64 ? rk4_step2d_steps_t<double, X64>(G, (double*)c, (double*)work, p, nsteps, st,
                                                                  l.tile - 1, pure)
                                : rk4_step2d_steps_t<float, false>(G, (float*)c, (float*)work, p, nsteps, st,
                                                                   l.tile - 1, pure);
    if (G->dim != 3) l.tile = 0;
    if (G->dim == 3 && l.tile >= kFirstZ && !zreg_ok(G, dtype)) l.tile = dtype == PHX_F32 ? 0 : 1;
    if (G->dim == 3 && is_pair_tile(l.tile)) {
        if (!pair_ok(G)) l.tile = dtype == PHX_F32 ? 0 : 1;
        else
            return dtype == PHX_F64 ? rk4_pair_steps_t<double, X64>(G, (double*)c, (double*)work, p, nsteps, st,
                                                                    l.zc, l.tile - kNumTiles3, pure)
                                    : rk4_pair_steps_t<float, false>(G, (float*)c, (float*)work, p, nsteps, st,
                                                                     l.zc, l.tile - kNumTiles3, pure);
    }
    int z3[3] = {l.zc1, l.zc, l.zc4}, t3[3] = {l.tile1, l.tile, l.tile4};
    for (int k = 0; k < 3; ++k) {
        // a class keeps its own launch only in a 3D one-stage schedule
        if (G->dim != 3 || !one_stage_tile(t3[k]) || (t3[k] >= kFirstZ && !zreg_ok(G, dtype))) {
            z3[k] = l.zc;
            t3[k] = l.tile;
        }
    }
    return dtype == PHX_F64 ? rk4_steps_t<double, X64>(G, (double*)c, (double*)work, p, nsteps, st, z3, t3)
                            : rk4_steps_t<float, false>(G, (float*)c, (float*)work, p, nsteps, st, z3, t3);
}

// Schedule by the launch's tile index; `pure` runs nsteps (even) two-stage or
// whole-step steps only (autotuner timing). FP64 takes the form g_fp64_exact
// selects.
static int run_rk4(const phx_grid* G, int dtype, void* c, void* work, const phx_ch_params* p, int nsteps,
                   hipStream_t st, Launch3 l, int pure) {
    return g_fp64_exact ? run_rk4_t<true>(G, dtype, c, work, p, nsteps, st, l, pure)
                        : run_rk4_t<false>(G, dtype, c, work, p, nsteps, st, l, pure);
}

extern "C" {

int phx_ch_rk4_steps(const phx_grid* G, int dtype, void* c, void* work, const phx_ch_params* p, int nsteps,
                     phx_stream s) {
    int rc = check_grid_fused(G, "phx_ch_rk4_steps");
    if (rc) return rc;
    if ((rc = check_dtype(dtype, "phx_ch_rk4_steps"))) return rc;
    if ((rc = check_ch_params(p, "phx_ch_rk4_steps"))) return rc;
    if (!c || !work) return fail(PHX_ERR_INVALID, "phx_ch_rk4_steps: NULL buffer");
    if (nsteps < 0) return fail(PHX_ERR_INVALID, "phx_ch_rk4_steps: nsteps < 0");
    if (nsteps == 0) return PHX_OK;
    return run_rk4(G, dtype, c, work, p, nsteps, as_stream(s), launch_for(G, dtype), 0);
}

// ------------------------------------------- RK4 without allocation ----
// phx_ch_rk4_steps_noalloc runs phx_ch_rk4_steps with its three-field
// workspace in this file-scope device array instead of a caller buffer. Each
// GPU holds its own copy (PHX_CH_RK4_NOALLOC_BYTES) from the library's load
// until the process ends. Host code must not pass the array's name to a
// kernel (on the host it is a stand-in, not a device address): its address
// on the current device comes from hipGetSymbolAddress, taken once per device
// on the first call there and kept in g_scratch_dev.
static __device__ __align__(256) double scratch[PHX_CH_RK4_NOALLOC_BYTES / sizeof(double)];

constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static void* g_scratch_dev[kMaxDevices];

static int scratch_on_current_device(void** out) {
    int dev = 0;
    PHX_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices)
        return fail(PHX_ERR_UNSUPPORTED, "phx_ch_rk4_steps_noalloc: device index beyond 63");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_dev[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(PHX_ERR_HIP, std::string("phx_ch_rk4_steps_noalloc: hipGetSymbolAddress: ") +
                                         hipGetErrorString(e));
        g_scratch_dev[dev] = p;
    }
    *out = g_scratch_dev[dev];
    return PHX_OK;
}

extern "C" int phx_ch_rk4_steps_noalloc(const phx_grid* G, int dtype, void* c, const phx_ch_params* p,
                                        int nsteps, phx_stream s) {
    int rc = check_grid_fused(G, "phx_ch_rk4_steps_noalloc");
    if (rc) return rc;
    if ((rc = check_dtype(dtype, "phx_ch_rk4_steps_noalloc"))) return rc;
    if ((rc = check_ch_params(p, "phx_ch_rk4_steps_noalloc"))) return rc;
    if (!c) return fail(PHX_ERR_INVALID, "phx_ch_rk4_steps_noalloc: NULL field");
    if (nsteps < 0) return fail(PHX_ERR_INVALID, "phx_ch_rk4_steps_noalloc: nsteps < 0");
    const size_t need = 3 * (size_t)make_layout(G).total() * (dtype == PHX_F64 ? 8 : 4);
    if (need > (size_t)PHX_CH_RK4_NOALLOC_BYTES)
        return fail(PHX_ERR_UNSUPPORTED, "phx_ch_rk4_steps_noalloc: the RK4 workspace of this grid needs " +
                                             std::to_string(need) + " bytes, more than the " +
                                             std::to_string((size_t)PHX_CH_RK4_NOALLOC_BYTES) +
                                             " the library holds per GPU; use phx_ch_rk4_steps");
    if (nsteps == 0) return PHX_OK;
    void* work = nullptr;
    if ((rc = scratch_on_current_device(&work))) return rc;
    // The stage kernels write every physical cell of a workspace field
    // before reading it and never read its ghost cells, so the array's
    // contents from earlier calls need no clearing.
    t_timer_off = 1;
    rc = run_rk4(G, dtype, c, work, p, nsteps, as_stream(s), launch_for(G, dtype), 0);
    t_timer_off = 0;
    return rc;
}

// nbatch independent fields of one 2D grid, stored back to back, advanced
// together: the whole-step schedule with the batch on the launch's z axis.
// Whole-step tile of phx_ch_rk4_steps_batch per (shape, dtype, nbatch), set
// by phx_ch_rk4_batch_autotune: a batch fills the chip, so the tile that is
// fastest for it can differ from the single field's (which is latency bound
// on small grids).
static std::map<std::tuple<int, int, int, int, int, int>, int> g_batch_tile;

static int batch_tile(const phx_grid* G, int dtype, int nbatch) {
    std::lock_guard<std::mutex> lk(g_zc_mu);
    auto it = g_batch_tile.find(std::make_tuple(G->dim, G->n[0], G->n[1], G->n[2], dtype, nbatch));
    return it == g_batch_tile.end() ? -1 : it->second;
}

int phx_ch_rk4_steps_batch(const phx_grid* G, int dtype, int nbatch, void* c, void* work, const phx_ch_params* p,
                           int nsteps, phx_stream s) {
    int rc = check_grid_fused(G, "phx_ch_rk4_steps_batch");
    if (rc) return rc;
    if ((rc = check_dtype(dtype, "phx_ch_rk4_steps_batch"))) return rc;
    if ((rc = check_ch_params(p, "phx_ch_rk4_steps_batch"))) return rc;
    if (!c || !work) return fail(PHX_ERR_INVALID, "phx_ch_rk4_steps_batch: NULL buffer");
    if (nsteps < 0) return fail(PHX_ERR_INVALID, "phx_ch_rk4_steps_batch: nsteps < 0");
    if (nbatch < 1 || nbatch > 65535) return fail(PHX_ERR_INVALID, "phx_ch_rk4_steps_batch: nbatch must be in [1, 65535]");
    if (!step2d_ok(G))
        return fail(PHX_ERR_UNSUPPORTED, "phx_ch_rk4_steps_batch: 2D grids with n >= 8 on both axes only");
    if (nsteps == 0) return PHX_OK;
    // the batch's own launch when phx_ch_rk4_batch_autotune chose one, else
    // the single-field launch of this shape
    int stile = batch_tile(G, dtype, nbatch);
    if (stile < 0) {
        Launch3 l = launch_for(G, dtype);
        stile = l.tile >= 1 ? l.tile - 1 : step2d_tile(4, dtype == PHX_F64 ? 8 : 4);
    }
    hipStream_t st = as_stream(s);
    if (dtype == PHX_F32)
        return rk4_step2d_steps_t<float, false>(G, (float*)c, (float*)work, p, nsteps, st, stile, 0, nbatch);
    return g_fp64_exact ? rk4_step2d_steps_t<double, true>(G, (double*)c, (double*)work, p, nsteps, st, stile, 0, nbatch)
                        : rk4_step2d_steps_t<double, false>(G, (double*)c, (double*)work, p, nsteps, st, stile, 0, nbatch);
}

// Times the whole-step tiles on `steps` steps of the caller's batch (advancing
// it) and keeps the fastest for phx_ch_rk4_steps_batch with this shape,
// dtype and nbatch.
int phx_ch_rk4_batch_autotune(const phx_grid* G, int dtype, int nbatch, void* c, void* work, const phx_ch_params* p,
                              int steps, double* best_ms_per_step, phx_stream s) {
    int rc = check_grid_fused(G, "phx_ch_rk4_batch_autotune");
    if (rc) return rc;
    if ((rc = check_dtype(dtype, "phx_ch_rk4_batch_autotune"))) return rc;
    if ((rc = check_ch_params(p, "phx_ch_rk4_batch_autotune"))) return rc;
    if (!c || !work) return fail(PHX_ERR_INVALID, "phx_ch_rk4_batch_autotune: NULL buffer");
    if (steps < 1) return fail(PHX_ERR_INVALID, "phx_ch_rk4_batch_autotune: steps < 1");
    if (nbatch < 1 || nbatch > 65535)
        return fail(PHX_ERR_INVALID, "phx_ch_rk4_batch_autotune: nbatch must be in [1, 65535]");
    if (!step2d_ok(G))
        return fail(PHX_ERR_UNSUPPORTED, "phx_ch_rk4_batch_autotune: 2D grids with n >= 8 on both axes only");
    hipStream_t st = as_stream(s);
    auto run = [&](int stile, int n) -> int {
        if (dtype == PHX_F32) return rk4_step2d_steps_t<float, false>(G, (float*)c, (float*)work, p, n, st, stile, 0, nbatch);
        return g_fp64_exact ? rk4_step2d_steps_t<double, true>(G, (double*)c, (double*)work, p, n, st, stile, 0, nbatch)
                            : rk4_step2d_steps_t<double, false>(G, (double*)c, (double*)work, p, n, st, stile, 0, nbatch);
    };
    hipEvent_t e0, e1;
    PHX_HIP(hipEventCreate(&e0));
    PHX_HIP(hipEventCreate(&e1));
    float best = 1e30f;
    int bt = 0;
    for (int stile = 0; stile < kNumTilesS2; ++stile) {
        if (step2d_tile(stile, dtype == PHX_F64 ? 8 : 4) != stile) continue;   // runs as another tile
        if ((rc = run(stile, 2))) return rc;
        PHX_HIP(hipEventRecord(e0, st));
        if ((rc = run(stile, steps))) return rc;
        PHX_HIP(hipEventRecord(e1, st));
        PHX_HIP(hipEventSynchronize(e1));
        float ms = 0.f;
        PHX_HIP(hipEventElapsedTime(&ms, e0, e1));
        ms /= steps;
        if (ms < best) {
            best = ms;
            bt = stile;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    {
        std::lock_guard<std::mutex> lk(g_zc_mu);
        g_batch_tile[std::make_tuple(G->dim, G->n[0], G->n[1], G->n[2], dtype, nbatch)] = bt;
    }
    if (best_ms_per_step) *best_ms_per_step = best;
    return PHX_OK;
}

// Times every launch candidate on `steps` steps of the caller's state
// (advancing it) and keeps the fastest per step for this shape: in 3D the
// one-stage kernel (5 tiles), the two-stage kernel (4 tiles) and the
// z-register kernel (9 tiles) at z-chunks {4, 8, 16, 32, 64} plus the chunk
// counts that fill whole waves of resident workgroups; in 2D the one-stage
// kernel and the whole-step kernels. Two-stage and whole-step candidates run
// 2*ceil(steps/2) steps. When a 3D one-stage launch wins, each stage class
// (stage 1, stages 2-3, stage 4) is then refined in turn over the one-stage
// candidates with the other classes held.
int phx_ch_rk4_autotune(const phx_grid* G, int dtype, void* c, void* work, const phx_ch_params* p, int steps,
                        int* best_zc, double* best_ms_per_step, phx_stream s) {
    int rc = check_grid_fused(G, "phx_ch_rk4_autotune");
    if (rc) return rc;
    if (steps < 1) return fail(PHX_ERR_INVALID, "phx_ch_rk4_autotune: steps < 1");
    hipStream_t st = as_stream(s);
    const int cand[] = {4, 8, 16, 32, 64};
    const int ntiles = G->dim == 3 ? kFirstZ + kNumTilesZ : (G->dim == 2 && step2d_ok(G) ? 1 + kNumTilesS2 : 1);
    const int psteps = 2 * ((steps + 1) / 2);
    int dev = 0, ncu = 0;
    PHX_HIP(hipGetDevice(&dev));
    PHX_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
    // candidates (zc, tile)
    std::vector<std::pair<int, int>> cands;
    for (int tile = 0; tile < ntiles; ++tile) {
        if (G->dim == 3 && is_pair_tile(tile) && !pair_ok(G)) continue;
        if (G->dim == 3 && tile >= kFirstZ && !zreg_ok(G, dtype)) continue;
        if (G->dim != 3) {
            cands.emplace_back(1, tile);
            continue;
        }
        // z-chunks: fixed sizes, plus chunk counts that fill whole waves of
        // the resident workgroup slots (no partly filled last wave)
        std::vector<int> zcs(cand, cand + sizeof(cand) / sizeof(cand[0]));
        const Tile3 tl = shape3(tile);
        const long long ncols = (long long)cdiv(G->n[0], tl.tx) * cdiv(G->n[1], tl.ty);
        const int bpc = dtype == PHX_F32 ? blocks_per_cu<float, false>(tile)
                        : (g_fp64_exact ? blocks_per_cu<double, true>(tile) : blocks_per_cu<double, false>(tile));
        for (int k = 1; k <= 3 && bpc > 0; ++k) {
            const long long nch = std::max(1LL, (long long)k * bpc * ncu / ncols);
            const int zc = (int)cdiv(G->n[2], (int)std::min<long long>(nch, G->n[2]));
            if (std::find(zcs.begin(), zcs.end(), zc) == zcs.end()) zcs.push_back(zc);
        }
        for (int zc : zcs)
            if (zc <= G->n[2]) cands.emplace_back(zc, tile);
    }
    hipEvent_t e0, e1;
    PHX_HIP(hipEventCreate(&e0));
    PHX_HIP(hipEventCreate(&e1));
    auto time_launch = [&](const Launch3& l, float& ms) -> int {
        const int pure = G->dim == 3 ? is_pair_tile(l.tile) : l.tile >= 1;
        const int n = pure ? psteps : steps;
        int r = run_rk4(G, dtype, c, work, p, pure ? 2 : 1, st, l, pure);
        if (r) return r;
        PHX_HIP(hipEventRecord(e0, st));
        if ((r = run_rk4(G, dtype, c, work, p, n, st, l, pure))) return r;
        PHX_HIP(hipEventRecord(e1, st));
        PHX_HIP(hipEventSynchronize(e1));
        PHX_HIP(hipEventElapsedTime(&ms, e0, e1));
        ms /= n;
        return PHX_OK;
    };
    float best = 1e30f;
    Launch3 bl = uniform_launch(16, 0);
    for (const auto& cz : cands) {
        const Launch3 l = uniform_launch(cz.first, cz.second);
        float ms = 0.f;
        if ((rc = time_launch(l, ms))) return rc;
        if (ms < best) {
            best = ms;
            bl = l;
        }
    }
    if (G->dim == 3 && one_stage_tile(bl.tile)) {
        for (int cls = 0; cls < 3; ++cls) {
            for (const auto& cz : cands) {
                if (!one_stage_tile(cz.second)) continue;
                Launch3 l = bl;
                int& zc = cls == 0 ? l.zc1 : (cls == 1 ? l.zc : l.zc4);
                int& tile = cls == 0 ? l.tile1 : (cls == 1 ? l.tile : l.tile4);
                if (zc == cz.first && tile == cz.second) continue;
                zc = cz.first;
                tile = cz.second;
                float ms = 0.f;
                if ((rc = time_launch(l, ms))) return rc;
                if (ms < best) {
                    best = ms;
                    bl = l;
                }
            }
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    set_launch(G, dtype, bl);
    if (best_zc) *best_zc = bl.zc;
    if (best_ms_per_step) *best_ms_per_step = best;
    return PHX_OK;
}

// Launch of one stage class (0: stage 1, 1: stages 2 and 3, 2: stage 4) of
// the 3D one-stage schedule.
int phx_ch_rk4_set_stage_launch(const phx_grid* G, int dtype, int cls, int tile, int zc) {
    int rc = check_grid(G, "phx_ch_rk4_set_stage_launch");
    if (rc) return rc;
    if (cls < 0 || cls > 2) return fail(PHX_ERR_INVALID, "phx_ch_rk4_set_stage_launch: class must be 0, 1 or 2");
    if (G->dim != 3) return fail(PHX_ERR_UNSUPPORTED, "phx_ch_rk4_set_stage_launch: 3D grids only");
    if (tile < 0 || tile >= kFirstZ + kNumTilesZ || !one_stage_tile(tile))
        return fail(PHX_ERR_INVALID, "phx_ch_rk4_set_stage_launch: not a one-stage 3D tile");
    if (tile >= kFirstZ && !zreg_ok(G, dtype))
        return fail(PHX_ERR_UNSUPPORTED, "phx_ch_rk4_set_stage_launch: z-register tiles need a stored plane below 4 GiB");
    if (zc < 1 || zc > 4096) return fail(PHX_ERR_INVALID, "phx_ch_rk4_set_stage_launch: zc must be in [1, 4096]");
    Launch3 l = launch_for(G, dtype);
    (cls == 0 ? l.zc1 : (cls == 1 ? l.zc : l.zc4)) = zc;
    (cls == 0 ? l.tile1 : (cls == 1 ? l.tile : l.tile4)) = tile;
    set_launch(G, dtype, l);
    return PHX_OK;
}

int phx_ch_rk4_get_stage_launch(const phx_grid* G, int dtype, int cls, int* zc, int* tile) {
    int rc = check_grid(G, "phx_ch_rk4_get_stage_launch");
    if (rc) return rc;
    if (cls < 0 || cls > 2) return fail(PHX_ERR_INVALID, "phx_ch_rk4_get_stage_launch: class must be 0, 1 or 2");
    const Launch3 l = launch_for(G, dtype);
    if (zc) *zc = cls == 0 ? l.zc1 : (cls == 1 ? l.zc : l.zc4);
    if (tile) *tile = cls == 0 ? l.tile1 : (cls == 1 ? l.tile : l.tile4);
    return PHX_OK;
}

int phx_ch_set_fp64_exact(int on) {
    g_fp64_exact = on ? 1 : 0;
    return PHX_OK;
}

int phx_ch_get_fp64_exact(void) { return g_fp64_exact; }

int phx_ch_timing(int on) {
    g_lt.on = on ? 1 : 0;
    g_lt.used = 0;
    return PHX_OK;
}

int phx_ch_timing_read(int kind, int* launches, double* total_ms) {
    if (kind < 0 || kind > 10) return fail(PHX_ERR_INVALID, "phx_ch_timing_read: kind must be in [0, 10]");
    int n = 0;
    double ms = 0.0;
    for (size_t i = 0; i < g_lt.used; ++i) {
        if (g_lt.kind[i] != kind) continue;
        PHX_HIP(hipEventSynchronize(g_lt.ev[2 * i + 1]));
        float t = 0.f;
        PHX_HIP(hipEventElapsedTime(&t, g_lt.ev[2 * i], g_lt.ev[2 * i + 1]));
        ms += t;
        ++n;
    }
    if (launches) *launches = n;
    if (total_ms) *total_ms = ms;
    return PHX_OK;
}

}  // extern "C"
