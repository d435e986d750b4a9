// ac_fused.hip -- fused multi-order-parameter Allen-Cahn RK4 stage kernel.
//
// The reference runs a VectorEquation as one Equation per component, each a
// memset plus one launch per Term and a device sync (src/equation/
// VectorEquation.cu:33-40), with pw() limited to three fields
// (include/equation/TermPW.inl:212-268), so an N-field coupling costs ~N^2
// launches per stage. Here one kernel evaluates every component's RHS for a
// tile and applies the RK4 stage epilogue to all of them:
//   all fields' halo tiles -> LDS behind one barrier (their loads overlap);
//   then per field: rhs_i = L*kappa*lap_i + L*(e_i - e_i^3)
//                           + sum_{j!=i} (-2*L*gamma)*(e_i*e_j*e_j)
//                   in that order (the order of the VectorEquation terms),
//                   and the stage update.
// That order is kept for FP64, which reproduces the FP64 oracle bit for bit.
// FP32 (held to a stated error against the oracle) evaluates the same
// expression with explicit FMAs and the coupling factored as
//   (-2*L*gamma) * e_i * sum_{j!=i} e_j^2,
// the sum over j != i formed from prefix and suffix sums of the squares (no
// cancellation: every term is non-negative), so a field costs ~13 operations
// instead of ~45 at 8 fields.
// BCs are index maps of the halo load (stage states are BC-filled in the
// reference, src/solver/VectorSolver.cu:149-179).
#pragma clang fp contract(off)

#include "common.hpp"

namespace phx {

constexpr int AC_MAXF = 8;

template <typename T>
struct ACArgs {
    const T* u[AC_MAXF];
    const T* c[AC_MAXF];
    T* acc[AC_MAXF];
    T* out[AC_MAXF];
    Layout L;
    AxisBC<T> bx, by;
    T idx2, idy2;
    T cl, cp, cg;   // L*kappa, L, -2*L*gamma
    T icc;          // -2 (idx2 + idy2): FP32 form of the Laplacian
    T a, dt6, dt3;
    T a2;            // two-stage kernel: shift of its second stage
    int stage;
    int xcd_remap;   // 1: XCD-aware tile order (grid size a multiple of 8)
};

// RHS of every field at one point of a block in LDS (field f at S + f*fs,
// row pitch P): the stage state e and the RHS k, formed exactly as described
// at the top of this file. Both kernels evaluate points only through this
// helper, so the one- and two-stage forms return the same bits.
template <typename T, int NF, int DIM, bool EXACT>
__device__ __forceinline__ void ac_rhs(const ACArgs<T>& A, const T* S, int fs, int c0, int P, T (&e)[NF],
                                       T (&k)[NF]) {
    T lp[NF];
#pragma unroll
    for (int f = 0; f < NF; ++f) {
        const T* U = S + f * fs;
        const T cc = U[c0];
        T val;
        if constexpr (EXACT) {
            val = (U[c0 + 1] - (T)2 * cc + U[c0 - 1]) * A.idx2;
            if (DIM >= 2) val += (U[c0 + P] - (T)2 * cc + U[c0 - P]) * A.idy2;
        } else {
            val = __builtin_fmaf(U[c0 + 1] + U[c0 - 1], A.idx2, cc * A.icc);
            if (DIM >= 2) val = __builtin_fmaf(U[c0 + P] + U[c0 - P], A.idy2, val);
        }
        e[f] = cc;
        lp[f] = val;
    }
    // FP32: squares, and their prefix / suffix sums (sum over j != i = pre[i] + suf[i])
    T e2[NF], pre[NF], suf[NF];
    if constexpr (!EXACT) {
#pragma unroll
        for (int f = 0; f < NF; ++f) e2[f] = e[f] * e[f];
        pre[0] = (T)0;
#pragma unroll
        for (int f = 1; f < NF; ++f) pre[f] = pre[f - 1] + e2[f - 1];
        suf[NF - 1] = (T)0;
#pragma unroll
        for (int f = NF - 2; f >= 0; --f) suf[f] = suf[f + 1] + e2[f + 1];
    }
#pragma unroll
    for (int f = 0; f < NF; ++f) {
        const T ef = e[f];
        T r;
        if constexpr (EXACT) {
            r = A.cl * lp[f];
            r = r + A.cp * (ef - ef * ef * ef);
#pragma unroll
            for (int h = 0; h < NF; ++h) {
                if (h == f) continue;
                r = r + A.cg * (ef * e[h] * e[h]);
            }
        } else {
            r = A.cl * lp[f];
            r = __builtin_fmaf(A.cp, __builtin_fmaf(-e2[f], ef, ef), r);
            if (NF > 1) r = __builtin_fmaf(A.cg * ef, pre[f] + suf[f], r);
        }
        k[f] = r;
    }
}

template <typename T, int DIM, int NF, int TX, int TY, bool EXACT>
__global__ __launch_bounds__(256) void k_ac_stage_2d(ACArgs<T> A) {
    constexpr int UX = TX + 2, UY = TY + 2;
    constexpr int R = (TX * TY) / 256;   // outputs per thread
    constexpr int NS = (UX * UY + 255) / 256;
    __shared__ T U[NF][UY * UX];
    const int tid = threadIdx.x;
    // XCD-aware tile order: consecutive workgroups go to the 8 XCDs in turn;
    // giving each XCD a contiguous range of tiles (row bands) makes the halo
    // rows shared by neighbouring tiles hit that XCD's own L2 instead of being
    // fetched once per XCD (measured: 1.28x the stage throughput at 1024^2 x 8)
    int bx = blockIdx.x, by = blockIdx.y;
    if (A.xcd_remap) {
        const int gx = gridDim.x, nb = gx * gridDim.y;
        const int h = bx + gx * by;
        const int lb = (h & 7) * (nb >> 3) + (h >> 3);
        bx = lb % gx;
        by = lb / gx;
    }
    const int x0 = bx * TX, y0 = by * TY;

    // Every load of the block -- all fields' halo tiles (offsets with the BC
    // maps computed first) and the stage epilogue's c_n and accumulator
    // values -- is issued unconditionally, from valid cells, before the single
    // barrier, so one global latency covers them all. FIXED halo slots and
    // cells beyond the domain are replaced or skipped afterwards.
    const int g = A.L.g;
    const long long zb = (long long)g * A.L.sxy;
    const int cell0 = g * A.L.sx + g;
    int off[NS];
    unsigned fm = 0;
    T fv[NS];
#pragma unroll
    for (int i = 0; i < NS; ++i) {
        const int t = min(tid + 256 * i, UX * UY - 1);
        const int ly = t / UX, lx = t - ly * UX;
        const int px = x0 - 1 + lx, py = DIM >= 2 ? y0 - 1 + ly : 0;
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
            off[i] = (qy + g) * A.L.sx + (qx + g);
        }
    }
    T hr[NF][NS];
#pragma unroll
    for (int f = 0; f < NF; ++f)
#pragma unroll
        for (int i = 0; i < NS; ++i) hr[f][i] = A.u[f][zb + off[i]];
    // epilogue operands; one the stage does not need is read at element 0 (below)
    const int st = A.stage;
    const bool need_c = st == 2 || st == 3 || (EXACT && st == 4);
    const bool need_a = st >= 2;
    // an operand the stage does not need: every lane reads element 0 (one line)
    const long long cmask = need_c ? -1LL : 0LL, amask = need_a ? -1LL : 0LL;
    T pc[NF][R], pa[NF][R];
    long long ogi[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int t = tid + r * 256;
        const int oy = t / TX, ox = t - oy * TX;
        const int x = min(x0 + ox, A.L.n0 - 1), y = min(y0 + oy, A.L.n1 - 1);
        ogi[r] = zb + (long long)(y + g) * A.L.sx + (x + g);
    }
#pragma unroll
    for (int f = 0; f < NF; ++f) {
        const T* cs = need_c ? A.c[f] : A.u[f];
        const T* as = need_a ? A.acc[f] : A.u[f];
#pragma unroll
        for (int r = 0; r < R; ++r) {
            pc[f][r] = cs[ogi[r] & cmask];
            pa[f][r] = as[ogi[r] & amask];
        }
    }
#pragma unroll
    for (int f = 0; f < NF; ++f)
#pragma unroll
        for (int i = 0; i < NS; ++i) {
            const int t = tid + 256 * i;
            if (t < UX * UY) U[f][t] = ((fm >> i) & 1u) ? fv[i] : hr[f][i];
        }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int t = tid + r * 256;
        const int oy = t / TX, ox = t - oy * TX;
        const int x = x0 + ox, y = y0 + oy;
        if (x >= A.L.n0 || y >= A.L.n1) continue;
        const long long gi = A.L.at(x, y, 0);
        const int c0 = (oy + 1) * UX + ox + 1;
        T e[NF], kk[NF];
        ac_rhs<T, NF, DIM, EXACT>(A, &U[0][0], UY * UX, c0, UX, e, kk);
#pragma unroll
        for (int f = 0; f < NF; ++f) {
            const T ef = e[f];
            const T k = kk[f];
            const int s = A.stage;
            if constexpr (EXACT) {
                if (s == 1) {
                    A.acc[f][gi] = k;
                    A.out[f][gi] = ef + A.a * k;
                } else if (s < 4) {
                    A.acc[f][gi] = pa[f][r] + (T)2 * k;
                    A.out[f][gi] = pc[f][r] + A.a * k;
                } else {
                    A.out[f][gi] = pc[f][r] + A.dt6 * (pa[f][r] + k);
                }
            } else {
                if (s == 1) {
                    A.acc[f][gi] = __builtin_fmaf(A.dt6, k, ef);
                    A.out[f][gi] = __builtin_fmaf(A.a, k, ef);
                } else if (s < 4) {
                    A.acc[f][gi] = __builtin_fmaf(A.dt3, k, pa[f][r]);
                    A.out[f][gi] = __builtin_fmaf(A.a, k, pc[f][r]);
                } else {
                    A.out[f][gi] = __builtin_fmaf(A.dt6, k, pa[f][r]);
                }
            }
        }
    }
}

// Two RK4 stages per launch (temporal blocking), 2D: stages (1,2)
// (SECOND = false) or (3,4) (SECOND = true). Per step, per field:
//   (1,2): read c_n;                          write acc and the stage-3 state;
//   (3,4): read the stage-3 state, c_n, acc;  write c_{n+1} (to a buffer other
//          than c_n: the launch reads c_n around its tile).
// 7 field passes per step instead of the one-stage kernel's 16. The first
// stage is evaluated on the tile plus a 1-cell ring (its input, radius 2, in
// LDS), its result -- the second stage's input state -- goes to LDS, and the
// second stage is evaluated on the tile.
// Ring positions outside the domain take the value the second stage's input
// has there after the BC fill, as in the one-stage kernel's halo load:
//   FIXED -> the constant; NOFLUX -> the first stage evaluated at the mirror
//   image (the cell whose value the ghost copies); PERIODIC -> the first stage
//   evaluated at the position itself from the wrapped input, which repeats the
//   image's operands in the same order.
// Every point goes through ac_rhs and the epilogues below repeat the
// one-stage kernel's operations, so the results equal two one-stage launches
// bit for bit. Stage-A slots: the TX*TY tile cells first, in the order of the
// outputs (a thread keeps its cells' first-stage accumulator in registers),
// then the ring.
template <typename T, int NF, int TX, int TY, bool EXACT, bool SECOND>
__global__ __launch_bounds__(256) void k_ac_pair_2d(ACArgs<T> A) {
    constexpr int VX = TX + 4, VY = TY + 4;   // first-stage input, radius 2
    constexpr int WX = TX + 2, WY = TY + 2;   // second-stage input, radius 1
    constexpr int R = (TX * TY) / 256;
    static_assert((TX * TY) % 256 == 0, "tile must be a multiple of 256 cells");
    constexpr int NV = (VX * VY + 255) / 256;
    constexpr int NBORD = 2 * WX + 2 * TY;
    constexpr int NW = R + (NBORD + 255) / 256;
    __shared__ T V[NF][VY * VX];
    __shared__ T W[NF][WY * WX];
    const int tid = threadIdx.x;
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
    const long long zb = (long long)g * A.L.sxy;
    const int cell0 = g * A.L.sx + g;
    const int n0 = A.L.n0, n1 = A.L.n1;

    // first-stage input block: offsets with the BC maps, FIXED slots flagged
    int voff[NV];
    unsigned vfm = 0;
    T vfv[NV];
#pragma unroll
    for (int i = 0; i < NV; ++i) {
        const int t = min(tid + 256 * i, VX * VY - 1);
        const int ly = t / VX, lx = t - ly * VX;
        const int px = x0 - 2 + lx, py = y0 - 2 + ly;
        const int qy = map_in(py, A.by), qx = map_in(px, A.bx);
        voff[i] = cell0;
        vfv[i] = (T)0;
        if (qy < 0) {
            vfm |= 1u << i;
            vfv[i] = fixed_value(py, A.by);
        } else if (qx < 0) {
            vfm |= 1u << i;
            vfv[i] = fixed_value(px, A.bx);
        } else {
            voff[i] = (qy + g) * A.L.sx + (qx + g);
        }
    }
    // first-stage slots: LDS index in W, evaluation centre in V, FIXED value,
    // and (second pair) the c_n offset of the slot
    int wl[NW], vc[NW], coff[NW];
    unsigned wact = 0, wfix = 0;
    T wfv[NW];
#pragma unroll
    for (int s = 0; s < NW; ++s) {
        int wx, wy;
        bool act = true;
        if (s < R) {
            const int q = tid + 256 * s;
            wy = q / TX + 1;
            wx = q - (wy - 1) * TX + 1;
        } else {
            int q = tid + 256 * (s - R);
            act = q < NBORD;
            q = min(q, NBORD - 1);
            if (q < WX) { wx = q; wy = 0; }
            else if (q < 2 * WX) { wx = q - WX; wy = WY - 1; }
            else if (q < 2 * WX + TY) { wx = 0; wy = 1 + q - 2 * WX; }
            else { wx = WX - 1; wy = 1 + q - 2 * WX - TY; }
        }
        if (act) wact |= 1u << s;
        wl[s] = wy * WX + wx;
        const int px = x0 - 1 + wx, py = y0 - 1 + wy;
        const int qy = map_in(py, A.by), qx = map_in(px, A.bx);
        wfv[s] = (T)0;
        coff[s] = cell0;
        int ix = px, iy = py;
        if (qy < 0) {
            wfix |= 1u << s;
            wfv[s] = fixed_value(py, A.by);
        } else if (qx < 0) {
            wfix |= 1u << s;
            wfv[s] = fixed_value(px, A.bx);
        } else {
            coff[s] = (qy + g) * A.L.sx + (qx + g);
            if ((py < 0 && A.by.klo == PHX_BC_NOFLUX) || (py >= n1 && A.by.khi == PHX_BC_NOFLUX)) iy = qy;
            if ((px < 0 && A.bx.klo == PHX_BC_NOFLUX) || (px >= n0 && A.bx.khi == PHX_BC_NOFLUX)) ix = qx;
        }
        // images that leave the ring belong to corners or to cells past the
        // first ghost layer, which feed no output: keep their stencil in V
        ix = min(max(ix, x0 - 1), x0 + TX);
        iy = min(max(iy, y0 - 1), y0 + TY);
        vc[s] = (iy - y0 + 2) * VX + (ix - x0 + 2);
    }
    // output cells (beyond the domain: read a valid cell, never store)
    long long ogi[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int t = tid + r * 256;
        const int oy = t / TX, ox = t - oy * TX;
        const int x = min(x0 + ox, n0 - 1), y = min(y0 + oy, n1 - 1);
        ogi[r] = zb + (long long)(y + g) * A.L.sx + (x + g);
    }

    // every load before the first barrier
    T hv[NF][NV];
#pragma unroll
    for (int f = 0; f < NF; ++f)
#pragma unroll
        for (int i = 0; i < NV; ++i) hv[f][i] = A.u[f][zb + voff[i]];
    T cb[NF][NW], pa[NF][R];
    if constexpr (SECOND) {
#pragma unroll
        for (int f = 0; f < NF; ++f) {
#pragma unroll
            for (int s = 0; s < NW; ++s) cb[f][s] = A.c[f][zb + coff[s]];
#pragma unroll
            for (int r = 0; r < R; ++r) pa[f][r] = A.acc[f][ogi[r]];
        }
    }
#pragma unroll
    for (int f = 0; f < NF; ++f)
#pragma unroll
        for (int i = 0; i < NV; ++i) {
            const int t = tid + 256 * i;
            if (t < VX * VY) V[f][t] = ((vfm >> i) & 1u) ? vfv[i] : hv[f][i];
        }
    __syncthreads();

    // first stage (1 or 3) on the tile and its ring
    T am[NF][R];   // accumulator after the first stage, at the thread's cells
#pragma unroll
    for (int s = 0; s < NW; ++s) {
        T e[NF], k[NF];
        ac_rhs<T, NF, 2, EXACT>(A, &V[0][0], VY * VX, vc[s], VX, e, k);
#pragma unroll
        for (int f = 0; f < NF; ++f) {
            T ub;
            if constexpr (!SECOND) {   // stage 1: state = c_n
                ub = EXACT ? e[f] + A.a * k[f] : __builtin_fmaf(A.a, k[f], e[f]);
                if (s < R) am[f][s] = EXACT ? k[f] : __builtin_fmaf(A.dt6, k[f], e[f]);
            } else {                   // stage 3
                ub = EXACT ? cb[f][s] + A.a * k[f] : __builtin_fmaf(A.a, k[f], cb[f][s]);
                if (s < R) am[f][s] = EXACT ? pa[f][s] + (T)2 * k[f] : __builtin_fmaf(A.dt3, k[f], pa[f][s]);
            }
            if ((wfix >> s) & 1u) ub = wfv[s];
            if ((wact >> s) & 1u) W[f][wl[s]] = ub;
        }
    }
    __syncthreads();

    // second stage (2 or 4) on the tile
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int t = tid + r * 256;
        const int oy = t / TX, ox = t - oy * TX;
        if (x0 + ox >= n0 || y0 + oy >= n1) continue;
        T e[NF], k[NF];
        ac_rhs<T, NF, 2, EXACT>(A, &W[0][0], WY * WX, (oy + 1) * WX + ox + 1, WX, e, k);
        const long long gi = ogi[r];
#pragma unroll
        for (int f = 0; f < NF; ++f) {
            if constexpr (!SECOND) {   // stage 2: acc, and the stage-3 state from c_n
                const T cn = V[f][(oy + 2) * VX + ox + 2];
                if constexpr (EXACT) {
                    A.acc[f][gi] = am[f][r] + (T)2 * k[f];
                    A.out[f][gi] = cn + A.a2 * k[f];
                } else {
                    A.acc[f][gi] = __builtin_fmaf(A.dt3, k[f], am[f][r]);
                    A.out[f][gi] = __builtin_fmaf(A.a2, k[f], cn);
                }
            } else {                   // stage 4: c_{n+1}
                if constexpr (EXACT) A.out[f][gi] = cb[f][r] + A.dt6 * (am[f][r] + k[f]);
                else A.out[f][gi] = __builtin_fmaf(A.dt6, k[f], am[f][r]);
            }
        }
    }
}

int bc_fill_impl(const phx_grid* G, int dtype, void* data, hipStream_t st);

template <typename T>
static AxisBC<T> ac_axis(const phx_grid* G, int a) {
    AxisBC<T> b;
    b.n = G->n[a];
    b.klo = a < G->dim ? G->bc[a][0] : PHX_BC_PERIODIC;
    b.khi = a < G->dim ? G->bc[a][1] : PHX_BC_PERIODIC;
    b.vlo = (T)G->bc_value[a][0];
    b.vhi = (T)G->bc_value[a][1];
    return b;
}

constexpr int ACX = 64, ACY = 8;
// two-stage kernel tile: 32x8 (25 KB of LDS in FP32, 49 KB in FP64). It is
// latency bound; at 1024^2 x 8 FP32 the small tile's occupancy beats 64x8,
// 64x4, 32x16 and 128x4 (55.7 us per step against 59-81 us).
template <typename T>
constexpr int acp_tx() { return 32; }
template <typename T>
constexpr int acp_ty() { return 8; }
static int g_ac_pair = 1;

template <typename T, int NF, bool EXACT>
static void ac_pair_launch(bool second, const ACArgs<T>& A, dim3 grid, hipStream_t st) {
    constexpr int TX = acp_tx<T>(), TY = acp_ty<T>();
    if (second) hipLaunchKernelGGL((k_ac_pair_2d<T, NF, TX, TY, EXACT, true>), grid, dim3(256), 0, st, A);
    else hipLaunchKernelGGL((k_ac_pair_2d<T, NF, TX, TY, EXACT, false>), grid, dim3(256), 0, st, A);
}

template <typename T, bool EXACT>
static void ac_pair_dispatch(int nf, bool second, const ACArgs<T>& A, dim3 grid, hipStream_t st) {
    switch (nf) {
        case 1: ac_pair_launch<T, 1, EXACT>(second, A, grid, st); break;
        case 2: ac_pair_launch<T, 2, EXACT>(second, A, grid, st); break;
        case 3: ac_pair_launch<T, 3, EXACT>(second, A, grid, st); break;
        case 4: ac_pair_launch<T, 4, EXACT>(second, A, grid, st); break;
        case 5: ac_pair_launch<T, 5, EXACT>(second, A, grid, st); break;
        case 6: ac_pair_launch<T, 6, EXACT>(second, A, grid, st); break;
        case 7: ac_pair_launch<T, 7, EXACT>(second, A, grid, st); break;
        default: ac_pair_launch<T, 8, EXACT>(second, A, grid, st); break;
    }
}

template <typename T, int NF, bool EXACT>
static void ac_launch(const phx_grid* G, const ACArgs<T>& A, dim3 grid, hipStream_t st) {
    if (G->dim == 2) hipLaunchKernelGGL((k_ac_stage_2d<T, 2, NF, ACX, ACY, EXACT>), grid, dim3(256), 0, st, A);
    else hipLaunchKernelGGL((k_ac_stage_2d<T, 1, NF, ACX, ACY, EXACT>), grid, dim3(256), 0, st, A);
}

template <typename T, bool EXACT>
static void ac_dispatch(int nf, const phx_grid* G, const ACArgs<T>& A, dim3 grid, hipStream_t st) {
    switch (nf) {
        case 1: ac_launch<T, 1, EXACT>(G, A, grid, st); break;
        case 2: ac_launch<T, 2, EXACT>(G, A, grid, st); break;
        case 3: ac_launch<T, 3, EXACT>(G, A, grid, st); break;
        case 4: ac_launch<T, 4, EXACT>(G, A, grid, st); break;
        case 5: ac_launch<T, 5, EXACT>(G, A, grid, st); break;
        case 6: ac_launch<T, 6, EXACT>(G, A, grid, st); break;
        case 7: ac_launch<T, 7, EXACT>(G, A, grid, st); break;
        default: ac_launch<T, 8, EXACT>(G, A, grid, st); break;
    }
}

template <typename T, bool EXACT>
static int ac_steps_t(const phx_grid* G, int nf, T* const* eta, T* work, const phx_ac_params* p, int nsteps,
                      hipStream_t st) {
    ACArgs<T> A;
    A.L = make_layout(G);
    const long long N = A.L.total();
    T* w0[AC_MAXF];
    T* w1[AC_MAXF];
    for (int f = 0; f < nf; ++f) {
        w0[f] = work + (long long)(3 * f + 0) * N;
        w1[f] = work + (long long)(3 * f + 1) * N;
        A.acc[f] = work + (long long)(3 * f + 2) * N;
        A.c[f] = eta[f];
    }
    A.bx = ac_axis<T>(G, 0);
    A.by = ac_axis<T>(G, 1);
    A.idx2 = (T)(1.0 / (G->d[0] * G->d[0]));
    A.idy2 = G->dim >= 2 ? (T)(1.0 / (G->d[1] * G->d[1])) : (T)0;
    A.cl = (T)(p->L * p->kappa);
    A.cp = (T)p->L;
    A.cg = (T)(-2.0 * p->L * p->gamma);
    A.icc = (T)(-2.0 * (1.0 / (G->d[0] * G->d[0]) + (G->dim >= 2 ? 1.0 / (G->d[1] * G->d[1]) : 0.0)));
    const double dt = p->dt, dt6 = dt / 6.0;
    A.dt6 = (T)dt6;
    A.dt3 = (T)(2.0 * dt6);
    const T shift[4] = {(T)(dt * 0.5), (T)(dt * 0.5), (T)dt, (T)0};
    if (G->dim == 2 && g_ac_pair) {
        // Two-stage launches. c_n alternates between two buffers, X and Y,
        // chosen so that the last step reads the caller's arrays: their ghost
        // cells then hold the BC fill of c_{n-1} (filled before the last step)
        // as on the one-stage path, and the last step's interior result is
        // copied back into them.
        const dim3 pgrid(cdiv(A.L.n0, acp_tx<T>()), cdiv(A.L.n1, acp_ty<T>()), 1);
        A.xcd_remap = (pgrid.x * pgrid.y) % 8 == 0;
        const bool even = nsteps % 2 == 0;
        const size_t bytes = (size_t)N * sizeof(T);
        T* X[AC_MAXF];
        T* Y[AC_MAXF];
        for (int f = 0; f < nf; ++f) {
            X[f] = even ? w0[f] : eta[f];
            Y[f] = even ? eta[f] : w0[f];
            if (even) PHX_HIP(hipMemcpyAsync(w0[f], eta[f], bytes, hipMemcpyDeviceToDevice, st));
        }
        for (int s = 0; s < nsteps; ++s) {
            T* const* cur = s % 2 == 0 ? X : Y;
            T* const* nxt = s % 2 == 0 ? Y : X;
            if (s == nsteps - 1)
                for (int f = 0; f < nf; ++f) {
                    int rc = bc_fill_impl(G, sizeof(T) == 8 ? PHX_F64 : PHX_F32, cur[f], st);
                    if (rc) return rc;
                }
            A.a = shift[0];
            A.a2 = shift[1];
            for (int f = 0; f < nf; ++f) {
                A.u[f] = cur[f];
                A.c[f] = cur[f];
                A.out[f] = w1[f];
            }
            launch_timer_begin(st);
            ac_pair_dispatch<T, EXACT>(nf, false, A, pgrid, st);
            launch_timer_end(5, st);
            A.a = shift[2];
            for (int f = 0; f < nf; ++f) {
                A.u[f] = w1[f];
                A.out[f] = nxt[f];
            }
            launch_timer_begin(st);
            ac_pair_dispatch<T, EXACT>(nf, true, A, pgrid, st);
            launch_timer_end(6, st);
        }
        // the last step wrote w0: its interior goes to the caller's arrays
        const size_t pitch = (size_t)A.L.sx * sizeof(T);
        const long long o0 = A.L.at(0, 0, 0);
        for (int f = 0; f < nf; ++f)
            PHX_HIP(hipMemcpy2DAsync(eta[f] + o0, pitch, w0[f] + o0, pitch, (size_t)A.L.n0 * sizeof(T), A.L.n1,
                                     hipMemcpyDeviceToDevice, st));
        PHX_LAUNCH_CHECK("phx_ac_rk4_steps");
        return PHX_OK;
    }
    dim3 grid(cdiv(A.L.n0, ACX), cdiv(A.L.n1, ACY), 1);
    A.xcd_remap = (grid.x * grid.y) % 8 == 0;
    for (int s = 0; s < nsteps; ++s) {
        if (s == nsteps - 1)
            for (int f = 0; f < nf; ++f) {
                int rc = bc_fill_impl(G, sizeof(T) == 8 ? PHX_F64 : PHX_F32, eta[f], st);
                if (rc) return rc;
            }
        for (int stg = 0; stg < 4; ++stg) {
            A.stage = stg + 1;
            A.a = shift[stg];
            for (int f = 0; f < nf; ++f) {
                A.u[f] = stg == 0 ? eta[f] : (stg == 2 ? w1[f] : w0[f]);
                A.out[f] = stg == 3 ? eta[f] : (stg == 1 ? w1[f] : w0[f]);
            }
            launch_timer_begin(st);
            ac_dispatch<T, EXACT>(nf, G, A, grid, st);
            launch_timer_end(4, st);
        }
    }
    PHX_LAUNCH_CHECK("phx_ac_rk4_steps");
    return PHX_OK;
}

}  // namespace phx

using namespace phx;

extern "C" {

size_t phx_ac_rk4_workspace_bytes(const phx_grid* G, int dtype, int nfields) {
    if (check_grid(G, "phx_ac_rk4_workspace_bytes") || check_dtype(dtype, "phx_ac_rk4_workspace_bytes")) return 0;
    if (nfields < 1 || nfields > AC_MAXF) return 0;
    return (size_t)3 * nfields * make_layout(G).total() * (dtype == PHX_F64 ? 8 : 4);
}

int phx_ac_set_two_stage(int on) {
    g_ac_pair = on ? 1 : 0;
    return PHX_OK;
}

int phx_ac_rk4_steps(const phx_grid* G, int dtype, int nfields, void* const* eta, void* work,
                     const phx_ac_params* p, int nsteps, phx_stream s) {
    int rc = check_grid_fused(G, "phx_ac_rk4_steps");
    if (rc) return rc;
    if ((rc = check_dtype(dtype, "phx_ac_rk4_steps"))) return rc;
    if (G->dim > 2) return fail(PHX_ERR_UNSUPPORTED, "phx_ac_rk4_steps: dim must be 1 or 2");
    if (nfields < 1 || nfields > AC_MAXF)
        return fail(PHX_ERR_UNSUPPORTED, "phx_ac_rk4_steps: nfields must be in [1, 8]");
    if (!eta || !work || !p) return fail(PHX_ERR_INVALID, "phx_ac_rk4_steps: NULL argument");
    for (int f = 0; f < nfields; ++f)
        if (!eta[f]) return fail(PHX_ERR_INVALID, "phx_ac_rk4_steps: NULL field pointer");
    if (!(p->dt > 0.0)) return fail(PHX_ERR_INVALID, "phx_ac_rk4_steps: dt must be > 0");
    if (nsteps < 0) return fail(PHX_ERR_INVALID, "phx_ac_rk4_steps: nsteps < 0");
    if (nsteps == 0) return PHX_OK;
    hipStream_t st = as_stream(s);
    return dtype == PHX_F64 ? ac_steps_t<double, true>(G, nfields, (double* const*)eta, (double*)work, p, nsteps, st)
                            : ac_steps_t<float, false>(G, nfields, (float* const*)eta, (float*)work, p, nsteps, st);
}

}  // extern "C"
