// ensemble.hip -- batched ensemble of independent 2D periodic Cahn-Hilliard
// runs of uneven sizes, all advanced by one launch per RK4 stage.
//
// No reference counterpart: PhiX runs one problem per process. Layout: the
// runs' physical N_r x N_r FP32 fields packed back to back in one arena (no
// ghost cells -- periodic halos are index maps), offsets in a device table.
// A stage is one launch of a persistent kernel: a fixed grid of workgroups
// walks a work list of (run, tile) entries ordered largest run first, so the
// uneven sizes load-balance across the 256 CUs. The step loop (4 stage
// launches per step) is captured once into a hipGraph and replayed.
// Per-run mass is reduced in FP64 with Neumaier-compensated sums, one
// workgroup per run.
// Tile shape and grid size come from phx_ensemble_autotune (timed candidates).
// Every kernel evaluates a cell through the same helpers (explicit FMAs,
// contraction off), so every tiling returns the same bits.
#pragma clang fp contract(off)

#include "common.hpp"

#include <algorithm>
#include <cstring>
#include <map>
#include <mutex>
#include <numeric>
#include <vector>

namespace phx {

struct EnsRun {
    long long off;   // element offset of the run in the arena
    int n;           // side length
    float kappa;     // per-run kappa (f = c^2(1-c)^2, M = 1 unless set)
    float M;
};

struct EnsTile {
    int run;
    int x0, y0;
};

struct EnsArgs {
    const float* u;
    const float* c;
    float* acc;
    float* out;
    const EnsRun* runs;
    const EnsTile* tiles;
    int ntiles;
    float a, dt6, dt3;
    float a2;       // two-stage kernel: shift of its second stage
    int stage;
    int xcd_walk;   // 1: XCD-aware tile walk (grid size a multiple of 8)
};

// Periodic wrap for |p - n| < n.
__device__ __forceinline__ int wrapp(int p, int n) { return p < 0 ? p + n : (p >= n ? p - n : p); }

// mu = f'(c) - kappa*lap(c) on a unit grid, f = c^2(1-c)^2:
// lap = (xp + xm) + (yp + ym) - 4c, f'(c) = 2c (c - 1) (2c - 1).
__device__ __forceinline__ float ens_mu(float c, float xp, float xm, float yp, float ym, float nk) {
    const float lap = __builtin_fmaf(-4.f, c, (xp + xm) + (yp + ym));
    const float fp = (2.f * c) * (c - 1.f) * __builtin_fmaf(2.f, c, -1.f);
    return __builtin_fmaf(nk, lap, fp);
}
// k = M * lap(mu)
__device__ __forceinline__ float ens_k(float m, float xp, float xm, float yp, float ym, float M) {
    return M * __builtin_fmaf(-4.f, m, (xp + xm) + (yp + ym));
}
// RK4 stage epilogue (FP32 16-pass form)
__device__ __forceinline__ void ens_epilogue(const EnsArgs& A, long long gi, int st, float uc, float k, float cn,
                                             float an) {
    if (st == 1) {
        A.acc[gi] = __builtin_fmaf(A.dt6, k, uc);
        A.out[gi] = __builtin_fmaf(A.a, k, uc);
    } else if (st < 4) {
        A.acc[gi] = __builtin_fmaf(A.dt3, k, an);
        A.out[gi] = __builtin_fmaf(A.a, k, cn);
    } else {
        A.out[gi] = __builtin_fmaf(A.dt6, k, an);
    }
}

// One persistent workgroup walks tiles w, w + G, w + 2G, ... (G = grid size).
// With the XCD-aware walk, the tile list is split into 8 contiguous ranges,
// one per XCD (consecutive workgroups are dispatched to the 8 XCDs in turn),
// and the G/8 workgroups of an XCD walk its range with stride G/8: tiles
// that share halo rows are then processed on the same XCD at about the same
// time, and the shared rows come from its own L2. Each workgroup is software
// pipelined across its tiles: the next tile's halo is in flight
// (registers) while the current tile's mu is computed, its epilogue operands
// while the current tile's outputs are written, and tile descriptors two tiles
// ahead. Every load is unconditional (wrapped, clamped offsets; clamped tile
// indices past the end re-read a valid tile), so each iteration issues the
// same loads and the compiler waits only for the ones it uses.
template <int TX, int TY>
__global__ __launch_bounds__(256) void k_ens_stage(EnsArgs A) {
    constexpr int UX = TX + 4, UY = TY + 4, MX = TX + 2, MY = TY + 2;
    constexpr int NS = (UX * UY + 255) / 256;
    constexpr int NO = (TX * TY) / 256;
    static_assert((TX * TY) % 256 == 0, "tile must be a multiple of 256 cells");
    __shared__ float U[UY * UX];
    __shared__ float MU[MY * MX];
    const int tid = threadIdx.x;
    int w, G, last;   // first tile, stride, last tile of this workgroup's walk
    if (A.xcd_walk) {
        const int x = blockIdx.x & 7;
        w = (int)((long long)x * A.ntiles / 8) + (blockIdx.x >> 3);
        G = gridDim.x >> 3;
        last = (int)((long long)(x + 1) * A.ntiles / 8) - 1;
    } else {
        w = blockIdx.x;
        G = gridDim.x;
        last = A.ntiles - 1;
    }
    if (w > last) return;
    // epilogue sources: an operand the stage does not need is still loaded
    // (every iteration issues the same loads), from the arena's first element
    // for every lane -- one cache line per wave instead of a second pass over
    // the stage state (the next tile's operands are in flight before its halo
    // has reached the cache)
    const int st = A.stage;
    const bool need_c = st == 2 || st == 3, need_a = st >= 2;
    const float* csrc = need_c ? A.c : A.u;
    const float* asrc = need_a ? A.acc : A.u;
    const long long cmask = need_c ? -1LL : 0LL, amask = need_a ? -1LL : 0LL;

    float h[NS], pc[NO], pa[NO];
    auto load_halo = [&](const EnsTile& T, const EnsRun& R) {
        const int n = R.n;
        const float* u = A.u + R.off;
#pragma unroll
        for (int i = 0; i < NS; ++i) {
            const int t = min(tid + 256 * i, UX * UY - 1);
            const int ly = t / UX, lx = t - ly * UX;
            const int px = min(max(T.x0 - 2 + lx, -2), n + 1), py = min(max(T.y0 - 2 + ly, -2), n + 1);
            h[i] = u[(long long)wrapp(py, n) * n + wrapp(px, n)];
        }
    };
    auto load_epi = [&](const EnsTile& T, const EnsRun& R) {
        const int n = R.n;
#pragma unroll
        for (int j = 0; j < NO; ++j) {
            const int t = tid + 256 * j;
            const int oy = t / TX, ox = t - oy * TX;
            const long long gi = R.off + (long long)min(T.y0 + oy, n - 1) * n + min(T.x0 + ox, n - 1);
            pc[j] = csrc[gi & cmask];
            pa[j] = asrc[gi & amask];
        }
    };

    // tile and run indices are workgroup-uniform: in scalar registers the
    // descriptor loads are scalar loads, whose waits do not drain the vector
    // loads in flight
    auto uidx = [](int v) { return __builtin_amdgcn_readfirstlane(v); };
    // the tables are read-only during the launch: read them through the
    // constant address space, which the compiler serves with scalar loads
    using CTile = const __attribute__((address_space(4))) EnsTile;
    using CRun = const __attribute__((address_space(4))) EnsRun;
    CTile* ctiles = (CTile*)A.tiles;
    CRun* cruns = (CRun*)A.runs;
    auto tile_at = [&](int i) {
        EnsTile t;
        t.run = ctiles[i].run;
        t.x0 = ctiles[i].x0;
        t.y0 = ctiles[i].y0;
        return t;
    };
    auto run_at = [&](int i) {
        EnsRun r;
        r.off = cruns[i].off;
        r.n = cruns[i].n;
        r.kappa = cruns[i].kappa;
        r.M = cruns[i].M;
        return r;
    };
    EnsTile T = tile_at(uidx(w));
    EnsRun R = run_at(uidx(T.run));
    EnsTile Tn = tile_at(uidx(min(w + G, last)));
    EnsRun Rn = run_at(uidx(Tn.run));
    load_halo(T, R);
    load_epi(T, R);
    for (;;) {
        // descriptors two tiles ahead
        const EnsTile T2 = tile_at(uidx(min(w + 2 * G, last)));
        const EnsRun R2 = run_at(uidx(T2.run));
#pragma unroll
        for (int i = 0; i < NS; ++i)
            if (tid + 256 * i < UX * UY) U[tid + 256 * i] = h[i];
        __syncthreads();
        load_halo(Tn, Rn);   // next tile's halo, in flight during mu and the outputs
        const int n = R.n;
        const long long off = R.off;
        const float nk = -R.kappa;
        for (int t = tid; t < MX * MY; t += 256) {
            int my = t / MX, mx = t - my * MX;
            const int c0 = (my + 1) * UX + mx + 1;
            MU[t] = ens_mu(U[c0], U[c0 + 1], U[c0 - 1], U[c0 + UX], U[c0 - UX], nk);
        }
        __syncthreads();
#pragma unroll
        for (int j = 0; j < NO; ++j) {
            const int t = tid + 256 * j;
            const int oy = t / TX, ox = t - oy * TX;
            const int x = T.x0 + ox, y = T.y0 + oy;
            if (x < n && y < n) {
                const int m0 = (oy + 1) * MX + ox + 1;
                const float k = ens_k(MU[m0], MU[m0 + 1], MU[m0 - 1], MU[m0 + MX], MU[m0 - MX], R.M);
                const long long gi = off + (long long)y * n + x;
                ens_epilogue(A, gi, st, U[(oy + 2) * UX + ox + 2], k, pc[j], pa[j]);
            }
        }
        w += G;
        if (w > last) break;
        load_epi(Tn, Rn);   // next tile's epilogue operands
        T = Tn;
        R = Rn;
        Tn = T2;
        Rn = R2;
        __syncthreads();   // LDS is reused by the next tile of this workgroup
    }
}

// Two RK4 stages per launch (temporal blocking): stages (1,2) (SECOND =
// false) or (3,4) (SECOND = true), persistent like k_ens_stage, each tile's
// loads issued while the previous tile is computed. Per step
//   (1,2): read c_n;                          write acc and the stage-3 state;
//   (3,4): read the stage-3 state, c_n, acc;  write c_{n+1} into a buffer other
//          than c_n (the launch reads c_n around its tile),
// 7 field passes instead of 16. The first stage's input is staged with a
// 4-cell periodic halo; mu and the first stage's result are formed on the
// rings the second stage needs (mu_A on tile+3, u_B on tile+2, mu_B on
// tile+1). Ring cells outside a run are evaluated from the wrapped input,
// which repeats the operands of the cell they wrap to, in the same order, so
// the results equal two one-stage launches bit for bit.
// p in [-4, n + 3] wrapped into [0, n) for any n >= 2.
__device__ __forceinline__ int wrap4(int p, int n) {
    p = p < 0 ? p + n : p;
    p = p < 0 ? p + n : p;
    p = p >= n ? p - n : p;
    return p >= n ? p - n : p;
}

template <int TX, int TY, bool SECOND>
__global__ __launch_bounds__(256) void k_ens_pair(EnsArgs A) {
    // Every level lives in one frame of pitch UX: tile position (x, y) sits at
    // (y + 4) * UX + x + 4 in U, MA, UB, MB, CB and KA alike. A level covers
    // whole frame rows (ring R: rows 4-R .. 4+TY+R-1, all UX columns), cut
    // into runs of 4 positions: a run's operands are three 16-byte LDS reads
    // and the two scalars beside it, its results one 16-byte store. Columns
    // outside a level's ring read finite entries of the level before (whose
    // rows reach one further) and feed only other out-of-ring entries.
    constexpr int UX = TX + 8, UY = TY + 8;   // the frame: first-stage input, radius 4
    constexpr int NF = UX * UY;
    constexpr int NS = (NF + 255) / 256;
    constexpr int CR0 = 2 * UX, NCB = (TY + 4) * UX;   // c_n (SECOND) on whole rows of ring 2
    constexpr int NC = SECOND ? (NCB + 255) / 256 : 1;
    constexpr int NO = (TX * TY) / 256;
    static_assert((TX * TY) % 256 == 0, "tile must be a multiple of 256 cells");
    static_assert(UX % 4 == 0, "frame rows hold whole runs");
    __shared__ __attribute__((aligned(16))) float U[NF];
    __shared__ __attribute__((aligned(16))) float MA[NF];
    __shared__ __attribute__((aligned(16))) float UB[NF];
    __shared__ __attribute__((aligned(16))) float MB[NF];
    __shared__ __attribute__((aligned(16))) float CB[SECOND ? NF : 4];
    __shared__ __attribute__((aligned(16))) float KA[NF];   // first-stage RHS (read at the tile cells)
    const int tid = threadIdx.x;
    // persistent walk, as k_ens_stage (XCD-aware when the grid allows)
    int w, G, last;
    if (A.xcd_walk) {
        const int x = blockIdx.x & 7;
        w = (int)((long long)x * A.ntiles / 8) + (blockIdx.x >> 3);
        G = gridDim.x >> 3;
        last = (int)((long long)(x + 1) * A.ntiles / 8) - 1;
    } else {
        w = blockIdx.x;
        G = gridDim.x;
        last = A.ntiles - 1;
    }
    if (w > last) return;
    auto uidx = [](int v) { return __builtin_amdgcn_readfirstlane(v); };
    using CTile = const __attribute__((address_space(4))) EnsTile;
    using CRun = const __attribute__((address_space(4))) EnsRun;
    CTile* ctiles = (CTile*)A.tiles;
    CRun* cruns = (CRun*)A.runs;
    auto tile_at = [&](int i) {
        EnsTile t;
        t.run = ctiles[i].run;
        t.x0 = ctiles[i].x0;
        t.y0 = ctiles[i].y0;
        return t;
    };
    auto run_at = [&](int i) {
        EnsRun r;
        r.off = cruns[i].off;
        r.n = cruns[i].n;
        r.kappa = cruns[i].kappa;
        r.M = cruns[i].M;
        return r;
    };
    // a tile's loads, issued one tile ahead (registers) so that they are in
    // flight while the current tile is computed
    float hv[NS], cv[NC], pa[NO];
    auto load = [&](const EnsTile& T, const EnsRun& R) {
        const int n = R.n;
        const long long off = R.off;
#pragma unroll
        for (int i = 0; i < NS; ++i) {
            const int t = min(tid + 256 * i, UX * UY - 1);
            const int ly = t / UX, lx = t - ly * UX;
            const int px = wrap4(min(T.x0 - 4 + lx, n + 3), n), py = wrap4(min(T.y0 - 4 + ly, n + 3), n);
            hv[i] = A.u[off + (long long)py * n + px];
        }
        if constexpr (SECOND) {
#pragma unroll
            for (int i = 0; i < NC; ++i) {
                const int t = min(tid + 256 * i, NCB - 1);
                const int ly = t / UX, lx = t - ly * UX;
                const int px = wrap4(min(T.x0 - 4 + lx, n + 3), n), py = wrap4(min(T.y0 - 2 + ly, n + 3), n);
                cv[i] = A.c[off + (long long)py * n + px];
            }
#pragma unroll
            for (int j = 0; j < NO; ++j) {
                const int t = tid + 256 * j;
                const int oy = t / TX, ox = t - oy * TX;
                pa[j] = A.acc[off + (long long)min(T.y0 + oy, n - 1) * n + min(T.x0 + ox, n - 1)];
            }
        }
    };
    EnsTile T = tile_at(uidx(w));
    EnsRun R = run_at(uidx(T.run));
    load(T, R);
    for (;;) {
        const int wn = w + G;
        const EnsTile Tn = tile_at(uidx(min(wn, last)));
        const EnsRun Rn = run_at(uidx(Tn.run));
#pragma unroll
        for (int i = 0; i < NS; ++i)
            if (tid + 256 * i < UX * UY) U[tid + 256 * i] = hv[i];
        float pc[NO];
        if constexpr (SECOND) {
#pragma unroll
            for (int i = 0; i < NC; ++i)
                if (tid + 256 * i < NCB) CB[CR0 + tid + 256 * i] = cv[i];
#pragma unroll
            for (int j = 0; j < NO; ++j) pc[j] = pa[j];
        }
        __syncthreads();
        load(Tn, Rn);   // next tile (the last tile again at the end: valid, unused)
        const int n = R.n;
        const long long off = R.off;
        const float nk = -R.kappa, M = R.M;
        typedef float f4 __attribute__((ext_vector_type(4)));
        // the run at f of a level held in src: centre, x and y neighbours
        auto run_ops = [&](const float* src, int f, f4& c, f4& xp, f4& xm, f4& yp, f4& ym) {
            c = *reinterpret_cast<const f4*>(&src[f]);
            yp = *reinterpret_cast<const f4*>(&src[f + UX]);
            ym = *reinterpret_cast<const f4*>(&src[f - UX]);
            const float l = src[f - 1], r = src[f + 4];
            xp = f4{c[1], c[2], c[3], r};
            xm = f4{l, c[0], c[1], c[2]};
        };
        // runs of the level on ring R: thread tid takes runs tid + 256 i (the
        // last partial pass reads a valid run and skips the store)
        auto level = [&](auto rr, auto body) {
            constexpr int R = decltype(rr)::value;
            constexpr int F0 = (4 - R) * UX, NQ = (TY + 2 * R) * UX / 4, NIT = (NQ + 255) / 256;
#pragma unroll
            for (int i = 0; i < NIT; ++i) {
                const int q = tid + 256 * i;
                const int f = F0 + 4 * (NQ % 256 != 0 && i == NIT - 1 ? min(q, NQ - 1) : q);
                body(f, NQ % 256 == 0 || i < NIT - 1 || q < NQ);
            }
        };
        // first-stage mu on ring 3
        level(std::integral_constant<int, 3>{}, [&](int f, bool st4) {
            f4 c, xp, xm, yp, ym, o;
            run_ops(U, f, c, xp, xm, yp, ym);
#pragma unroll
            for (int k = 0; k < 4; ++k) o[k] = ens_mu(c[k], xp[k], xm[k], yp[k], ym[k], nk);
            if (st4) *reinterpret_cast<f4*>(&MA[f]) = o;
        });
        __syncthreads();
        // first stage: u_B = base + a k on ring 2; k kept (read at the tile cells)
        level(std::integral_constant<int, 2>{}, [&](int f, bool st4) {
            f4 m, xp, xm, yp, ym, o, kk;
            run_ops(MA, f, m, xp, xm, yp, ym);
            const f4 base = *reinterpret_cast<const f4*>(SECOND ? &CB[f] : &U[f]);
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                kk[k] = ens_k(m[k], xp[k], xm[k], yp[k], ym[k], M);
                o[k] = __builtin_fmaf(A.a, kk[k], base[k]);
            }
            if (st4) {
                *reinterpret_cast<f4*>(&UB[f]) = o;
                *reinterpret_cast<f4*>(&KA[f]) = kk;
            }
        });
        __syncthreads();
        // second-stage mu on ring 1
        level(std::integral_constant<int, 1>{}, [&](int f, bool st4) {
            f4 c, xp, xm, yp, ym, o;
            run_ops(UB, f, c, xp, xm, yp, ym);
#pragma unroll
            for (int k = 0; k < 4; ++k) o[k] = ens_mu(c[k], xp[k], xm[k], yp[k], ym[k], nk);
            if (st4) *reinterpret_cast<f4*>(&MB[f]) = o;
        });
        __syncthreads();
        // second stage and the RK4 epilogue at the tile cells
#pragma unroll
        for (int j = 0; j < NO; ++j) {
            const int t = tid + 256 * j;
            const int oy = t / TX, ox = t - oy * TX;
            const int x = T.x0 + ox, y = T.y0 + oy;
            if (x >= n || y >= n) continue;
            const int m0 = (oy + 4) * UX + ox + 4;
            const float k = ens_k(MB[m0], MB[m0 + 1], MB[m0 - 1], MB[m0 + UX], MB[m0 - UX], M);
            const long long gi = off + (long long)y * n + x;
            if constexpr (!SECOND) {
                const float cn = U[m0];
                const float am = __builtin_fmaf(A.dt6, KA[m0], cn);   // accumulator after stage 1
                A.acc[gi] = __builtin_fmaf(A.dt3, k, am);
                A.out[gi] = __builtin_fmaf(A.a2, k, cn);
            } else {
                const float am = __builtin_fmaf(A.dt3, KA[m0], pc[j]);   // after stage 3
                A.out[gi] = __builtin_fmaf(A.dt6, k, am);
            }
        }
        if (wn > last) break;
        w = wn;
        T = Tn;
        R = Rn;
        __syncthreads();   // LDS is reused by the next tile
    }
}

// FP64 compensated per-run sum: one workgroup per run.
__global__ __launch_bounds__(256) void k_ens_mass(const float* __restrict__ c, const EnsRun* __restrict__ runs,
                                                  double* __restrict__ out) {
    const EnsRun R = runs[blockIdx.x];
    const long long cnt = (long long)R.n * R.n;
    double s = 0.0, comp = 0.0;
    for (long long q = threadIdx.x; q < cnt; q += 256) {
        double x = (double)c[R.off + q];
        double t = s + x;
        comp += fabs(s) >= fabs(x) ? (s - t) + x : (x - t) + s;
        s = t;
    }
    __shared__ double S[256], Cc[256];
    S[threadIdx.x] = s;
    Cc[threadIdx.x] = comp;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if (threadIdx.x < h) {
            double a = S[threadIdx.x], b = S[threadIdx.x + h];
            double t = a + b;
            double e = fabs(a) >= fabs(b) ? (a - t) + b : (b - t) + a;
            S[threadIdx.x] = t;
            Cc[threadIdx.x] += Cc[threadIdx.x + h] + e;
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) out[blockIdx.x] = S[0] + Cc[0];
}

// ------------------------------------------------------------------ host --
struct TileShape {
    int tx, ty;
};
static const TileShape kShapes[] = {{64, 16}, {32, 16}, {64, 8}, {128, 8}, {32, 32}};
constexpr int kNumShapes = sizeof(kShapes) / sizeof(kShapes[0]);

template <int TX, int TY>
static void launch_stage(int grid, const EnsArgs& A, hipStream_t st) {
    hipLaunchKernelGGL((k_ens_stage<TX, TY>), dim3(grid), dim3(256), 0, st, A);
}

template <int TX, int TY>
static void launch_pair_t(bool second, int grid, const EnsArgs& A, hipStream_t st) {
    if (second) hipLaunchKernelGGL((k_ens_pair<TX, TY, true>), dim3(grid), dim3(256), 0, st, A);
    else hipLaunchKernelGGL((k_ens_pair<TX, TY, false>), dim3(grid), dim3(256), 0, st, A);
}

static void launch_pair(int shape, bool second, int grid, const EnsArgs& A, hipStream_t st) {
    switch (shape) {
        case 0: launch_pair_t<64, 16>(second, grid, A, st); break;
        case 1: launch_pair_t<32, 16>(second, grid, A, st); break;
        case 2: launch_pair_t<64, 8>(second, grid, A, st); break;
        case 3: launch_pair_t<128, 8>(second, grid, A, st); break;
        default: launch_pair_t<32, 32>(second, grid, A, st); break;
    }
}

static void launch_shape(int shape, int grid, const EnsArgs& A, hipStream_t st) {
    switch (shape) {
        case 0: launch_stage<64, 16>(grid, A, st); break;
        case 1: launch_stage<32, 16>(grid, A, st); break;
        case 2: launch_stage<64, 8>(grid, A, st); break;
        case 3: launch_stage<128, 8>(grid, A, st); break;
        default: launch_stage<32, 32>(grid, A, st); break;
    }
}

}  // namespace phx

using namespace phx;

// Opaque handle: device tables for one ensemble and one tile shape, plus the
// captured graph. Allocation happens in phx_ensemble_create only.
struct phx_ensemble {
    int nruns = 0;
    long long total = 0;
    std::vector<EnsRun> runs;
    EnsRun* d_runs = nullptr;
    EnsTile* d_tiles[kNumShapes] = {};
    int ntiles[kNumShapes] = {};
    int shape = 0;
    int grid = 0;
    int two_stage = 1;   // 1: two RK4 stages per launch (k_ens_pair)
    int cus = 256;
    hipGraphExec_t graph = nullptr;
    int graph_steps = 0;
    const void* graph_key[4] = {};
    double dt = 0.0;
};

extern "C" {

int phx_ensemble_create(int nruns, const int* sides, const double* kappa, const double* mobility,
                        phx_ensemble** out) {
    if (!out || nruns <= 0 || !sides || !kappa) return fail(PHX_ERR_INVALID, "phx_ensemble_create: bad arguments");
    *out = nullptr;
    phx_ensemble* E = new phx_ensemble();
    E->nruns = nruns;
    E->runs.resize(nruns);
    long long off = 0;
    for (int r = 0; r < nruns; ++r) {
        if (sides[r] < 2 || sides[r] > 46340) {
            delete E;
            return fail(PHX_ERR_INVALID, "phx_ensemble_create: side lengths must be in [2, 46340]");
        }
        E->runs[r] = EnsRun{off, sides[r], (float)kappa[r], (float)(mobility ? mobility[r] : 1.0)};
        off += (long long)sides[r] * sides[r];
    }
    E->total = off;
    int dev = 0;
    hipDeviceProp_t prop;
    if (hipGetDevice(&dev) == hipSuccess && hipGetDeviceProperties(&prop, dev) == hipSuccess)
        E->cus = prop.multiProcessorCount;
    // work lists per tile shape, largest runs first
    std::vector<int> order(nruns);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return sides[a] > sides[b]; });
    if (hipMalloc(&E->d_runs, sizeof(EnsRun) * nruns) != hipSuccess ||
        hipMemcpy(E->d_runs, E->runs.data(), sizeof(EnsRun) * nruns, hipMemcpyHostToDevice) != hipSuccess) {
        phx_ensemble_destroy(E);
        return fail(PHX_ERR_NOMEM, "phx_ensemble_create: device tables");
    }
    for (int s = 0; s < kNumShapes; ++s) {
        std::vector<EnsTile> tl;
        for (int r : order) {
            int n = sides[r];
            for (int y = 0; y < n; y += kShapes[s].ty)
                for (int x = 0; x < n; x += kShapes[s].tx) tl.push_back(EnsTile{r, x, y});
        }
        E->ntiles[s] = (int)tl.size();
        if (hipMalloc(&E->d_tiles[s], sizeof(EnsTile) * tl.size()) != hipSuccess ||
            hipMemcpy(E->d_tiles[s], tl.data(), sizeof(EnsTile) * tl.size(), hipMemcpyHostToDevice) != hipSuccess) {
            phx_ensemble_destroy(E);
            return fail(PHX_ERR_NOMEM, "phx_ensemble_create: device tables");
        }
    }
    E->shape = 0;
    E->grid = std::min(E->ntiles[0], 8 * E->cus);
    *out = E;
    return PHX_OK;
}

int phx_ensemble_destroy(phx_ensemble* E) {
    if (!E) return PHX_OK;
    if (E->graph) (void)hipGraphExecDestroy(E->graph);
    if (E->d_runs) (void)hipFree(E->d_runs);
    for (int s = 0; s < kNumShapes; ++s)
        if (E->d_tiles[s]) (void)hipFree(E->d_tiles[s]);
    delete E;
    return PHX_OK;
}

long long phx_ensemble_cells(const phx_ensemble* E) { return E ? E->total : -1; }

long long phx_ensemble_offset(const phx_ensemble* E, int run) {
    if (!E || run < 0 || run >= E->nruns) return -1;
    return E->runs[run].off;
}

int phx_ensemble_set_tiling(phx_ensemble* E, int shape, int grid) {
    if (!E || shape < 0 || shape >= kNumShapes) return fail(PHX_ERR_INVALID, "phx_ensemble_set_tiling: bad shape");
    E->shape = shape;
    E->grid = grid > 0 ? std::min(grid, E->ntiles[shape]) : std::min(E->ntiles[shape], 8 * E->cus);
    if (E->graph) {
        (void)hipGraphExecDestroy(E->graph);
        E->graph = nullptr;
    }
    return PHX_OK;
}

int phx_ensemble_set_two_stage(phx_ensemble* E, int on) {
    if (!E) return fail(PHX_ERR_INVALID, "phx_ensemble_set_two_stage: NULL");
    E->two_stage = on ? 1 : 0;
    if (E->graph) {
        (void)hipGraphExecDestroy(E->graph);
        E->graph = nullptr;
    }
    return PHX_OK;
}

int phx_ensemble_get_two_stage(const phx_ensemble* E) { return E ? E->two_stage : -1; }

int phx_ensemble_get_tiling(const phx_ensemble* E, int* shape, int* grid, int* tx, int* ty) {
    if (!E) return fail(PHX_ERR_INVALID, "phx_ensemble_get_tiling: NULL");
    if (shape) *shape = E->shape;
    if (grid) *grid = E->grid;
    if (tx) *tx = kShapes[E->shape].tx;
    if (ty) *ty = kShapes[E->shape].ty;
    return PHX_OK;
}

static void enqueue_steps(phx_ensemble* E, float* c, float* work, double dt, int nsteps, hipStream_t st) {
    const long long N = E->total;
    float* w0 = work;
    float* w1 = work + N;
    float* acc = work + 2 * N;
    EnsArgs A;
    A.c = c;
    A.acc = acc;
    A.runs = E->d_runs;
    A.tiles = E->d_tiles[E->shape];
    A.ntiles = E->ntiles[E->shape];
    A.xcd_walk = E->grid % 8 == 0;
    const double dt6 = dt / 6.0;
    A.dt6 = (float)dt6;
    A.dt3 = (float)(2.0 * dt6);
    const float shift[4] = {(float)(dt * 0.5), (float)(dt * 0.5), (float)dt, 0.f};
    if (E->two_stage) {
        // c_n alternates between c and w0; after an odd step count the state
        // is copied back into c
        for (int s = 0; s < nsteps; ++s) {
            float* cur = s % 2 == 0 ? c : w0;
            float* nxt = s % 2 == 0 ? w0 : c;
            A.u = cur;
            A.c = cur;
            A.out = w1;
            A.a = shift[0];
            A.a2 = shift[1];
            launch_timer_begin(st);
            launch_pair(E->shape, false, E->grid, A, st);
            launch_timer_end(8, st);
            A.u = w1;
            A.out = nxt;
            A.a = shift[2];
            launch_timer_begin(st);
            launch_pair(E->shape, true, E->grid, A, st);
            launch_timer_end(9, st);
        }
        if (nsteps % 2) (void)hipMemcpyAsync(c, w0, (size_t)N * sizeof(float), hipMemcpyDeviceToDevice, st);
        return;
    }
    const float* uin[4] = {c, w0, w1, w0};
    float* uout[4] = {w0, w1, w0, c};
    for (int s = 0; s < nsteps; ++s)
        for (int k = 0; k < 4; ++k) {
            A.stage = k + 1;
            A.a = shift[k];
            A.u = uin[k];
            A.out = uout[k];
            launch_timer_begin(st);
            launch_shape(E->shape, E->grid, A, st);
            launch_timer_end(7, st);
        }
}

int phx_ensemble_rk4_steps(phx_ensemble* E, float* c, float* work, double dt, int nsteps, phx_stream s) {
    if (!E || !c || !work) return fail(PHX_ERR_INVALID, "phx_ensemble_rk4_steps: NULL argument");
    if (nsteps < 0 || !(dt > 0.0)) return fail(PHX_ERR_INVALID, "phx_ensemble_rk4_steps: bad nsteps or dt");
    enqueue_steps(E, c, work, dt, nsteps, as_stream(s));
    PHX_LAUNCH_CHECK("phx_ensemble_rk4_steps");
    return PHX_OK;
}

int phx_ensemble_graph_steps(phx_ensemble* E, float* c, float* work, double dt, int steps_per_graph, int nreplays,
                             phx_stream s) {
    if (!E || !c || !work || steps_per_graph <= 0 || nreplays < 0 || !(dt > 0.0))
        return fail(PHX_ERR_INVALID, "phx_ensemble_graph_steps: bad arguments");
    hipStream_t st = as_stream(s);
    const void* key[4] = {c, work, (const void*)(intptr_t)steps_per_graph, (const void*)(intptr_t)E->shape};
    if (!E->graph || std::memcmp(key, E->graph_key, sizeof(key)) != 0 || E->dt != dt) {
        if (E->graph) {
            (void)hipGraphExecDestroy(E->graph);
            E->graph = nullptr;
        }
        hipStream_t cap;
        PHX_HIP(hipStreamCreateWithFlags(&cap, hipStreamNonBlocking));
        hipGraph_t g;
        PHX_HIP(hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal));
        enqueue_steps(E, c, work, dt, steps_per_graph, cap);
        PHX_HIP(hipStreamEndCapture(cap, &g));
        PHX_HIP(hipGraphInstantiate(&E->graph, g, nullptr, nullptr, 0));
        (void)hipGraphDestroy(g);
        (void)hipStreamDestroy(cap);
        std::memcpy(E->graph_key, key, sizeof(key));
        E->dt = dt;
        E->graph_steps = steps_per_graph;
    }
    for (int r = 0; r < nreplays; ++r) PHX_HIP(hipGraphLaunch(E->graph, st));
    return PHX_OK;
}

int phx_ensemble_mass(const phx_ensemble* E, const float* c, double* out_dev, phx_stream s) {
    if (!E || !c || !out_dev) return fail(PHX_ERR_INVALID, "phx_ensemble_mass: NULL argument");
    hipLaunchKernelGGL(k_ens_mass, dim3(E->nruns), dim3(256), 0, as_stream(s), c, E->d_runs, out_dev);
    PHX_LAUNCH_CHECK("phx_ensemble_mass");
    return PHX_OK;
}

// Times every (shape, grid) candidate on `steps` RK4 steps of the caller's
// state (which it advances: pass a scratch copy) and keeps the fastest.
int phx_ensemble_autotune(phx_ensemble* E, float* c, float* work, double dt, int steps, double* best_ms,
                          phx_stream s) {
    if (!E || !c || !work || steps <= 0) return fail(PHX_ERR_INVALID, "phx_ensemble_autotune: bad arguments");
    hipStream_t st = as_stream(s);
    hipEvent_t e0, e1;
    PHX_HIP(hipEventCreate(&e0));
    PHX_HIP(hipEventCreate(&e1));
    float best = 1e30f;
    int bs = 0, bg = 0, b2 = 0;
    // candidates: kernel (two-stage, one-stage) x shape x persistent grid
    const int mult[] = {2, 4, 8, 16};
    for (int two = 1; two >= 0; --two)
    for (int sh = 0; sh < kNumShapes; ++sh)
        for (int m : mult) {
            int grid = std::min(E->ntiles[sh], m * E->cus);
            E->two_stage = two;
            E->shape = sh;
            E->grid = grid;
            enqueue_steps(E, c, work, dt, 1, st);   // warm
            PHX_HIP(hipEventRecord(e0, st));
            enqueue_steps(E, c, work, dt, steps, st);
            PHX_HIP(hipEventRecord(e1, st));
            PHX_HIP(hipEventSynchronize(e1));
            float ms = 0.f;
            PHX_HIP(hipEventElapsedTime(&ms, e0, e1));
            if (ms < best) {
                best = ms;
                bs = sh;
                bg = grid;
                b2 = E->two_stage;
            }
        }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (E->graph) {
        (void)hipGraphExecDestroy(E->graph);
        E->graph = nullptr;
    }
    E->shape = bs;
    E->grid = bg;
    E->two_stage = b2;
    if (best_ms) *best_ms = best / steps;
    PHX_LAUNCH_CHECK("phx_ensemble_autotune");
    return PHX_OK;
}

}  // extern "C"
