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
//     NOFLUX: mu at ghost -1 equals mu(0); with the mirrored halo the same
//     operands appear with the +/- neighbours of the ghost axis exchanged, so
//     the kernel exchanges them back -- again bit for bit.
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

namespace phx {

// ------------------------------------------------------------- parameters --
template <typename T>
struct CHArgs {
    const T* u;     // stage state (stencil input)
    const T* c;     // c_n
    T* acc;         // RK4 accumulator
    T* out;         // next stage state, or c for stage 4
    Layout L;
    AxisBC<T> bx, by, bz;
    T idx2, idy2, idz2;
    T two_rho, ca, cb, nkappa, M;
    T a;            // stage shift dt/2, dt/2, dt
    T dt6, dt3;
    int stage;      // 1..4
    int zc;         // z planes per block (3D)
    int xcd_remap;  // 1: XCD-aware tile order (grid size a multiple of 8)
};

template <typename T>
__device__ __forceinline__ T dwell(T v, T two_rho, T ca, T cb) {
    // 2.0*rho*(c-ca)*(c-cb)*(2.0*c-ca-cb), evaluated left to right
    return two_rho * (v - ca) * (v - cb) * ((T)2 * v - ca - cb);
}

// Halo value of the stage state at position (px,py,pz) in [-2, n+1]^dim, with
// the BC maps. FIXED faces return their constant; z, then y, then x decide
// first, the reverse of the X, Y, Z order in which the reference fills faces.
template <typename T, int DIM>
__device__ __forceinline__ T load_halo(const T* __restrict__ u, const CHArgs<T>& A, int px, int py, int pz) {
    int qz = 0, qy = 0, qx;
    if (DIM >= 3) {
        qz = map_in(pz, A.bz);
        if (qz < 0) return fixed_value(pz, A.bz);
    }
    if (DIM >= 2) {
        qy = map_in(py, A.by);
        if (qy < 0) return fixed_value(py, A.by);
    }
    qx = map_in(px, A.bx);
    if (qx < 0) return fixed_value(px, A.bx);
    return u[A.L.at(qx, qy, qz)];
}

// mu at halo position p from the stage state's 7 (5) point star.
//  c: centre; x/y/z (p,m): the +/- neighbours as loaded.
template <typename T, int DIM>
__device__ __forceinline__ T mu_at(const CHArgs<T>& A, int px, int py, int pz, T c, T xp, T xm, T yp, T ym, T zp,
                                   T zm) {
    const bool gx = px < 0 || px >= A.L.n0;
    const bool gy = DIM >= 2 && (py < 0 || py >= A.L.n1);
    const bool gz = DIM >= 3 && (pz < 0 || pz >= A.L.n2);
    if (gz) {
        int k = pz < 0 ? A.bz.klo : A.bz.khi;
        if (k == PHX_BC_FIXED) return fixed_value(pz, A.bz);
        if (k == PHX_BC_NOFLUX) { T t = zp; zp = zm; zm = t; }
    }
    if (gy) {
        int k = py < 0 ? A.by.klo : A.by.khi;
        if (k == PHX_BC_FIXED) return fixed_value(py, A.by);
        if (k == PHX_BC_NOFLUX) { T t = yp; yp = ym; ym = t; }
    }
    if (gx) {
        int k = px < 0 ? A.bx.klo : A.bx.khi;
        if (k == PHX_BC_FIXED) return fixed_value(px, A.bx);
        if (k == PHX_BC_NOFLUX) { T t = xp; xp = xm; xm = t; }
    }
    T val = (xp - (T)2 * c + xm) * A.idx2;
    if (DIM >= 2) val += (yp - (T)2 * c + ym) * A.idy2;
    if (DIM >= 3) val += (zp - (T)2 * c + zm) * A.idz2;
    return dwell(c, A.two_rho, A.ca, A.cb) + A.nkappa * val;
}

// Stage epilogue for one physical output cell.
template <typename T, bool EXACT>
__device__ __forceinline__ void epilogue(const CHArgs<T>& A, long long gi, T uc, T k) {
    const int s = A.stage;
    if (EXACT) {
        if (s == 1) {
            A.acc[gi] = k;
            A.out[gi] = uc + A.a * k;
        } else if (s < 4) {
            A.acc[gi] = A.acc[gi] + (T)2 * k;
            A.out[gi] = A.c[gi] + A.a * k;
        } else {
            A.out[gi] = A.c[gi] + A.dt6 * (A.acc[gi] + k);
        }
    } else {
        if (s == 1) {
            A.acc[gi] = uc + A.dt6 * k;
            A.out[gi] = uc + A.a * k;
        } else if (s < 4) {
            A.acc[gi] = A.acc[gi] + A.dt3 * k;
            A.out[gi] = A.c[gi] + A.a * k;
        } else {
            A.out[gi] = A.acc[gi] + A.dt6 * k;
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
    const int tid = threadIdx.x;
    const int x0 = blockIdx.x * TX, y0 = blockIdx.y * TY;
    const int kz = 0;

    const bool interior = (x0 >= 2) && (x0 + TX + 1 < A.L.n0) && (DIM < 2 || ((y0 >= 2) && (y0 + TY + 1 < A.L.n1)));
    if (interior) {
        const T* base = A.u + A.L.at(x0 - 2, y0 - 2, kz);
        for (int t = tid; t < UX * UY; t += 256) {
            int ly = t / UX, lx = t - ly * UX;
            U[ly][lx] = base[(long long)ly * A.L.sx + lx];
        }
    } else {
        for (int t = tid; t < UX * UY; t += 256) {
            int ly = t / UX, lx = t - ly * UX;
            int py = DIM >= 2 ? y0 - 2 + ly : 0;
            U[ly][lx] = load_halo<T, DIM>(A.u, A, x0 - 2 + lx, py, kz);
        }
    }
    __syncthreads();
    for (int t = tid; t < MX * MY; t += 256) {
        int my = t / MX, mx = t - my * MX;
        int px = x0 - 1 + mx, py = DIM >= 2 ? y0 - 1 + my : 0;
        T c = U[my + 1][mx + 1];
        T m = mu_at<T, DIM>(A, px, py, kz, c, U[my + 1][mx + 2], U[my + 1][mx], DIM >= 2 ? U[my + 2][mx + 1] : c,
                            DIM >= 2 ? U[my][mx + 1] : c, c, c);
        MU[my][mx] = m;
    }
    __syncthreads();
    for (int t = tid; t < TX * TY; t += 256) {
        int oy = t / TX, ox = t - oy * TX;
        int x = x0 + ox, y = y0 + oy;
        if (x >= A.L.n0 || y >= A.L.n1) continue;
        T m = MU[oy + 1][ox + 1];
        T val = (MU[oy + 1][ox + 2] - (T)2 * m + MU[oy + 1][ox]) * A.idx2;
        if (DIM >= 2) val += (MU[oy + 2][ox + 1] - (T)2 * m + MU[oy][ox + 1]) * A.idy2;
        T k = A.M * val;
        epilogue<T, EXACT>(A, A.L.at(x, y, kz), U[oy + 2][ox + 2], k);
    }
}

// ------------------------------------------------------------------- 3D ----
// One block = TX x TY columns of a z-chunk of A.zc planes, streaming along z
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
//   by output o-1 before barrier A of iteration o.
template <typename T, int TX, int TY, bool EXACT>
__global__ __launch_bounds__(256) void k_ch_stage_3d(CHArgs<T> A) {
    constexpr int UX = TX + 4, UY = TY + 4, MX = TX + 2, MY = TY + 2;
    constexpr int NU = (UX * UY + 255) / 256;
    constexpr int NM = (MX * MY + 255) / 256;
    constexpr int NO = (TX * TY) / 256;
    static_assert((TX * TY) % 256 == 0, "tile must be a multiple of 256 cells");
    __shared__ T U[4][UY * UX];
    __shared__ T MU[3][MY * MX];
    const int tid = threadIdx.x;
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
    const int z0 = bz * A.zc;
    const int z1 = min(z0 + A.zc, A.L.n2);
    const int g = A.L.g;
    const long long sxy = A.L.sxy;

    // halo slots: in-plane offset (x/y maps applied) or -1 for a FIXED constant
    int hoff[NU];
    T hfix[NU];
#pragma unroll
    for (int i = 0; i < NU; ++i) {
        const int t = tid + 256 * i;
        hoff[i] = 0;
        hfix[i] = (T)0;
        if (t < UX * UY) {
            const int ly = t / UX, lx = t - ly * UX;
            const int px = x0 - 2 + lx, py = y0 - 2 + ly;
            const int qy = map_in(py, A.by), qx = map_in(px, A.bx);
            if (qy < 0) {
                hoff[i] = -1;
                hfix[i] = fixed_value(py, A.by);
            } else if (qx < 0) {
                hoff[i] = -1;
                hfix[i] = fixed_value(px, A.bx);
            } else {
                hoff[i] = (qy + g) * A.L.sx + (qx + g);
            }
        }
    }
    // output slots
    int ooff[NO];
    bool oval[NO];
#pragma unroll
    for (int j = 0; j < NO; ++j) {
        const int t = tid + 256 * j;
        const int oy = t / TX, ox = t - oy * TX;
        const int x = x0 + ox, y = y0 + oy;
        oval[j] = x < A.L.n0 && y < A.L.n1;
        ooff[j] = (y + g) * A.L.sx + (x + g);
    }

    auto load_plane = [&](int pz, T (&r)[NU]) {
        const int qz = map_in(pz, A.bz);
        if (qz < 0) {
            const T v = fixed_value(pz, A.bz);
#pragma unroll
            for (int i = 0; i < NU; ++i) r[i] = v;
        } else {
            const T* base = A.u + (long long)(qz + g) * sxy;
#pragma unroll
            for (int i = 0; i < NU; ++i)
                if (tid + 256 * i < UX * UY) r[i] = hoff[i] >= 0 ? base[hoff[i]] : hfix[i];
        }
    };
    auto store_plane = [&](int pz, const T (&r)[NU]) {
        T* dst = U[(pz + 4) & 3];
#pragma unroll
        for (int i = 0; i < NU; ++i)
            if (tid + 256 * i < UX * UY) dst[tid + 256 * i] = r[i];
    };
    // mu slots: LDS centre offset in the halo plane, and halo position
    int mc0[NM], mpx[NM], mpy[NM];
#pragma unroll
    for (int i = 0; i < NM; ++i) {
        const int t = tid + 256 * i;
        const int my = t / MX, mx = t - my * MX;
        mc0[i] = (my + 1) * UX + mx + 1;
        mpx[i] = x0 - 1 + mx;
        mpy[i] = y0 - 1 + my;
    }
    // tiles whose mu ring lies inside the domain in x and y need no BC logic
    const bool xy_in = x0 >= 1 && x0 + TX < A.L.n0 && y0 >= 1 && y0 + TY < A.L.n1;
    auto compute_mu = [&](int m) {
        const T* ul = U[(m + 3) & 3];
        const T* um = U[(m + 4) & 3];
        const T* uh = U[(m + 5) & 3];
        T* dst = MU[(m + 3) % 3];
        if (xy_in && m >= 0 && m < A.L.n2) {
#pragma unroll
            for (int i = 0; i < NM; ++i) {
                const int t = tid + 256 * i;
                if (t < MX * MY) {
                    const int c0 = mc0[i];
                    const T c = um[c0];
                    T val = (um[c0 + 1] - (T)2 * c + um[c0 - 1]) * A.idx2;
                    val += (um[c0 + UX] - (T)2 * c + um[c0 - UX]) * A.idy2;
                    val += (uh[c0] - (T)2 * c + ul[c0]) * A.idz2;
                    dst[t] = dwell(c, A.two_rho, A.ca, A.cb) + A.nkappa * val;
                }
            }
        } else {
#pragma unroll
            for (int i = 0; i < NM; ++i) {
                const int t = tid + 256 * i;
                if (t < MX * MY) {
                    const int c0 = mc0[i];
                    dst[t] = mu_at<T, 3>(A, mpx[i], mpy[i], m, um[c0], um[c0 + 1], um[c0 - 1], um[c0 + UX],
                                         um[c0 - UX], uh[c0], ul[c0]);
                }
            }
        }
    };
    // epilogue operands of one output plane
    const int st = A.stage;
    const bool need_c = st == 2 || st == 3 || (EXACT && st == 4);
    const bool need_a = st >= 2;
    T pc[NO], pa[NO], nc[NO], na[NO];
    auto load_epi = [&](int o, T (&rc)[NO], T (&ra)[NO]) {
        const long long pb = (long long)(o + g) * sxy;
#pragma unroll
        for (int j = 0; j < NO; ++j) {
            if (!oval[j]) continue;
            if (need_c) rc[j] = A.c[pb + ooff[j]];
            if (need_a) ra[j] = A.acc[pb + ooff[j]];
        }
    };

    // ---- prologue: halo planes z0-2 .. z0+1, mu planes z0-1, z0
    T pre[NU];
    load_plane(z0 - 2, pre);
    store_plane(z0 - 2, pre);
    load_plane(z0 - 1, pre);
    store_plane(z0 - 1, pre);
    load_plane(z0, pre);
    store_plane(z0, pre);
    load_plane(z0 + 1, pre);
    store_plane(z0 + 1, pre);
    load_plane(z0 + 2, pre);
    load_epi(z0, pc, pa);
    __syncthreads();
    compute_mu(z0 - 1);
    compute_mu(z0);
    __syncthreads();

    for (int o = z0; o < z1; ++o) {
        store_plane(o + 2, pre);
        __syncthreads();   // A
        if (o + 3 <= z1 + 1) load_plane(o + 3, pre);
        if (o + 1 < z1) load_epi(o + 1, nc, na);
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
            T val = (mmid[m0 + 1] - (T)2 * mc + mmid[m0 - 1]) * A.idx2;
            val += (mmid[m0 + MX] - (T)2 * mc + mmid[m0 - MX]) * A.idy2;
            val += (mhi[m0] - (T)2 * mc + mlo[m0]) * A.idz2;
            const T k = A.M * val;
            const long long gi = pb + ooff[j];
            const T uc = uo[(oy + 2) * UX + ox + 2];
            if (EXACT) {
                if (st == 1) {
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
                    A.acc[gi] = uc + A.dt6 * k;
                    A.out[gi] = uc + A.a * k;
                } else if (st < 4) {
                    A.acc[gi] = pa[j] + A.dt3 * k;
                    A.out[gi] = pc[j] + A.a * k;
                } else {
                    A.out[gi] = pa[j] + A.dt6 * k;
                }
            }
        }
#pragma unroll
        for (int j = 0; j < NO; ++j) {
            pc[j] = nc[j];
            pa[j] = na[j];
        }
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

template <typename T>
static int euler_steps_t(const phx_grid* G, T* c, T* mu, const phx_ch_params* p, int nsteps, hipStream_t st) {
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

// Tile shapes (autotuned later; these are the defaults).
constexpr int T2X = 64, T2Y = 16;
// 3D tile shapes the autotuner chooses from (TX x TY columns, 256 threads).
struct Tile3 {
    int tx, ty;
};
static const Tile3 kTiles3[] = {{64, 16}, {64, 8}, {128, 8}, {32, 16}, {64, 4}};
constexpr int kNumTiles3 = sizeof(kTiles3) / sizeof(kTiles3[0]);

template <typename T, bool EXACT>
static void launch3(int tile, dim3 grid, const CHArgs<T>& A, hipStream_t st) {
    switch (tile) {
        case 0: hipLaunchKernelGGL((k_ch_stage_3d<T, 64, 16, EXACT>), grid, dim3(256), 0, st, A); break;
        case 1: hipLaunchKernelGGL((k_ch_stage_3d<T, 64, 8, EXACT>), grid, dim3(256), 0, st, A); break;
        case 2: hipLaunchKernelGGL((k_ch_stage_3d<T, 128, 8, EXACT>), grid, dim3(256), 0, st, A); break;
        case 3: hipLaunchKernelGGL((k_ch_stage_3d<T, 32, 16, EXACT>), grid, dim3(256), 0, st, A); break;
        default: hipLaunchKernelGGL((k_ch_stage_3d<T, 64, 4, EXACT>), grid, dim3(256), 0, st, A); break;
    }
}

static int g_xcd_remap = 1;

template <typename T, bool EXACT>
static int rk4_steps_t(const phx_grid* G, T* c, T* work, const phx_ch_params* p, int nsteps, hipStream_t st,
                       int zc, int tile) {
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
    const double dt = p->dt, dt6 = dt / 6.0;
    A.dt6 = (T)dt6;
    A.dt3 = (T)(2.0 * dt6);
    A.c = c;
    A.acc = acc;
    A.zc = zc;
    A.xcd_remap = 0;
    const T shift[4] = {(T)(dt * 0.5), (T)(dt * 0.5), (T)dt, (T)0};
    const T* uin[4] = {c, w0, w1, w0};
    T* uout[4] = {w0, w1, w0, c};
    dim3 grid;
    if (G->dim == 3) {
        grid = dim3(cdiv(A.L.n0, kTiles3[tile].tx), cdiv(A.L.n1, kTiles3[tile].ty), cdiv(A.L.n2, zc));
        A.xcd_remap = g_xcd_remap && (grid.x * grid.y * grid.z) % 8 == 0;
    }
    else grid = dim3(cdiv(A.L.n0, T2X), cdiv(A.L.n1, T2Y), 1);
    for (int s = 0; s < nsteps; ++s) {
        if (s == nsteps - 1) {
            int rc = bc_fill_impl(G, sizeof(T) == 8 ? PHX_F64 : PHX_F32, c, st);
            if (rc) return rc;
        }
        for (int stg = 0; stg < 4; ++stg) {
            A.stage = stg + 1;
            A.u = uin[stg];
            A.out = uout[stg];
            A.a = shift[stg];
            if (G->dim == 3) launch3<T, EXACT>(tile, grid, A, st);
            else if (G->dim == 2) hipLaunchKernelGGL((k_ch_stage_2d<T, 2, T2X, T2Y, EXACT>), grid, dim3(256), 0, st, A);
            else hipLaunchKernelGGL((k_ch_stage_2d<T, 1, T2X, T2Y, EXACT>), grid, dim3(256), 0, st, A);
        }
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

// z planes per block of the 3D stage kernel: tuned per (shape, dtype) by
// phx_ch_rk4_autotune, else 16.
struct Launch3 {
    int zc, tile;
};
static std::mutex g_zc_mu;
static std::map<std::tuple<int, int, int, int>, Launch3> g_zc;

static Launch3 launch_for(const phx_grid* G, int dtype) {
    std::lock_guard<std::mutex> lk(g_zc_mu);
    auto it = g_zc.find(std::make_tuple(G->n[0], G->n[1], G->n[2], dtype));
    if (it != g_zc.end()) return it->second;
    return Launch3{16, dtype == PHX_F32 ? 0 : 1};
}

static void set_launch(const phx_grid* G, int dtype, Launch3 l) {
    std::lock_guard<std::mutex> lk(g_zc_mu);
    g_zc[std::make_tuple(G->n[0], G->n[1], G->n[2], dtype)] = l;
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
    l.zc = zc;
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
    if (tile < 0 || tile >= kNumTiles3) return fail(PHX_ERR_INVALID, "phx_ch_rk4_set_tile: tile index out of range");
    Launch3 l = launch_for(G, dtype);
    l.tile = tile;
    set_launch(G, dtype, l);
    return PHX_OK;
}

int phx_ch_rk4_steps(const phx_grid* G, int dtype, void* c, void* work, const phx_ch_params* p, int nsteps,
                     phx_stream s) {
    int rc = check_grid_fused(G, "phx_ch_rk4_steps");
    if (rc) return rc;
    if ((rc = check_dtype(dtype, "phx_ch_rk4_steps"))) return rc;
    if ((rc = check_ch_params(p, "phx_ch_rk4_steps"))) return rc;
    if (!c || !work) return fail(PHX_ERR_INVALID, "phx_ch_rk4_steps: NULL buffer");
    if (nsteps < 0) return fail(PHX_ERR_INVALID, "phx_ch_rk4_steps: nsteps < 0");
    if (nsteps == 0) return PHX_OK;
    const Launch3 l = G->dim == 3 ? launch_for(G, dtype) : Launch3{1, 0};
    return dtype == PHX_F64
               ? rk4_steps_t<double, true>(G, (double*)c, (double*)work, p, nsteps, as_stream(s), l.zc, l.tile)
               : rk4_steps_t<float, false>(G, (float*)c, (float*)work, p, nsteps, as_stream(s), l.zc, l.tile);
}

// Times the 3D stage kernel at z-chunks {4, 8, 16, 32, 64} on `steps` steps of
// the caller's state (advancing it) and keeps the fastest for this shape.
int phx_ch_rk4_autotune(const phx_grid* G, int dtype, void* c, void* work, const phx_ch_params* p, int steps,
                        int* best_zc, double* best_ms_per_step, phx_stream s) {
    int rc = check_grid_fused(G, "phx_ch_rk4_autotune");
    if (rc) return rc;
    if (steps < 1) return fail(PHX_ERR_INVALID, "phx_ch_rk4_autotune: steps < 1");
    hipStream_t st = as_stream(s);
    const int cand[] = {4, 8, 16, 32, 64};
    float best = 1e30f;
    int bz = 16, bt = 0;
    hipEvent_t e0, e1;
    PHX_HIP(hipEventCreate(&e0));
    PHX_HIP(hipEventCreate(&e1));
    for (int tile = 0; tile < (G->dim == 3 ? kNumTiles3 : 1); ++tile)
    for (int zc : cand) {
        if (G->dim == 3 && zc > G->n[2]) continue;
        set_launch(G, dtype, Launch3{zc, tile});
        if ((rc = phx_ch_rk4_steps(G, dtype, c, work, p, 1, s))) return rc;
        PHX_HIP(hipEventRecord(e0, st));
        if ((rc = phx_ch_rk4_steps(G, dtype, c, work, p, steps, s))) return rc;
        PHX_HIP(hipEventRecord(e1, st));
        PHX_HIP(hipEventSynchronize(e1));
        float ms = 0.f;
        PHX_HIP(hipEventElapsedTime(&ms, e0, e1));
        if (ms < best) {
            best = ms;
            bz = zc;
            bt = tile;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    set_launch(G, dtype, Launch3{bz, bt});
    if (best_zc) *best_zc = bz;
    if (best_ms_per_step) *best_ms_per_step = best / steps;
    return PHX_OK;
}

}  // extern "C"
