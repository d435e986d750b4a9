// reduce.hip -- FP64 compensated mass and free-energy reductions.
//
// The reference has only a host mean (src/field/ScalarField.cu:260-280) and a
// numpy energy (develop/Spinodal Decomposition/computeFreeEnergy.py:76-98).
// Cahn-Hilliard conserves mass to round-off, so the diagnostics accumulate in
// FP64 with Neumaier-compensated partial sums whatever the storage precision:
//   pass 1: each block takes whole x-rows and each thread accumulates
//           (sum, compensation) pairs over its cells of them; pairs merge
//           exactly (TwoSum) across the wave by shuffles and across the
//           block in LDS; one pair per block.
//   pass 2: one block merges the per-block pairs in a fixed order.
// The grid size depends only on the grid's shape, so results are bitwise
// deterministic run to run.
#pragma clang fp contract(off)

#include "common.hpp"

namespace phx {

constexpr int RED_THREADS = 256;
constexpr int RED_MAX_BLOCKS = 1024;

struct NSum {
    double s, c;
};

__device__ __forceinline__ void nadd(NSum& a, double x) {
    double t = a.s + x;
    if (fabs(a.s) >= fabs(x)) a.c += (a.s - t) + x;
    else a.c += (x - t) + a.s;
    a.s = t;
}

__device__ __forceinline__ NSum nmerge(NSum a, NSum b) {
    NSum r;
    r.s = a.s + b.s;
    double e = fabs(a.s) >= fabs(b.s) ? (a.s - r.s) + b.s : (b.s - r.s) + a.s;
    r.c = a.c + b.c + e;
    return r;
}

__device__ __forceinline__ NSum wave_merge(NSum v) {
    for (int off = 32; off >= 1; off >>= 1) {
        NSum o;
        o.s = __shfl_xor(v.s, off, 64);
        o.c = __shfl_xor(v.c, off, 64);
        v = nmerge(v, o);
    }
    return v;
}

template <typename T>
struct RedArgs {
    const T* c;
    Layout L;
    int bk[3][2];
    double bv[3][2];
    double rho, ca, cb, kappa;
    double twod[3];
    long long ncells;
    long long bstride;   // elements between the fields of a batch
};

// Neighbour index along one axis for the central difference: p off the grid
// maps by the face's condition (PERIODIC wraps, NOFLUX takes the face cell);
// FIXED returns -1 (the constant is used instead).
__device__ __forceinline__ int nb_index(int p, int n, int klo, int khi) {
    if (p >= 0 && p < n) return p;
    const int kind = p < 0 ? klo : khi;
    if (kind == PHX_BC_FIXED) return -1;
    if (kind == PHX_BC_PERIODIC) return p < 0 ? p + n : p - n;
    return p < 0 ? 0 : n - 1;
}

// Pass 1 over rows: a block takes the x-rows (j, k) = blockIdx.x, +gridDim.x,
// ... (row index r = j + n1 k), and its threads the cells x = tid, +256, ...
// of each row. The neighbour rows (j +- 1, k +- 1, BC-mapped) are resolved
// once per row, so a cell costs its loads and the arithmetic: no 64-bit
// division or BC test per cell except at the row ends.
template <typename T>
__global__ __launch_bounds__(RED_THREADS) void k_mass_energy(RedArgs<T> A, double* __restrict__ part) {
    __shared__ NSum sh[2][RED_THREADS / 64];
    NSum m = {0.0, 0.0}, e = {0.0, 0.0};
    const int n0 = A.L.n0, n1 = A.L.n1, n2 = A.L.n2, dim = A.L.dim;
    const long long rows = (long long)n1 * n2;
    const double kh = 0.5 * A.kappa;
    const T* cf = A.c + (long long)blockIdx.y * A.bstride;   // field blockIdx.y of a batch
    part += 4LL * gridDim.x * blockIdx.y;
    for (long long r = blockIdx.x; r < rows; r += gridDim.x) {
        const int j = (int)(r % n1), k = (int)(r / n1);
        const T* row = cf + A.L.at(0, j, k);
        // neighbour rows along y and z (nullptr: FIXED, the constant instead)
        const T* ny[2] = {nullptr, nullptr};
        const T* nz[2] = {nullptr, nullptr};
        double vy[2] = {0.0, 0.0}, vz[2] = {0.0, 0.0};
        if (dim >= 2)
#pragma unroll
            for (int sd = 0; sd < 2; ++sd) {
                const int q = nb_index(j + (sd ? 1 : -1), n1, A.bk[1][0], A.bk[1][1]);
                if (q < 0) vy[sd] = A.bv[1][sd];
                else ny[sd] = cf + A.L.at(0, q, k);
            }
        if (dim >= 3)
#pragma unroll
            for (int sd = 0; sd < 2; ++sd) {
                const int q = nb_index(k + (sd ? 1 : -1), n2, A.bk[2][0], A.bk[2][1]);
                if (q < 0) vz[sd] = A.bv[2][sd];
                else nz[sd] = cf + A.L.at(0, j, q);
            }
        for (int i = threadIdx.x; i < n0; i += RED_THREADS) {
            const double v = (double)row[i];
            nadd(m, v);
            const double fchem = A.rho * (v - A.ca) * (v - A.ca) * (A.cb - v) * (A.cb - v);
            double xp, xm;
            if (i > 0 && i < n0 - 1) {
                xp = (double)row[i + 1];
                xm = (double)row[i - 1];
            } else {
                const int qp = nb_index(i + 1, n0, A.bk[0][0], A.bk[0][1]);
                const int qm = nb_index(i - 1, n0, A.bk[0][0], A.bk[0][1]);
                xp = qp < 0 ? A.bv[0][1] : (double)row[qp];
                xm = qm < 0 ? A.bv[0][0] : (double)row[qm];
            }
            const double dx = (xp - xm) / A.twod[0];
            double gs = dx * dx;
            if (dim >= 2) {
                const double yp = ny[1] ? (double)ny[1][i] : vy[1], ym = ny[0] ? (double)ny[0][i] : vy[0];
                const double dy = (yp - ym) / A.twod[1];
                gs += dy * dy;
            }
            if (dim >= 3) {
                const double zp = nz[1] ? (double)nz[1][i] : vz[1], zm = nz[0] ? (double)nz[0][i] : vz[0];
                const double dz = (zp - zm) / A.twod[2];
                gs += dz * dz;
            }
            nadd(e, fchem + kh * gs);
        }
    }
    m = wave_merge(m);
    e = wave_merge(e);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) {
        sh[0][wid] = m;
        sh[1][wid] = e;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        NSum tm = sh[0][0], te = sh[1][0];
        for (int w = 1; w < RED_THREADS / 64; ++w) {
            tm = nmerge(tm, sh[0][w]);
            te = nmerge(te, sh[1][w]);
        }
        part[4 * blockIdx.x + 0] = tm.s;
        part[4 * blockIdx.x + 1] = tm.c;
        part[4 * blockIdx.x + 2] = te.s;
        part[4 * blockIdx.x + 3] = te.c;
    }
}

__global__ __launch_bounds__(64) void k_mass_energy_final(const double* __restrict__ part, int nb, double dV,
                                                          double* __restrict__ out) {
    // single wave: lane l merges blocks l, l+64, ... in order, then the wave
    // merges; block b of the grid finishes field b of a batch
    part += 4LL * nb * blockIdx.x;
    out += 2 * blockIdx.x;
    NSum m = {0.0, 0.0}, e = {0.0, 0.0};
    for (int b = threadIdx.x; b < nb; b += 64) {
        m = nmerge(m, NSum{part[4 * b], part[4 * b + 1]});
        e = nmerge(e, NSum{part[4 * b + 2], part[4 * b + 3]});
    }
    m = wave_merge(m);
    e = wave_merge(e);
    if (threadIdx.x == 0) {
        out[0] = m.s + m.c;
        out[1] = (e.s + e.c) * dV;
    }
}

// One block per x-row up to RED_MAX_BLOCKS (a function of the shape only:
// bitwise deterministic).
static int red_blocks(const phx_grid* G) {
    long long b = (long long)G->n[1] * G->n[2];
    if (b < 1) b = 1;
    if (b > RED_MAX_BLOCKS) b = RED_MAX_BLOCKS;
    return (int)b;
}

template <typename T>
static int mass_energy_t(const phx_grid* G, const T* c, const phx_ch_params* p, double* work, double* out,
                         hipStream_t st, int nbatch = 1, long long bstride = 0) {
    RedArgs<T> A;
    A.c = c;
    A.bstride = bstride;
    A.L = make_layout(G);
    for (int a = 0; a < 3; ++a)
        for (int s = 0; s < 2; ++s) {
            A.bk[a][s] = a < G->dim ? G->bc[a][s] : PHX_BC_PERIODIC;
            A.bv[a][s] = G->bc_value[a][s];
        }
    A.rho = p->rho;
    A.ca = p->ca;
    A.cb = p->cb;
    A.kappa = p->kappa;
    double dV = 1.0;
    for (int a = 0; a < 3; ++a) {
        A.twod[a] = a < G->dim ? 2.0 * G->d[a] : 1.0;
        if (a < G->dim) dV *= G->d[a];
    }
    A.ncells = (long long)G->n[0] * G->n[1] * G->n[2];
    int nb = red_blocks(G);
    hipLaunchKernelGGL(k_mass_energy<T>, dim3(nb, nbatch), dim3(RED_THREADS), 0, st, A, work);
    hipLaunchKernelGGL(k_mass_energy_final, dim3(nbatch), dim3(64), 0, st, work, nb, dV, out);
    PHX_LAUNCH_CHECK("phx_mass_energy");
    return PHX_OK;
}

}  // namespace phx

using namespace phx;

extern "C" {

size_t phx_reduce_workspace_bytes(const phx_grid* G) {
    if (check_grid(G, "phx_reduce_workspace_bytes")) return 0;
    return (size_t)4 * sizeof(double) * RED_MAX_BLOCKS;
}

int phx_mass_energy(const phx_grid* G, int dtype, const void* c, const phx_ch_params* p, void* work,
                    double* out_dev, phx_stream s) {
    int rc = check_grid(G, "phx_mass_energy");
    if (rc) return rc;
    if ((rc = check_dtype(dtype, "phx_mass_energy"))) return rc;
    if (!c || !p || !work || !out_dev) return fail(PHX_ERR_INVALID, "phx_mass_energy: NULL argument");
    for (int a = 0; a < G->dim; ++a) {
        int lo = G->bc[a][0], hi = G->bc[a][1];
        if (lo < PHX_BC_PERIODIC || lo > PHX_BC_FIXED || hi < PHX_BC_PERIODIC || hi > PHX_BC_FIXED)
            return fail(PHX_ERR_UNSUPPORTED, "phx_mass_energy: every face of the active axes needs a condition");
    }
    return dtype == PHX_F64
               ? mass_energy_t<double>(G, (const double*)c, p, (double*)work, out_dev, as_stream(s))
               : mass_energy_t<float>(G, (const float*)c, p, (double*)work, out_dev, as_stream(s));
}

int phx_mass_energy_batch(const phx_grid* G, int dtype, int nbatch, const void* c, const phx_ch_params* p,
                          void* work, double* out_dev, phx_stream s) {
    int rc = check_grid(G, "phx_mass_energy_batch");
    if (rc) return rc;
    if ((rc = check_dtype(dtype, "phx_mass_energy_batch"))) return rc;
    if (!c || !p || !work || !out_dev) return fail(PHX_ERR_INVALID, "phx_mass_energy_batch: NULL argument");
    if (nbatch < 1 || nbatch > 65535) return fail(PHX_ERR_INVALID, "phx_mass_energy_batch: nbatch must be in [1, 65535]");
    const long long bs = make_layout(G).total();
    for (int a = 0; a < G->dim; ++a) {
        int lo = G->bc[a][0], hi = G->bc[a][1];
        if (lo < PHX_BC_PERIODIC || lo > PHX_BC_FIXED || hi < PHX_BC_PERIODIC || hi > PHX_BC_FIXED)
            return fail(PHX_ERR_UNSUPPORTED, "phx_mass_energy_batch: every face of the active axes needs a condition");
    }
    return dtype == PHX_F64
               ? mass_energy_t<double>(G, (const double*)c, p, (double*)work, out_dev, as_stream(s), nbatch, bs)
               : mass_energy_t<float>(G, (const float*)c, p, (double*)work, out_dev, as_stream(s), nbatch, bs);
}

}  // extern "C"
