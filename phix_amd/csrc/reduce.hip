// reduce.hip -- FP64 compensated mass and free-energy reductions.
//
// The reference has only a host mean (src/field/ScalarField.cu:260-280) and a
// numpy energy (develop/Spinodal Decomposition/computeFreeEnergy.py:76-98).
// Cahn-Hilliard conserves mass to round-off, so the diagnostics accumulate in
// FP64 with Neumaier-compensated partial sums whatever the storage precision:
//   pass 1: each thread accumulates (sum, compensation) pairs over a
//           grid-stride slice; pairs merge exactly (TwoSum) across the wave
//           by shuffles and across the block in LDS; one pair per block.
//   pass 2: one block merges the per-block pairs in a fixed order.
// The grid size depends only on the cell count, so results are bitwise
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
};

template <typename T>
__device__ __forceinline__ double cval(const RedArgs<T>& A, int q0, int q1, int q2) {
    return (double)A.c[A.L.at(q0, q1, q2)];
}

// c at physical q shifted by +-1 along `axis`, off-grid neighbours by the BC table.
template <typename T>
__device__ __forceinline__ double nbv(const RedArgs<T>& A, int q[3], int axis, int step, const int n[3]) {
    int r[3] = {q[0], q[1], q[2]};
    int p = q[axis] + step;
    if (p < 0 || p >= n[axis]) {
        int side = p < 0 ? 0 : 1, kind = A.bk[axis][side];
        if (kind == PHX_BC_FIXED) return A.bv[axis][side];
        if (kind == PHX_BC_PERIODIC) p = p < 0 ? p + n[axis] : p - n[axis];
        else p = p < 0 ? 0 : n[axis] - 1;
    }
    r[axis] = p;
    return cval(A, r[0], r[1], r[2]);
}

template <typename T>
__global__ __launch_bounds__(RED_THREADS) void k_mass_energy(RedArgs<T> A, double* __restrict__ part) {
    __shared__ NSum sh[2][RED_THREADS / 64];
    NSum m = {0.0, 0.0}, e = {0.0, 0.0};
    const int n[3] = {A.L.n0, A.L.n1, A.L.n2};
    const long long stride = (long long)gridDim.x * RED_THREADS;
    for (long long q = blockIdx.x * (long long)RED_THREADS + threadIdx.x; q < A.ncells; q += stride) {
        int i = (int)(q % n[0]);
        long long r = q / n[0];
        int j = (int)(r % n[1]);
        int k = (int)(r / n[1]);
        int qq[3] = {i, j, k};
        double v = cval(A, i, j, k);
        nadd(m, v);
        double fchem = A.rho * (v - A.ca) * (v - A.ca) * (A.cb - v) * (A.cb - v);
        double gs = 0.0;
        for (int a = 0; a < A.L.dim; ++a) {
            double dd = (nbv(A, qq, a, +1, n) - nbv(A, qq, a, -1, n)) / A.twod[a];
            gs += dd * dd;
        }
        nadd(e, fchem + 0.5 * A.kappa * gs);
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
    // single wave: lane l merges blocks l, l+64, ... in order, then the wave merges.
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

static int red_blocks(long long ncells) {
    long long b = (ncells + RED_THREADS * 8 - 1) / (RED_THREADS * 8);
    if (b < 1) b = 1;
    if (b > RED_MAX_BLOCKS) b = RED_MAX_BLOCKS;
    return (int)b;
}

template <typename T>
static int mass_energy_t(const phx_grid* G, const T* c, const phx_ch_params* p, double* work, double* out,
                         hipStream_t st) {
    RedArgs<T> A;
    A.c = c;
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
    int nb = red_blocks(A.ncells);
    hipLaunchKernelGGL(k_mass_energy<T>, dim3(nb), dim3(RED_THREADS), 0, st, A, work);
    hipLaunchKernelGGL(k_mass_energy_final, dim3(1), dim3(64), 0, st, work, nb, dV, out);
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

}  // extern "C"
