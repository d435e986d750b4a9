// core.hip -- error plumbing, grid checks, boundary kernels and the FP64
// per-term kernels of the generic (unfused) DSL path.
//
// The per-term kernels keep the operation order of the reference launchers
// (src/equation/Equation.cu:40-124, 386-404; include/equation/TermPW.inl:18-37)
// so the generic path agrees with the FP64 oracle bit for bit; FMA contraction is
// off in this file for that reason.
#pragma clang fp contract(off)

#include "common.hpp"

#include <cstring>

namespace phx {

static thread_local std::string g_last_error;

void set_error(const std::string& msg) { g_last_error = msg; }

int fail(int code, const std::string& msg) {
    g_last_error = msg;
    return code;
}

int check_grid(const phx_grid* G, const char* who) {
    if (!G) return fail(PHX_ERR_INVALID, std::string(who) + ": grid is NULL");
    if (G->dim < 1 || G->dim > 3) return fail(PHX_ERR_INVALID, std::string(who) + ": dim must be 1, 2 or 3");
    if (G->ghost < 0) return fail(PHX_ERR_INVALID, std::string(who) + ": ghost must be >= 0");
    for (int a = 0; a < 3; ++a) {
        if (a < G->dim) {
            if (G->n[a] <= 0) return fail(PHX_ERR_INVALID, std::string(who) + ": n[" + std::to_string(a) + "] must be > 0");
            if (!(G->d[a] > 0.0)) return fail(PHX_ERR_INVALID, std::string(who) + ": d[" + std::to_string(a) + "] must be > 0");
        } else if (G->n[a] != 1) {
            return fail(PHX_ERR_INVALID, std::string(who) + ": n[" + std::to_string(a) + "] must be 1 for unused axes");
        }
    }
    return PHX_OK;
}

int check_grid_fused(const phx_grid* G, const char* who) {
    int rc = check_grid(G, who);
    if (rc) return rc;
    if (G->ghost < 1) return fail(PHX_ERR_UNSUPPORTED, std::string(who) + ": ghost must be >= 1");
    for (int a = 0; a < G->dim; ++a) {
        int lo = G->bc[a][0], hi = G->bc[a][1];
        if (lo < PHX_BC_PERIODIC || lo > PHX_BC_FIXED || hi < PHX_BC_PERIODIC || hi > PHX_BC_FIXED)
            return fail(PHX_ERR_UNSUPPORTED, std::string(who) + ": every face of axis " + std::to_string(a) +
                                                 " needs a PERIODIC, NOFLUX or FIXED condition");
        if ((lo == PHX_BC_PERIODIC) != (hi == PHX_BC_PERIODIC))
            return fail(PHX_ERR_INVALID, std::string(who) + ": periodic condition on one side only of axis " +
                                             std::to_string(a));
        if (G->n[a] < 2) return fail(PHX_ERR_UNSUPPORTED, std::string(who) + ": active axes need n >= 2");
    }
    // the fused kernels address a stored x-y plane with 32-bit offsets
    const long long sx = G->n[0] + 2LL * G->ghost, sy = G->dim >= 2 ? G->n[1] + 2LL * G->ghost : 1;
    if (sx * sy >= (1LL << 31))
        return fail(PHX_ERR_UNSUPPORTED, std::string(who) + ": a stored x-y plane must hold fewer than 2^31 cells");
    return PHX_OK;
}

int check_dtype(int dtype, const char* who) {
    if (dtype != PHX_F64 && dtype != PHX_F32)
        return fail(PHX_ERR_INVALID, std::string(who) + ": dtype must be PHX_F64 or PHX_F32");
    return PHX_OK;
}

// ------------------------------------------------------------------ BCs ----
// One thread per (face0, face1) position, looping over ghost layers, as the
// reference face kernels do (Boundary.cu:95-195). Consecutive threads walk the
// faster stored axis of the face, so the Y and Z faces are coalesced.
template <typename T>
__global__ void k_bc_face(T* __restrict__ data, int nf0, int nf1, long long astride, long long fs0,
                          long long fs1, int na, int g, int kind, int do_lo, int do_hi, T value) {
    int t0 = blockIdx.x * blockDim.x + threadIdx.x;
    int t1 = blockIdx.y * blockDim.y + threadIdx.y;
    if (t0 >= nf0 || t1 >= nf1) return;
    long long off = t0 * fs0 + t1 * fs1;
    if (kind == PHX_BC_PERIODIC) {
        for (int l = 1; l <= g; ++l) {
            data[(g - l) * astride + off] = data[(g + na - l) * astride + off];
            data[(g + na + l - 1) * astride + off] = data[(g + l - 1) * astride + off];
        }
    } else if (kind == PHX_BC_NOFLUX) {
        if (do_lo) {
            T v = data[g * astride + off];
            for (int l = 1; l <= g; ++l) data[(g - l) * astride + off] = v;
        }
        if (do_hi) {
            T v = data[(g + na - 1) * astride + off];
            for (int l = 1; l <= g; ++l) data[(g + na + l - 1) * astride + off] = v;
        }
    } else {
        if (do_lo)
            for (int l = 1; l <= g; ++l) data[(g - l) * astride + off] = value;
        if (do_hi)
            for (int l = 1; l <= g; ++l) data[(g + na + l - 1) * astride + off] = value;
    }
}

template <typename T>
static int bc_apply_t(const phx_grid* G, T* data, int axis, int kind, int side, double value, hipStream_t st) {
    Layout L = make_layout(G);
    long long astride, fs0, fs1;
    int nf0, nf1, na;
    if (axis == 0) { astride = 1; nf0 = L.sy; nf1 = L.sz; fs0 = L.sx; fs1 = L.sxy; na = L.n0; }
    else if (axis == 1) { astride = L.sx; nf0 = L.sx; nf1 = L.sz; fs0 = 1; fs1 = L.sxy; na = L.n1; }
    else { astride = L.sxy; nf0 = L.sx; nf1 = L.sy; fs0 = 1; fs1 = L.sx; na = L.n2; }
    int lo = (side == PHX_SIDE_LOW || side == PHX_SIDE_BOTH);
    int hi = (side == PHX_SIDE_HIGH || side == PHX_SIDE_BOTH);
    dim3 blk(64, 4);
    dim3 grd(cdiv(nf0, 64), cdiv(nf1, 4));
    if (grd.y > 65535) return fail(PHX_ERR_UNSUPPORTED, "phx_bc_apply: face too large");
    hipLaunchKernelGGL(k_bc_face<T>, grd, blk, 0, st, data, nf0, nf1, astride, fs0, fs1, na, L.g, kind, lo, hi,
                       (T)value);
    PHX_LAUNCH_CHECK("phx_bc_apply");
    return PHX_OK;
}

// All faces in one launch: every ghost cell of the active axes gets the value
// the face fills leave when applied axis by axis (x, then y, then z), each
// over the full stored extent of the other axes. Walking the active axes from
// the last to the first, the first one on which the cell is a ghost decides:
// a FIXED face gives its value; otherwise the coordinate moves to its source
// (periodic: wrap; zero-flux: the face cell) and the earlier axes are looked
// at in turn. Sources have every active coordinate physical, and no such cell
// is written here, so the one pass equals the sequence. Threads enumerate the
// ghost cells once: the x slabs over the full (y, z) extent, the y slabs over
// physical x, the z slabs over physical x and y. Needs ghost <= n on every
// active axis and a BC on every active face.
struct FillGeom {
    Layout L;
    int kind[3][2];
    double value[3][2];
    long long nx_slab, ny_slab, nz_slab;   // ghost cells in each group
    long long bstride;                     // elements between the fields of a batch (blockIdx.y)
};

template <typename T>
__global__ void k_bc_fill_all(T* __restrict__ data, FillGeom F) {
    data += (long long)blockIdx.y * F.bstride;
    const Layout& L = F.L;
    const int g = L.g, g2 = 2 * g;
    long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    int c[3];
    if (q < F.nx_slab) {                     // x ghosts: (2g) x sy x sz
        const int gi = (int)(q % g2);
        const long long r = q / g2;
        c[0] = gi < g ? gi - g : L.n0 + gi - g;
        c[1] = (int)(r % L.sy) - g;
        c[2] = (int)(r / L.sy) - g;
    } else if ((q -= F.nx_slab) < F.ny_slab) {   // y ghosts: n0 x (2g) x sz
        c[0] = (int)(q % L.n0);
        const long long r = q / L.n0;
        const int gj = (int)(r % g2);
        c[1] = gj < g ? gj - g : L.n1 + gj - g;
        c[2] = (int)(r / g2) - g;
    } else if ((q -= F.ny_slab) < F.nz_slab) {   // z ghosts: n0 x n1 x (2g)
        c[0] = (int)(q % L.n0);
        const long long r = q / L.n0;
        c[1] = (int)(r % L.n1);
        const int gk = (int)(r / L.n1);
        c[2] = gk < g ? gk - g : L.n2 + gk - g;
    } else {
        return;
    }
    const long long dst = L.at(c[0], c[1], c[2]);
    const int n[3] = {L.n0, L.n1, L.n2};
    for (int a = L.dim - 1; a >= 0; --a) {
        const bool lo = c[a] < 0, hi = c[a] >= n[a];
        if (!lo && !hi) continue;
        const int s = hi ? 1 : 0;
        const int k = F.kind[a][s];
        if (k == PHX_BC_FIXED) {
            data[dst] = (T)F.value[a][s];
            return;
        }
        if (k == PHX_BC_PERIODIC) c[a] += lo ? n[a] : -n[a];
        else c[a] = lo ? 0 : n[a] - 1;
    }
    data[dst] = data[L.at(c[0], c[1], c[2])];
}

static bool fill_all_ok(const phx_grid* G) {
    for (int a = 0; a < G->dim; ++a) {
        if (G->ghost > G->n[a]) return false;
        for (int s = 0; s < 2; ++s)
            if (G->bc[a][s] != PHX_BC_PERIODIC && G->bc[a][s] != PHX_BC_NOFLUX && G->bc[a][s] != PHX_BC_FIXED)
                return false;
        if ((G->bc[a][0] == PHX_BC_PERIODIC) != (G->bc[a][1] == PHX_BC_PERIODIC)) return false;
    }
    return true;
}

template <typename T>
static int bc_fill_all_t(const phx_grid* G, T* data, hipStream_t st, int nbatch = 1) {
    FillGeom F{};
    F.L = make_layout(G);
    for (int a = 0; a < 3; ++a)
        for (int s = 0; s < 2; ++s) {
            F.kind[a][s] = G->bc[a][s];
            F.value[a][s] = G->bc_value[a][s];
        }
    const Layout& L = F.L;
    const long long g2 = 2LL * L.g;
    F.nx_slab = g2 * L.sy * L.sz;
    F.ny_slab = G->dim >= 2 ? (long long)L.n0 * g2 * L.sz : 0;
    F.nz_slab = G->dim >= 3 ? (long long)L.n0 * L.n1 * g2 : 0;
    const long long total = F.nx_slab + F.ny_slab + F.nz_slab;
    const long long blocks = (total + 255) / 256;
    if (blocks > 0x7fffffffLL) return fail(PHX_ERR_UNSUPPORTED, "phx_bc_fill: field too large");
    F.bstride = L.total();
    hipLaunchKernelGGL(k_bc_fill_all<T>, dim3((unsigned)blocks, (unsigned)nbatch), dim3(256), 0, st, data, F);
    PHX_LAUNCH_CHECK("phx_bc_fill");
    return PHX_OK;
}

int bc_fill_impl(const phx_grid* G, int dtype, void* data, hipStream_t st);

// The BC fill of nbatch fields of G's stored shape back to back: one launch
// when the one-launch fill applies, else field by field.
int bc_fill_batch_impl(const phx_grid* G, int dtype, void* data, int nbatch, hipStream_t st) {
    if (fill_all_ok(G))
        return dtype == PHX_F64 ? bc_fill_all_t(G, (double*)data, st, nbatch)
                                : bc_fill_all_t(G, (float*)data, st, nbatch);
    const size_t bytes = (size_t)make_layout(G).total() * (dtype == PHX_F64 ? 8 : 4);
    for (int b = 0; b < nbatch; ++b) {
        int rc = bc_fill_impl(G, dtype, (char*)data + b * bytes, st);
        if (rc) return rc;
    }
    return PHX_OK;
}

int bc_fill_impl(const phx_grid* G, int dtype, void* data, hipStream_t st) {
    if (fill_all_ok(G))
        return dtype == PHX_F64 ? bc_fill_all_t(G, (double*)data, st) : bc_fill_all_t(G, (float*)data, st);
    for (int a = 0; a < G->dim; ++a) {
        int klo = G->bc[a][0], khi = G->bc[a][1];
        int rc = 0;
        if (klo == PHX_BC_PERIODIC && khi == PHX_BC_PERIODIC) {
            rc = dtype == PHX_F64 ? bc_apply_t(G, (double*)data, a, PHX_BC_PERIODIC, PHX_SIDE_BOTH, 0.0, st)
                                  : bc_apply_t(G, (float*)data, a, PHX_BC_PERIODIC, PHX_SIDE_BOTH, 0.0, st);
            if (rc) return rc;
            continue;
        }
        if (klo == PHX_BC_PERIODIC || khi == PHX_BC_PERIODIC)
            return fail(PHX_ERR_INVALID, "phx_bc_fill: periodic condition on one side only of axis " + std::to_string(a));
        for (int s = 0; s < 2; ++s) {
            int k = G->bc[a][s];
            if (k == PHX_BC_NONE) continue;
            int side = s == 0 ? PHX_SIDE_LOW : PHX_SIDE_HIGH;
            rc = dtype == PHX_F64 ? bc_apply_t(G, (double*)data, a, k, side, G->bc_value[a][s], st)
                                  : bc_apply_t(G, (float*)data, a, k, side, G->bc_value[a][s], st);
            if (rc) return rc;
        }
    }
    return PHX_OK;
}

// ---------------------------------------------------------- term kernels ----
struct TermGeom {
    Layout L;
    double idx2, idy2, idz2;
};

static inline TermGeom term_geom(const phx_grid* G) {
    TermGeom t;
    t.L = make_layout(G);
    t.idx2 = 1.0 / (G->d[0] * G->d[0]);
    t.idy2 = G->dim >= 2 ? 1.0 / (G->d[1] * G->d[1]) : 0.0;
    t.idz2 = G->dim >= 3 ? 1.0 / (G->d[2] * G->d[2]) : 0.0;
    return t;
}

static inline dim3 phys_grid(const Layout& L) { return dim3(cdiv(L.n0, 64), L.n1, L.n2); }

__global__ void k_lap_acc(double* __restrict__ rhs, const double* __restrict__ src, double coeff, TermGeom t) {
    int i = blockIdx.x * 64 + threadIdx.x, j = blockIdx.y, k = blockIdx.z;
    if (i >= t.L.n0) return;
    long long c = t.L.at(i, j, k);
    double val = (src[c + 1] - 2.0 * src[c] + src[c - 1]) * t.idx2;
    if (t.L.dim >= 2) val += (src[c + t.L.sx] - 2.0 * src[c] + src[c - t.L.sx]) * t.idy2;
    if (t.L.dim >= 3) val += (src[c + t.L.sxy] - 2.0 * src[c] + src[c - t.L.sxy]) * t.idz2;
    rhs[c] += coeff * val;
}

__global__ void k_grad_acc(double* __restrict__ rhs, const double* __restrict__ src, double coeff, long long st,
                           double inv_2d, Layout L) {
    int i = blockIdx.x * 64 + threadIdx.x, j = blockIdx.y, k = blockIdx.z;
    if (i >= L.n0) return;
    long long c = L.at(i, j, k);
    rhs[c] += coeff * (src[c + st] - src[c - st]) * inv_2d;
}

__global__ void k_mul_acc(double* __restrict__ rhs, const double* __restrict__ s1, const double* __restrict__ s2,
                          double coeff, Layout L) {
    int i = blockIdx.x * 64 + threadIdx.x, j = blockIdx.y, k = blockIdx.z;
    if (i >= L.n0) return;
    long long c = L.at(i, j, k);
    rhs[c] += coeff * s1[c] * s2[c];
}

__global__ void k_pw_dw_acc(double* __restrict__ rhs, const double* __restrict__ src, double rho, double ca,
                            double cb, double coeff, Layout L) {
    int i = blockIdx.x * 64 + threadIdx.x, j = blockIdx.y, k = blockIdx.z;
    if (i >= L.n0) return;
    long long c = L.at(i, j, k);
    double v = src[c];
    rhs[c] += coeff * (2.0 * rho * (v - ca) * (v - cb) * (2.0 * v - ca - cb));
}

// Whole-array kernels: grid-stride over 64-bit n.
__global__ void k_fill(double* __restrict__ d, double v, size_t n) {
    for (size_t q = blockIdx.x * (size_t)blockDim.x + threadIdx.x; q < n; q += (size_t)gridDim.x * blockDim.x) d[q] = v;
}
__global__ void k_copy(double* __restrict__ d, const double* __restrict__ s, size_t n) {
    for (size_t q = blockIdx.x * (size_t)blockDim.x + threadIdx.x; q < n; q += (size_t)gridDim.x * blockDim.x) d[q] = s[q];
}
__global__ void k_axpy(double* __restrict__ d, const double* __restrict__ s, double a, size_t n) {
    for (size_t q = blockIdx.x * (size_t)blockDim.x + threadIdx.x; q < n; q += (size_t)gridDim.x * blockDim.x)
        d[q] += a * s[q];
}
__global__ void k_rk4_tmp(double* __restrict__ t, const double* __restrict__ p, const double* __restrict__ kk,
                          double a, size_t n) {
    for (size_t q = blockIdx.x * (size_t)blockDim.x + threadIdx.x; q < n; q += (size_t)gridDim.x * blockDim.x)
        t[q] = p[q] + a * kk[q];
}
__global__ void k_rk4_update(double* __restrict__ p, const double* __restrict__ k1, const double* __restrict__ k2,
                             const double* __restrict__ k3, const double* __restrict__ k4, double dt6, size_t n) {
    for (size_t q = blockIdx.x * (size_t)blockDim.x + threadIdx.x; q < n; q += (size_t)gridDim.x * blockDim.x)
        p[q] += dt6 * (k1[q] + 2.0 * k2[q] + 2.0 * k3[q] + k4[q]);
}

__global__ void k_f64_to_f32(float* __restrict__ d, const double* __restrict__ s, size_t n) {
    for (size_t q = blockIdx.x * (size_t)blockDim.x + threadIdx.x; q < n; q += (size_t)gridDim.x * blockDim.x)
        d[q] = (float)s[q];
}
__global__ void k_f32_to_f64(double* __restrict__ d, const float* __restrict__ s, size_t n) {
    for (size_t q = blockIdx.x * (size_t)blockDim.x + threadIdx.x; q < n; q += (size_t)gridDim.x * blockDim.x)
        d[q] = (double)s[q];
}

static inline int flat_blocks(size_t n) {
    size_t b = (n + 255) / 256;
    return (int)(b > 8192 ? 8192 : (b == 0 ? 1 : b));
}

}  // namespace phx

using namespace phx;

extern "C" {

int phx_abi_version(void) { return PHX_ABI_VERSION; }

const char* phx_last_error(void) { return g_last_error.c_str(); }

size_t phx_stored_size(const phx_grid* G) {
    if (check_grid(G, "phx_stored_size")) return 0;
    return (size_t)make_layout(G).total();
}

int phx_stream_sync(phx_stream s) {
    PHX_HIP(hipStreamSynchronize(as_stream(s)));
    return PHX_OK;
}

int phx_bc_apply(const phx_grid* G, int dtype, void* data, int axis, int kind, int side, double value,
                 phx_stream s) {
    int rc = check_grid(G, "phx_bc_apply");
    if (rc) return rc;
    if ((rc = check_dtype(dtype, "phx_bc_apply"))) return rc;
    if (!data) return fail(PHX_ERR_INVALID, "phx_bc_apply: data is NULL");
    if (axis < 0 || axis > 2) return fail(PHX_ERR_INVALID, "phx_bc_apply: axis must be 0, 1 or 2");
    if (kind < PHX_BC_PERIODIC || kind > PHX_BC_FIXED) return fail(PHX_ERR_INVALID, "phx_bc_apply: bad kind");
    if (side < PHX_SIDE_LOW || side > PHX_SIDE_BOTH) return fail(PHX_ERR_INVALID, "phx_bc_apply: bad side");
    if (G->ghost == 0) return PHX_OK;
    return dtype == PHX_F64 ? bc_apply_t(G, (double*)data, axis, kind, side, value, as_stream(s))
                            : bc_apply_t(G, (float*)data, axis, kind, side, value, as_stream(s));
}

int phx_bc_fill(const phx_grid* G, int dtype, void* data, phx_stream s) {
    int rc = check_grid(G, "phx_bc_fill");
    if (rc) return rc;
    if ((rc = check_dtype(dtype, "phx_bc_fill"))) return rc;
    if (!data) return fail(PHX_ERR_INVALID, "phx_bc_fill: data is NULL");
    if (G->ghost == 0) return PHX_OK;
    return bc_fill_impl(G, dtype, data, as_stream(s));
}

int phx_lap_accumulate(const phx_grid* G, double* rhs, const double* src, double coeff, phx_stream s) {
    int rc = check_grid(G, "phx_lap_accumulate");
    if (rc) return rc;
    if (G->ghost < 1) return fail(PHX_ERR_INVALID, "phx_lap_accumulate: needs ghost >= 1");
    if (!rhs || !src) return fail(PHX_ERR_INVALID, "phx_lap_accumulate: NULL buffer");
    TermGeom t = term_geom(G);
    hipLaunchKernelGGL(k_lap_acc, phys_grid(t.L), dim3(64), 0, as_stream(s), rhs, src, coeff, t);
    PHX_LAUNCH_CHECK("phx_lap_accumulate");
    return PHX_OK;
}

int phx_grad_accumulate(const phx_grid* G, double* rhs, const double* src, int axis, double coeff, phx_stream s) {
    int rc = check_grid(G, "phx_grad_accumulate");
    if (rc) return rc;
    if (axis < 0 || axis >= G->dim)
        return fail(PHX_ERR_INVALID, "grad: axis out of range for this mesh dimension");
    if (G->ghost < 1) return fail(PHX_ERR_INVALID, "phx_grad_accumulate: needs ghost >= 1");
    if (!rhs || !src) return fail(PHX_ERR_INVALID, "phx_grad_accumulate: NULL buffer");
    Layout L = make_layout(G);
    long long st = axis == 0 ? 1 : (axis == 1 ? L.sx : L.sxy);
    hipLaunchKernelGGL(k_grad_acc, phys_grid(L), dim3(64), 0, as_stream(s), rhs, src, coeff, st, 0.5 / G->d[axis], L);
    PHX_LAUNCH_CHECK("phx_grad_accumulate");
    return PHX_OK;
}

int phx_mul_accumulate(const phx_grid* G, double* rhs, const double* s1, const double* s2, double coeff,
                       phx_stream s) {
    int rc = check_grid(G, "phx_mul_accumulate");
    if (rc) return rc;
    if (!rhs || !s1 || !s2) return fail(PHX_ERR_INVALID, "phx_mul_accumulate: NULL buffer");
    Layout L = make_layout(G);
    hipLaunchKernelGGL(k_mul_acc, phys_grid(L), dim3(64), 0, as_stream(s), rhs, s1, s2, coeff, L);
    PHX_LAUNCH_CHECK("phx_mul_accumulate");
    return PHX_OK;
}

int phx_pw_doublewell_accumulate(const phx_grid* G, double* rhs, const double* src, double rho, double ca,
                                 double cb, double coeff, phx_stream s) {
    int rc = check_grid(G, "phx_pw_doublewell_accumulate");
    if (rc) return rc;
    if (!rhs || !src) return fail(PHX_ERR_INVALID, "phx_pw_doublewell_accumulate: NULL buffer");
    Layout L = make_layout(G);
    hipLaunchKernelGGL(k_pw_dw_acc, phys_grid(L), dim3(64), 0, as_stream(s), rhs, src, rho, ca, cb, coeff, L);
    PHX_LAUNCH_CHECK("phx_pw_doublewell_accumulate");
    return PHX_OK;
}

#define PHX_FLAT_ENTRY(name, kernel, check, ...)                                            \
    do {                                                                                    \
        if (n == 0) return PHX_OK;                                                          \
        if (check) return fail(PHX_ERR_INVALID, std::string(name) + ": NULL buffer");       \
        hipLaunchKernelGGL(kernel, dim3(flat_blocks(n)), dim3(256), 0, as_stream(s), __VA_ARGS__); \
        PHX_LAUNCH_CHECK(name);                                                             \
        return PHX_OK;                                                                      \
    } while (0)

int phx_fill_f64(double* dst, double value, size_t n, phx_stream s) {
    PHX_FLAT_ENTRY("phx_fill_f64", k_fill, !dst, dst, value, n);
}
int phx_copy_f64(double* dst, const double* src, size_t n, phx_stream s) {
    PHX_FLAT_ENTRY("phx_copy_f64", k_copy, !dst || !src, dst, src, n);
}
int phx_axpy_f64(double* dst, const double* src, double coeff, size_t n, phx_stream s) {
    PHX_FLAT_ENTRY("phx_axpy_f64", k_axpy, !dst || !src, dst, src, coeff, n);
}
int phx_rk4_tmp_f64(double* tmp, const double* phi, const double* k, double coeff, size_t n, phx_stream s) {
    PHX_FLAT_ENTRY("phx_rk4_tmp_f64", k_rk4_tmp, !tmp || !phi || !k, tmp, phi, k, coeff, n);
}
int phx_rk4_update_f64(double* phi, const double* k1, const double* k2, const double* k3, const double* k4,
                       double dt6, size_t n, phx_stream s) {
    PHX_FLAT_ENTRY("phx_rk4_update_f64", k_rk4_update, !phi || !k1 || !k2 || !k3 || !k4, phi, k1, k2, k3, k4, dt6,
                   n);
}

int phx_convert_f64_to_f32(float* dst, const double* src, size_t n, phx_stream s) {
    PHX_FLAT_ENTRY("phx_convert_f64_to_f32", k_f64_to_f32, !dst || !src, dst, src, n);
}
int phx_convert_f32_to_f64(double* dst, const float* src, size_t n, phx_stream s) {
    PHX_FLAT_ENTRY("phx_convert_f32_to_f64", k_f32_to_f64, !dst || !src, dst, src, n);
}

}  // extern "C"
