// probe.hip -- HBM streaming probes: the bandwidth that simple access
// patterns reach on this GPU, as the practical ceiling the stencil kernels
// are compared with (their roofline is quoted against the 8 TB/s peak).
//
// No reference counterpart. Kernels (FP32, float4 where the pattern allows):
//   kind 0: read a                    (grid-stride, result kept only behind a sentinel test)
//   kind 1: write a                   (grid-stride)
//   kind 2: copy a -> b               (one float4 per thread, no loop)
//   kind 3: d = b + 0.5 a, e = c + 0.25 a  (3 reads, 2 writes: an RK4 stage's streams)
//   kind 4: the 3D stage kernels' access pattern without compute or LDS:
//           per workgroup TX x TY columns of a z-chunk of a side^3 grid stored
//           with one ghost layer; per plane the (TX+4) x (TY+4) halo of a,
//           the cells of b and c read, the cells of d and c written
//   kind 5: as 4 without the halo (the cells of a only)
#include "common.hpp"

namespace phx {

typedef float f4 __attribute__((ext_vector_type(4)));

__global__ __launch_bounds__(256) void k_probe_read(const f4* __restrict__ a, long long n4, float* __restrict__ sink) {
    f4 s = {0.f, 0.f, 0.f, 0.f};
    for (long long i = blockIdx.x * 256LL + threadIdx.x; i < n4; i += (long long)gridDim.x * 256) s += a[i];
    const float t = s.x + s.y + s.z + s.w;
    if (t == 1234.5678f) sink[blockIdx.x] = t;
}

__global__ __launch_bounds__(256) void k_probe_write(f4* __restrict__ a, long long n4) {
    for (long long i = blockIdx.x * 256LL + threadIdx.x; i < n4; i += (long long)gridDim.x * 256)
        a[i] = f4{1.f, 1.f, 1.f, 1.f};
}

__global__ __launch_bounds__(256) void k_probe_copy(const f4* __restrict__ a, f4* __restrict__ b, long long n4) {
    const long long i = blockIdx.x * 256LL + threadIdx.x;
    if (i < n4) b[i] = a[i];
}

__global__ __launch_bounds__(256) void k_probe_stage(const f4* __restrict__ a, const f4* __restrict__ b,
                                                     const f4* __restrict__ c, f4* __restrict__ d,
                                                     f4* __restrict__ e, long long n4) {
    for (long long i = blockIdx.x * 256LL + threadIdx.x; i < n4; i += (long long)gridDim.x * 256) {
        const f4 x = a[i];
        d[i] = b[i] + 0.5f * x;
        e[i] = c[i] + 0.25f * x;
    }
}

template <int TX, int TY, bool HALO>
__global__ __launch_bounds__(256) void k_probe_tiles(const float* __restrict__ u, const float* __restrict__ c,
                                                     float* __restrict__ acc, float* __restrict__ out, int side) {
    constexpr int UX = TX + 4, UY = TY + 4, NU = (UX * UY + 255) / 256, NO = TX * TY / 256;
    const int S = side + 2;
    const long long sxy = (long long)S * S;
    const int tid = threadIdx.x;
    const int x0 = blockIdx.x * TX, y0 = blockIdx.y * TY;
    const int z0 = (int)((long long)blockIdx.z * side / gridDim.z), z1 = (int)((long long)(blockIdx.z + 1) * side / gridDim.z);
    int hoff[NU];
#pragma unroll
    for (int i = 0; i < NU; ++i) {
        const int t = min(tid + 256 * i, UX * UY - 1);
        const int ly = t / UX, lx = t - ly * UX;
        const int px = min(max(x0 - 2 + lx, -1), side), py = min(max(y0 - 2 + ly, -1), side);
        hoff[i] = (py + 1) * S + (px + 1);
    }
    int ooff[NO];
#pragma unroll
    for (int j = 0; j < NO; ++j) {
        const int t = tid + 256 * j;
        const int oy = t / TX, ox = t - oy * TX;
        ooff[j] = (y0 + oy + 1) * S + (x0 + ox + 1);
    }
    float hs = 0.f;
    for (int z = z0; z < z1; ++z) {
        const long long pb = (long long)(z + 1) * sxy;
        float r[NU], cv[NO], av[NO], uv[NO];
        if (HALO) {
#pragma unroll
            for (int i = 0; i < NU; ++i) r[i] = u[pb + hoff[i]];
        }
#pragma unroll
        for (int j = 0; j < NO; ++j) {
            cv[j] = c[pb + ooff[j]];
            av[j] = acc[pb + ooff[j]];
            if (!HALO) uv[j] = u[pb + ooff[j]];
        }
        if (HALO) {
#pragma unroll
            for (int i = 0; i < NU; ++i) hs += r[i];
        }
#pragma unroll
        for (int j = 0; j < NO; ++j) {
            const float uu = HALO ? hs : uv[j];
            out[pb + ooff[j]] = cv[j] + 0.5f * uu;
            acc[pb + ooff[j]] = av[j] + 0.25f * uu;
        }
    }
}

}  // namespace phx

using namespace phx;

extern "C" {

int phx_hbm_probe(int kind, float* a, float* b, float* c, float* d, float* e, long long n, int side, int tile,
                  int blocks, phx_stream s) {
    hipStream_t st = as_stream(s);
    if (kind < 0 || kind > 5) return fail(PHX_ERR_INVALID, "phx_hbm_probe: kind must be in [0, 5]");
    if (kind <= 3) {
        if (n <= 0 || n % 4) return fail(PHX_ERR_INVALID, "phx_hbm_probe: n must be a positive multiple of 4");
        if (!a || (kind >= 2 && !b) || (kind == 3 && (!c || !d || !e)))
            return fail(PHX_ERR_INVALID, "phx_hbm_probe: NULL buffer");
        if (blocks < 0) return fail(PHX_ERR_INVALID, "phx_hbm_probe: blocks < 0");
        const long long n4 = n / 4;
        const int gs = blocks > 0 ? blocks : 8192;
        switch (kind) {
            // kind 0 writes its sink (one float per workgroup) into b when the
            // sentinel sum occurs; b must hold `gs` floats
            case 0:
                if (!b || n < gs) return fail(PHX_ERR_INVALID, "phx_hbm_probe: kind 0 needs b with >= blocks floats");
                hipLaunchKernelGGL(k_probe_read, dim3(gs), dim3(256), 0, st, (const f4*)a, n4, b);
                break;
            case 1: hipLaunchKernelGGL(k_probe_write, dim3(gs), dim3(256), 0, st, (f4*)a, n4); break;
            case 2:
                hipLaunchKernelGGL(k_probe_copy, dim3((unsigned)((n4 + 255) / 256)), dim3(256), 0, st, (const f4*)a,
                                   (f4*)b, n4);
                break;
            default:
                hipLaunchKernelGGL(k_probe_stage, dim3(gs), dim3(256), 0, st, (const f4*)a, (const f4*)b,
                                   (const f4*)c, (f4*)d, (f4*)e, n4);
        }
    } else {
        static const int tw[3][2] = {{64, 16}, {128, 8}, {256, 4}};
        if (tile < 0 || tile > 2) return fail(PHX_ERR_INVALID, "phx_hbm_probe: tile must be 0 (64x16), 1 (128x8), 2 (256x4)");
        const int TX = tw[tile][0], TY = tw[tile][1];
        if (side < 8 || side % TX || side % TY)
            return fail(PHX_ERR_INVALID, "phx_hbm_probe: side must be >= 8 and a multiple of the tile");
        if (!a || !b || !c || !d) return fail(PHX_ERR_INVALID, "phx_hbm_probe: NULL buffer");
        const long long S = side + 2;
        if (n < S * S * S) return fail(PHX_ERR_INVALID, "phx_hbm_probe: buffers must hold (side+2)^3 floats");
        if (blocks < 1 || blocks > side) return fail(PHX_ERR_INVALID, "phx_hbm_probe: z-chunks must be in [1, side]");
        const dim3 grid(side / TX, side / TY, blocks);
        const bool halo = kind == 4;
#define PHX_PROBE_TILE(TXV, TYV)                                                                             \
    if (halo) hipLaunchKernelGGL((k_probe_tiles<TXV, TYV, true>), grid, dim3(256), 0, st, a, b, c, d, side); \
    else hipLaunchKernelGGL((k_probe_tiles<TXV, TYV, false>), grid, dim3(256), 0, st, a, b, c, d, side);
        if (tile == 0) { PHX_PROBE_TILE(64, 16) }
        else if (tile == 1) { PHX_PROBE_TILE(128, 8) }
        else { PHX_PROBE_TILE(256, 4) }
#undef PHX_PROBE_TILE
    }
    PHX_LAUNCH_CHECK("phx_hbm_probe");
    return PHX_OK;
}

}  // extern "C"
