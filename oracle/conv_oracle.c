/*
 * conv_oracle.c -- FP64 CPU oracle of the surrogate's 3x3 circular-padding
 * convolution (forward, data gradient, weight/bias gradient). TEST
 * INFRASTRUCTURE ONLY. The reference has no learned part; this is the naive
 * direct sum from the definition
 *     y[n,o,h,w] = b[o] + sum_{i,dh,dw} w[o,i,dh,dw] * x[n,i,(h+dh-1) mod H,(w+dw-1) mod W]
 * pinned in tests/test_oracle_conv.py by a numpy FFT (circular convolution
 * theorem), identity and shift kernels, the adjoint identity (data gradient)
 * and linearity in (w, b) (weight and bias gradients).
 */
#include <stddef.h>

static inline int wrap(int a, int n) { a %= n; return a < 0 ? a + n : a; }

void orc_conv3x3_fwd(int N, int Ci, int Co, int H, int W, const double* x, const double* w,
                     const double* b, double* y) {
    for (int n = 0; n < N; ++n)
        for (int o = 0; o < Co; ++o)
            for (int h = 0; h < H; ++h)
                for (int q = 0; q < W; ++q) {
                    double s = b ? b[o] : 0.0;
                    for (int i = 0; i < Ci; ++i)
                        for (int dh = 0; dh < 3; ++dh)
                            for (int dw = 0; dw < 3; ++dw)
                                s += w[((o * Ci + i) * 3 + dh) * 3 + dw] *
                                     x[((size_t)(n * Ci + i) * H + wrap(h + dh - 1, H)) * W + wrap(q + dw - 1, W)];
                    y[((size_t)(n * Co + o) * H + h) * W + q] = s;
                }
}

/* dx[n,i,h,w] = sum_{o,dh,dw} w[o,i,dh,dw] * dy[n,o,(h-dh+1) mod H,(w-dw+1) mod W] */
void orc_conv3x3_dgrad(int N, int Ci, int Co, int H, int W, const double* dy, const double* w, double* dx) {
    for (int n = 0; n < N; ++n)
        for (int i = 0; i < Ci; ++i)
            for (int h = 0; h < H; ++h)
                for (int q = 0; q < W; ++q) {
                    double s = 0.0;
                    for (int o = 0; o < Co; ++o)
                        for (int dh = 0; dh < 3; ++dh)
                            for (int dw = 0; dw < 3; ++dw)
                                s += w[((o * Ci + i) * 3 + dh) * 3 + dw] *
                                     dy[((size_t)(n * Co + o) * H + wrap(h - dh + 1, H)) * W + wrap(q - dw + 1, W)];
                    dx[((size_t)(n * Ci + i) * H + h) * W + q] = s;
                }
}

/* dw[o,i,dh,dw] = sum_{n,h,w} dy[n,o,h,w] * x[n,i,h+dh-1,w+dw-1];  db[o] = sum dy[n,o,:,:] */
void orc_conv3x3_wgrad(int N, int Ci, int Co, int H, int W, const double* x, const double* dy,
                       double* dw, double* db) {
    for (int o = 0; o < Co; ++o) {
        if (db) {
            double s = 0.0;
            for (int n = 0; n < N; ++n)
                for (size_t p = 0; p < (size_t)H * W; ++p) s += dy[(size_t)(n * Co + o) * H * W + p];
            db[o] = s;
        }
        for (int i = 0; i < Ci; ++i)
            for (int kh = 0; kh < 3; ++kh)
                for (int kw = 0; kw < 3; ++kw) {
                    double s = 0.0;
                    for (int n = 0; n < N; ++n)
                        for (int h = 0; h < H; ++h)
                            for (int q = 0; q < W; ++q)
                                s += dy[((size_t)(n * Co + o) * H + h) * W + q] *
                                     x[((size_t)(n * Ci + i) * H + wrap(h + kh - 1, H)) * W + wrap(q + kw - 1, W)];
                    dw[((o * Ci + i) * 3 + kh) * 3 + kw] = s;
                }
    }
}
