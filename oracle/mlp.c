/* mlp.c -- CPU restatement of the tile-cost MLP (TEST INFRASTRUCTURE ONLY:
 * the checker for the library's HIP forward / backward kernels, never a
 * product path).
 *
 * The model has no counterpart in the reference (whose tuner only searches,
 * autotune/StandardEM.cpp:69-121); this file is the mathematical definition
 * the GPU kernels must agree with: x -> relu(x W1 + b1) -> relu(. W2 + b2) ->
 * relu(. W3 + b3) -> . w4 + b4, loss = mean_b (y_b - t_b)^2, evaluated in
 * double from the f32 parameters.  Pinned by finite differences
 * (tests/test_oracle_mlp.py). */
#include <stddef.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "chai_oracle.h"

#define H 64

typedef struct { size_t W1, b1, W2, b2, W3, b3, w4, b4, P; } Lay;

static Lay lay(int F)
{
    Lay L;
    L.W1 = 0; L.b1 = (size_t)F * H; L.W2 = L.b1 + H; L.b2 = L.W2 + H * H; L.W3 = L.b2 + H;
    L.b3 = L.W3 + H * H; L.w4 = L.b3 + H; L.b4 = L.w4 + H; L.P = L.b4 + 1;
    return L;
}

size_t chai_or_mlp_num_params(int F) { return lay(F).P; }

/* one sample: activations z1..z3 (pre-ReLU), output y */
static double fwd(const float* p, const Lay* L, int F, const float* x, double* z1, double* z2, double* z3)
{
    for (int j = 0; j < H; j++) {
        double s = p[L->b1 + j];
        for (int f = 0; f < F; f++) s += (double)x[f] * p[L->W1 + (size_t)f * H + j];
        z1[j] = s;
    }
    for (int j = 0; j < H; j++) {
        double s = p[L->b2 + j];
        for (int k = 0; k < H; k++) s += (z1[k] > 0 ? z1[k] : 0) * p[L->W2 + (size_t)k * H + j];
        z2[j] = s;
    }
    for (int j = 0; j < H; j++) {
        double s = p[L->b3 + j];
        for (int k = 0; k < H; k++) s += (z2[k] > 0 ? z2[k] : 0) * p[L->W3 + (size_t)k * H + j];
        z3[j] = s;
    }
    double y = p[L->b4];
    for (int j = 0; j < H; j++) y += (z3[j] > 0 ? z3[j] : 0) * p[L->w4 + j];
    return y;
}

int chai_or_mlp_forward(const float* params, int F, const float* X, size_t B, double* out)
{
    Lay L = lay(F);
    double z1[H], z2[H], z3[H];
    for (size_t b = 0; b < B; b++) out[b] = fwd(params, &L, F, X + b * F, z1, z2, z3);
    return 0;
}

/* grad = d/dparams of mean_b (y_b - Y_b)^2; returns the loss */
double chai_or_mlp_gradient(const float* params, int F, const float* X, const float* Y, size_t B, double* grad)
{
    Lay L = lay(F);
    memset(grad, 0, sizeof(double) * L.P);
    double z1[H], z2[H], z3[H], d3[H], d2[H], d1[H], loss = 0;
    for (size_t b = 0; b < B; b++) {
        const float* x = X + b * F;
        const double y = fwd(params, &L, F, x, z1, z2, z3);
        const double dy = 2.0 * (y - Y[b]) / (double)B;
        loss += (y - Y[b]) * (y - Y[b]) / (double)B;
        grad[L.b4] += dy;
        for (int j = 0; j < H; j++) {
            grad[L.w4 + j] += (z3[j] > 0 ? z3[j] : 0) * dy;
            d3[j] = z3[j] > 0 ? params[L.w4 + j] * dy : 0;
        }
        for (int k = 0; k < H; k++) {
            double s = 0;
            for (int j = 0; j < H; j++) {
                grad[L.W3 + (size_t)k * H + j] += (z2[k] > 0 ? z2[k] : 0) * d3[j];
                s += params[L.W3 + (size_t)k * H + j] * d3[j];
            }
            d2[k] = z2[k] > 0 ? s : 0;
        }
        for (int j = 0; j < H; j++) grad[L.b3 + j] += d3[j];
        for (int k = 0; k < H; k++) {
            double s = 0;
            for (int j = 0; j < H; j++) {
                grad[L.W2 + (size_t)k * H + j] += (z1[k] > 0 ? z1[k] : 0) * d2[j];
                s += params[L.W2 + (size_t)k * H + j] * d2[j];
            }
            d1[k] = z1[k] > 0 ? s : 0;
        }
        for (int j = 0; j < H; j++) grad[L.b2 + j] += d2[j];
        for (int f = 0; f < F; f++)
            for (int j = 0; j < H; j++) grad[L.W1 + (size_t)f * H + j] += (double)x[f] * d1[j];
        for (int j = 0; j < H; j++) grad[L.b1 + j] += d1[j];
    }
    return loss;
}
