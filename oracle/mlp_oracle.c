/*
 * mlp_oracle.c -- CPU oracle (test infrastructure only) for the search
 * policy's MLP: 14 -> 64 -> 64 -> 2, ReLU, two sigmoid heads, loss = sum over
 * masked outputs of binary cross-entropy with logits, divided by the number of
 * masked outputs.  Plain double-precision forward and backward passes; the
 * reference has no learned component, so this is pinned by central finite
 * differences (tests/test_mlp_oracle.py).
 *
 * Parameter layout: W1[64][14] b1[64] W2[64][64] b2[64] W3[2][64] b3[2].
 */
#include <math.h>
#include <stdint.h>
#include <string.h>

#define IN 14
#define H 64
#define OUT 2
#define O_W1 0
#define O_B1 (O_W1 + H * IN)
#define O_W2 (O_B1 + H)
#define O_B2 (O_W2 + H * H)
#define O_W3 (O_B2 + H)
#define O_B3 (O_W3 + OUT * H)
#define NPARAM (O_B3 + OUT)

int oracle_mlp_nparam(void) { return NPARAM; }

/* logits[n][2]; returns the loss; grads (may be NULL) receive dLoss/dparam. */
double oracle_mlp_loss_grad(const double *P, const float *X, const float *Y, const float *M, int n,
                            double *logits, double *grads) {
    double cnt = 0;
    for (int i = 0; i < n * OUT; i++) cnt += M[i];
    const double inv = 1.0 / (cnt > 1.0 ? cnt : 1.0);
    if (grads) memset(grads, 0, sizeof(double) * NPARAM);
    double loss = 0;
    for (int s = 0; s < n; s++) {
        const float *x = X + s * IN;
        double z1[H], h1[H], z2[H], h2[H], dy[OUT], d2[H], d1[H];
        for (int j = 0; j < H; j++) {
            double v = P[O_B1 + j];
            for (int i = 0; i < IN; i++) v += P[O_W1 + j * IN + i] * (double)x[i];
            z1[j] = v;
            h1[j] = v > 0 ? v : 0;
        }
        for (int j = 0; j < H; j++) {
            double v = P[O_B2 + j];
            for (int i = 0; i < H; i++) v += P[O_W2 + j * H + i] * h1[i];
            z2[j] = v;
            h2[j] = v > 0 ? v : 0;
        }
        for (int o = 0; o < OUT; o++) {
            double v = P[O_B3 + o];
            for (int j = 0; j < H; j++) v += P[O_W3 + o * H + j] * h2[j];
            if (logits) logits[s * OUT + o] = v;
            const double m = M[s * OUT + o], l = Y[s * OUT + o];
            const double p = 1.0 / (1.0 + exp(-v));
            loss += m * (fmax(v, 0.0) - v * l + log1p(exp(-fabs(v)))) * inv;
            dy[o] = m * (p - l) * inv;
        }
        if (!grads) continue;
        for (int j = 0; j < H; j++) {
            double v = 0;
            for (int o = 0; o < OUT; o++) v += P[O_W3 + o * H + j] * dy[o];
            d2[j] = z2[j] > 0 ? v : 0;
        }
        for (int i = 0; i < H; i++) {
            double v = 0;
            for (int j = 0; j < H; j++) v += P[O_W2 + j * H + i] * d2[j];
            d1[i] = z1[i] > 0 ? v : 0;
        }
        for (int o = 0; o < OUT; o++) {
            grads[O_B3 + o] += dy[o];
            for (int j = 0; j < H; j++) grads[O_W3 + o * H + j] += dy[o] * h2[j];
        }
        for (int j = 0; j < H; j++) {
            grads[O_B2 + j] += d2[j];
            for (int i = 0; i < H; i++) grads[O_W2 + j * H + i] += d2[j] * h1[i];
            grads[O_B1 + j] += d1[j];
            for (int i = 0; i < IN; i++) grads[O_W1 + j * IN + i] += d1[j] * (double)x[i];
        }
    }
    return loss;
}
