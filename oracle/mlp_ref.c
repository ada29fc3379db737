/*
 * mlp_ref.c -- ORACLE / TEST INFRASTRUCTURE ONLY.
 *
 * FP64 restatement of the learned-proposal network of libelp.so
 * (include/elp_gpu.h, elp_net_*): forward pass, softmax over the 11 moves,
 * mean cross-entropy and its gradient by backpropagation, all in double from
 * the fp32 parameter vector W1[144][64] W2[64][64] b2[64] W3[64][16] b3[16]:
 *   h1 = relu(x W1),  h2 = relu(h1 W2 + b2),  z = h2 W3 + b3,
 *   q = softmax(z[0..10]),  loss = -mean log q[y].
 * Pinned by central finite differences of the loss (tests/test_learned_oracle.py).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define NIN 144
#define NH 64
#define NOUT 16
#define NMOVE 11
#define OFF_W1 0
#define OFF_W2 (OFF_W1 + NIN * NH)
#define OFF_B2 (OFF_W2 + NH * NH)
#define OFF_W3 (OFF_B2 + NH)
#define OFF_B3 (OFF_W3 + NH * NOUT)
#define NPARAM (OFF_B3 + NOUT)

static void forward1(const double* p, const int8_t* x, double* h1, double* h2, double* q) {
    for (int j = 0; j < NH; j++) {
        double a = 0;
        for (int i = 0; i < NIN; i++) a += (double)x[i] * p[OFF_W1 + i * NH + j];
        h1[j] = a > 0 ? a : 0;
    }
    for (int j = 0; j < NH; j++) {
        double a = p[OFF_B2 + j];
        for (int i = 0; i < NH; i++) a += h1[i] * p[OFF_W2 + i * NH + j];
        h2[j] = a > 0 ? a : 0;
    }
    double z[NMOVE], mx = -1e300, s = 0;
    for (int o = 0; o < NMOVE; o++) {
        double a = p[OFF_B3 + o];
        for (int i = 0; i < NH; i++) a += h2[i] * p[OFF_W3 + i * NOUT + o];
        z[o] = a;
        if (a > mx) mx = a;
    }
    for (int o = 0; o < NMOVE; o++) { q[o] = exp(z[o] - mx); s += q[o]; }
    for (int o = 0; o < NMOVE; o++) q[o] /= s;
}

static double* to_double(const float* prm) {
    double* p = (double*)malloc(sizeof(double) * NPARAM);
    for (int i = 0; i < NPARAM; i++) p[i] = prm[i];
    return p;
}

int orc_mlp_nparam(void) { return NPARAM; }

/* prob[n][11] */
int orc_mlp_forward(const float* prm, const int8_t* x, int n, double* prob) {
    double* p = to_double(prm);
    double h1[NH], h2[NH];
    for (int b = 0; b < n; b++) forward1(p, x + (size_t)b * NIN, h1, h2, prob + (size_t)b * NMOVE);
    free(p);
    return 0;
}

/* loss and gradient of the mean cross-entropy; parameters taken in double
   from prm_d (so finite differences can perturb them exactly) */
int orc_mlp_grad_d(const double* p, const int8_t* x, const int32_t* y, int n, double* grad, double* loss) {
    memset(grad, 0, sizeof(double) * NPARAM);
    double h1[NH], h2[NH], q[NMOVE], d3[NMOVE], d2[NH], d1[NH];
    double L = 0;
    for (int b = 0; b < n; b++) {
        const int8_t* xb = x + (size_t)b * NIN;
        forward1(p, xb, h1, h2, q);
        L -= log(q[y[b]]);
        for (int o = 0; o < NMOVE; o++) d3[o] = (q[o] - (o == y[b] ? 1.0 : 0.0)) / n;
        for (int o = 0; o < NMOVE; o++) {
            grad[OFF_B3 + o] += d3[o];
            for (int i = 0; i < NH; i++) grad[OFF_W3 + i * NOUT + o] += h2[i] * d3[o];
        }
        for (int i = 0; i < NH; i++) {
            double a = 0;
            for (int o = 0; o < NMOVE; o++) a += d3[o] * p[OFF_W3 + i * NOUT + o];
            d2[i] = h2[i] > 0 ? a : 0;
        }
        for (int j = 0; j < NH; j++) {
            grad[OFF_B2 + j] += d2[j];
            for (int i = 0; i < NH; i++) grad[OFF_W2 + i * NH + j] += h1[i] * d2[j];
        }
        for (int i = 0; i < NH; i++) {
            double a = 0;
            for (int j = 0; j < NH; j++) a += d2[j] * p[OFF_W2 + i * NH + j];
            d1[i] = h1[i] > 0 ? a : 0;
        }
        for (int i = 0; i < NIN; i++) {
            if (!xb[i]) continue;
            for (int j = 0; j < NH; j++) grad[OFF_W1 + i * NH + j] += (double)xb[i] * d1[j];
        }
    }
    if (loss) *loss = L / n;
    return 0;
}

int orc_mlp_grad(const float* prm, const int8_t* x, const int32_t* y, int n, double* grad, double* loss) {
    double* p = to_double(prm);
    orc_mlp_grad_d(p, x, y, n, grad, loss);
    free(p);
    return 0;
}
