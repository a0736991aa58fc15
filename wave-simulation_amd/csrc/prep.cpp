#include "prep.hpp"

#include <algorithm>
#include <cmath>

namespace wv {

namespace {

// Solve for the half-offset Taylor weights w_k, k = 1..M, of
//   f'(x) ~ sum_k w_k (f(x + (k-1/2)h) - f(x - (k-1/2)h)) / h
// exactly in rational arithmetic would be overkill; the closed form
//   w_k = (-1)^(k+1) [(2M-1)!!]^2 / ((2k-1)^2 (M+k-1)! (M-k)! 2^(2M-2))
// is evaluated with integers that stay exact in double for M <= 6.
double half_weight(int M, int k)
{
    double dfact = 1;  // (2M-1)!!
    for (int i = 2 * M - 1; i > 1; i -= 2) dfact *= i;
    double f1 = 1, f2 = 1;
    for (int i = 2; i <= M + k - 1; i++) f1 *= i;
    for (int i = 2; i <= M - k; i++) f2 *= i;
    double p2 = std::ldexp(1.0, 2 * M - 2);
    double num = dfact * dfact;
    double den = (2.0 * k - 1) * (2.0 * k - 1) * f1 * f2 * p2;
    double w = num / den;
    return (k % 2 == 1) ? w : -w;
}

}  // namespace

double fd_coefficient(int order, int j)
{
    const int M = order / 2;
    // reference order: j = M-1 .. 0 -> -w_{M-j}, j = M .. 2M-1 -> w_{j-M+1}
    if (j >= M) return half_weight(M, j - M + 1);
    return -half_weight(M, M - j);
}

static float coef_f(int order, int j) { return (float)fd_coefficient(order, j); }

void build_taps(std::vector<float> &out, int n, int padded_n, int order, int H, bool forward, bool stencil, float DH,
                float DT)
{
    const int K = 2 * H + 1;
    out.assign((size_t)padded_n * K, 0.f);
    for (int c = 0; c < n; c++) {
        float *row = &out[(size_t)c * K];
        if (stencil) {
            const float s = DT / DH;
            for (int j = 0; j < order; j++) {
                if (forward) {
                    const int off = j - order / 2 + 1;
                    if (c + off >= 0 && c + off < n) row[off + H] = coef_f(order, j) * s;
                } else {
                    // row c of -(forward)^T
                    const int off = -(j - order / 2 + 1);
                    if (c + off >= 0 && c + off < n) row[off + H] = -coef_f(order, j) * s;
                }
            }
            continue;
        }
        // sparse operator: the local order shrinks to the largest even order
        // whose window fits the axis; where even order 2 does not fit, the
        // two-point stencil of the inward neighbour is used.
        int N = order, center = c;
        if (forward) {
            if (c - N / 2 + 1 < 0) N = 2 * c + 2;
            if (center + N / 2 >= n) {
                N = 2 * (n - 1 - center);
                if (N == 0) {
                    N = 2;
                    center = c - 1;
                }
            }
        } else {
            if (c - N / 2 < 0) {
                N = 2 * c;
                if (N == 0) {
                    N = 2;
                    center = c + 1;
                }
            }
            if (center + N / 2 - 1 >= n) N = 2 * (n - center);
        }
        for (int j = 0; j < N; j++) {
            const int off = (center - c) + (forward ? (j - N / 2 + 1) : (j - N / 2));
            const float v = coef_f(N, j) / DH;
            row[off + H] = v * DT;
        }
    }
}

void build_taps_free_surface(std::vector<float> &out, int n, int padded_n, int order, int H, bool forward, float DH,
                             float DT)
{
    const int K = 2 * H + 1;
    out.assign((size_t)padded_n * K, 0.f);
    const int N = order;
    for (int y = 0; y < n; y++) {
        float *row = &out[(size_t)y * K];
        for (int j = 0; j < N; j++) {
            const int off = forward ? (j - N / 2 + 1) : (j - N / 2);
            // image point of tap j mirrored at the surface (y = 0 for stresses
            // on full rows, y = -1/2 for the staggered ones)
            const int img = forward ? (N - 2 - 2 * y - j) : (N - 1 - 2 * y - j);
            const float fd = coef_f(N, j);
            const float diff = img >= 0 ? coef_f(N, img) : 0.f;
            if (y + off >= 0 && y + off < n) {
                const float v = (fd - diff) / DH;
                row[off + H] = v * DT;
            }
        }
    }
}

void cpml_profile(int W, double npower, double fc, double vmax, double DT, double DH, bool half,
                  std::vector<double> &a, std::vector<double> &b)
{
    a.assign(W, 0.0);
    b.assign(W, 0.0);
    const double R = 0.0008;
    const double alpha_max = M_PI * fc;
    const double d0 = -(npower + 1.0) * vmax * std::log(R) / (2.0 * W * DH);
    for (int i = 0; i < W; i++) {
        const double pos = (W - i - (half ? 0.5 : 0.0)) / W;
        const double d = d0 * std::pow(pos, npower);
        const double alpha = alpha_max * (1.0 - pos);
        b[i] = std::exp(-(d + alpha) * DT);
        a[i] = std::fabs(d) > 1.0e-6 ? d * (b[i] - 1.0) / (d + alpha) : 0.0;
    }
}

void cpml_axis(int n, int padded_n, int W, bool skip_near_side, double npower, double fc, double vmax, double DT,
               double DH, std::vector<float> &a, std::vector<float> &b, std::vector<float> &ah, std::vector<float> &bh,
               std::vector<int> &idx)
{
    std::vector<double> pa, pb, pah, pbh;
    cpml_profile(W, npower, fc, vmax, DT, DH, false, pa, pb);
    cpml_profile(W, npower, fc, vmax, DT, DH, true, pah, pbh);
    a.assign(padded_n, 0.f);
    b.assign(padded_n, 0.f);
    ah.assign(padded_n, 0.f);
    bh.assign(padded_n, 0.f);
    idx.assign(padded_n, -1);
    for (int p = 0; p < n; p++) {
        const int d = std::min(p, n - 1 - p);
        if (d >= W) continue;
        const bool near = p < W;
        if (near && skip_near_side) continue;
        // compact index: near side columns 0..W-1, far side W..2W-1
        idx[p] = near ? p : W + (W - 1 - d);
        if (near) {
            a[p] = (float)pa[d];
            b[p] = (float)pb[d];
            ah[p] = (float)pah[d];
            bh[p] = (float)pbh[d];
        } else {
            a[p] = (float)pah[d];
            b[p] = (float)pbh[d];
            ah[p] = (float)pa[d];
            bh[p] = (float)pb[d];
        }
    }
}

std::vector<double> abs_profile(int W, double damping_coeff)
{
    std::vector<double> c(W);
    const double amp = 1.0 - damping_coeff / 100.0;
    const double a = std::sqrt(-std::log(amp) / ((double)W * W));
    for (int j = 0; j < W; j++) c[j] = std::exp(-(a * a * (double)(W - j) * (W - j)));
    return c;
}

void em_ca_cb(float eps, float sigma, float DT, float &ca, float &cb)
{
    float q = 0.5f / eps;
    q *= sigma;
    q *= DT;
    const float cinv = 1.f / (q + 1.f);
    float a = 1.f - q;
    a *= cinv;
    float b = 1.f / eps;
    b *= cinv;
    ca = a;
    cb = b;
}

void em_visco_coef(float eps_s, float sig_s, float tE, float tS, float DT, float inv_tau_mean, int L,
                   const float *cdl, float &ca, float &cb, float *cd)
{
    const float eps0 = 8.8541878176e-12f;
    float e = 1.f - tE;
    e *= eps_s;
    const float t = sig_s * tS;
    e += t;
    if (e < eps0) e = eps0;  // Common::searchAndReplace(..., eps0, eps0, 1)
    float sg = tE * inv_tau_mean;
    sg *= eps_s;
    sg += sig_s;
    em_ca_cb(e, sg, DT, ca, cb);
    for (int l = 0; l < L; l++) {
        float v = tE * cdl[l];
        v *= eps_s;
        v *= -DT;
        cd[l] = v;
    }
}

std::vector<float> transpose_taps(const std::vector<float> &t, int H)
{
    const int K = 2 * H + 1;
    const int np = (int)(t.size() / K);
    std::vector<float> r(t.size(), 0.f);
    for (int j = 0; j < np; j++)
        for (int k = 0; k < K; k++) {
            const int i = j + k - H;
            if (i >= 0 && i < np) r[(size_t)j * K + k] = t[(size_t)i * K + (2 * H - k)];
        }
    return r;
}

}  // namespace wv
