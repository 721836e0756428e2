/*
 * moments.c -- CPU restatement of the sliding-window coverage moments
 * (TEST INFRASTRUCTURE ONLY; see oracle.h).
 *
 *  - orc_compact_index / orc_compact_ls: src/fourier/moments.rs:58-66 and the
 *    stored-L sequence of :225-247 (L = 1..=min(100, n-1), then l_f *= 1.01,
 *    l = ceil(l_f), kept when l > prev, stop when l > n-1), evaluated in IEEE
 *    FP64 with the C library's log / ceil.
 *  - orc_depth_signal_f32: src/fourier/depth.rs:109-136 (fractional impulse
 *    train, spikes clipped at n_bins*100 by build_spikes :206-215) followed by
 *    the 8-wide chunked f32 prefix scan of simd_cumsum_f32 (:143-204), in the
 *    same operation order, so g matches the reference's f32 values.
 *  - orc_depth_exact: the same impulse train in integer base pairs (x100), so
 *    G[j] is the exact bp coverage of bin j (g = G/100 up to f32 rounding).
 *  - orc_moments_reference: moments_from_signal (:192-254) -- d*, d**, d2 with
 *    the 4-wide chunked FP64 scans of :262-386, R = autocorrelation of d* via
 *    a zero-padded FP64 FFT of length next_pow2(2(n+1)) (:421-466), and
 *    window_moments (:392-411).  Only R's rounding differs from the reference
 *    (this file's radix-2 FFT instead of rustfft 6.4.1 / realfft 3.5.0).
 *  - orc_moments_exact: S1(L) = sum_w sum_{j in w} G[j] and
 *    S2(L) = sum_w (sum_{j in w} G[j])^2 over the n-L+1 windows, in exact
 *    64/128-bit integers; mean = S1/(100 n_w), var = (S2 n_w - S1^2)/(1e4 n_w^2)
 *    with an exact int128 numerator.  This is the FP64 evaluation of the
 *    reference method that the GPU moments are compared against.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define T_THRESHOLD 100u
#define EPS 0.01

size_t orc_compact_index(size_t l) {
    if (l <= T_THRESHOLD) return l - 1;
    size_t k = (size_t)(log((double)l / (double)T_THRESHOLD) / log(1.0 + EPS));
    return T_THRESHOLD + k - 1;
}

size_t orc_compact_ls(uint32_t n_bins, uint32_t* ls) {
    if (n_bins <= 1) return 0;
    size_t max_l = n_bins - 1, k = 0;
    for (size_t l = 1; l <= (T_THRESHOLD < max_l ? T_THRESHOLD : max_l); l++) {
        if (ls) ls[k] = (uint32_t)l;
        k++;
    }
    double l_f = (double)T_THRESHOLD;
    size_t prev = T_THRESHOLD;
    for (;;) {
        l_f *= 1.0 + EPS;
        size_t l = (size_t)ceil(l_f);
        if (l > max_l) break;
        if (l > prev) {
            if (ls) ls[k] = (uint32_t)l;
            k++;
            prev = l;
        }
    }
    return k;
}

/* ------------------------------------------------------------ depth */

static void simd_cumsum_f32(float* g, size_t n) {
    if (n <= 1) return;
    const size_t W = 8;
    size_t full = n / W, tail = full * W;
    for (size_t c = 0; c < full; c++)
        for (size_t i = 1; i < W; i++) g[c * W + i] += g[c * W + i - 1];
    for (size_t i = tail + 1; i < n; i++) g[i] += g[i - 1];
    if (full == 0) return;
    float carry = 0.0f;
    float* carries = (float*)malloc(full * sizeof(float));
    for (size_t c = 0; c < full; c++) {
        carries[c] = carry;
        carry += g[c * W + W - 1];
    }
    for (size_t c = 1; c < full; c++)
        for (size_t i = 0; i < W; i++) g[c * W + i] += carries[c];
    if (tail < n && carry != 0.0f)
        for (size_t i = tail; i < n; i++) g[i] += carry;
    free(carries);
}

void orc_depth_signal_f32(size_t n, const uint32_t* start, const uint32_t* end,
                          uint32_t n_bins, float* g) {
    uint32_t max_bp = n_bins * ORC_BIN_SIZE;
    memset(g, 0, (size_t)n_bins * sizeof(float));
    for (int pass = 0; pass < 2; pass++) { /* all pos spikes, then all neg */
        for (size_t i = 0; i < n; i++) {
            uint32_t x = pass == 0 ? start[i] : end[i];
            if (x > max_bp) x = max_bp;
            size_t bin = x / ORC_BIN_SIZE;
            float frac = (float)(x % ORC_BIN_SIZE) / (float)ORC_BIN_SIZE;
            if (pass == 0) {
                if (bin < n_bins) g[bin] += 1.0f - frac;
                if (frac > 0.0f && bin + 1 < n_bins) g[bin + 1] += frac;
            } else {
                if (bin < n_bins) g[bin] -= 1.0f - frac;
                if (frac > 0.0f && bin + 1 < n_bins) g[bin + 1] -= frac;
            }
        }
    }
    simd_cumsum_f32(g, n_bins);
}

void orc_depth_exact(size_t n, const uint32_t* start, const uint32_t* end, uint32_t n_bins,
                     int64_t* G) {
    uint32_t max_bp = n_bins * ORC_BIN_SIZE;
    memset(G, 0, (size_t)n_bins * sizeof(int64_t));
    for (size_t i = 0; i < n; i++) {
        for (int pass = 0; pass < 2; pass++) {
            uint32_t x = pass == 0 ? start[i] : end[i];
            int64_t sgn = pass == 0 ? 1 : -1;
            if (x > max_bp) x = max_bp;
            size_t bin = x / ORC_BIN_SIZE;
            int64_t r = x % ORC_BIN_SIZE;
            if (bin < n_bins) G[bin] += sgn * (int64_t)(ORC_BIN_SIZE - r);
            if (r > 0 && bin + 1 < n_bins) G[bin + 1] += sgn * r;
        }
    }
    for (size_t j = 1; j < n_bins; j++) G[j] += G[j - 1];
}

/* ------------------------------------------------------------ FFT */

static void fft_inplace(double* re, double* im, size_t N, int inverse) {
    for (size_t i = 1, j = 0; i < N; i++) {
        size_t bit = N >> 1;
        for (; j & bit; bit >>= 1) j ^= bit;
        j ^= bit;
        if (i < j) {
            double t = re[i]; re[i] = re[j]; re[j] = t;
            t = im[i]; im[i] = im[j]; im[j] = t;
        }
    }
    for (size_t len = 2; len <= N; len <<= 1) {
        size_t half = len / 2;
        double ang = (inverse ? 2.0 : -2.0) * M_PI / (double)len;
        for (size_t k = 0; k < half; k++) {
            double wr = cos(ang * (double)k), wi = sin(ang * (double)k);
            for (size_t i = k; i < N; i += len) {
                size_t j = i + half;
                double xr = re[j] * wr - im[j] * wi, xi = re[j] * wi + im[j] * wr;
                re[j] = re[i] - xr; im[j] = im[i] - xi;
                re[i] += xr; im[i] += xi;
            }
        }
    }
}

/* Linear autocorrelation r[L] = sum_k x[k] x[k+L], L = 0..n-1, via a
 * zero-padded FFT of length next_pow2(2n) (moments.rs:421-466). */
void orc_autocorr_fft(const double* x, size_t n, double* r) {
    if (n == 0) return;
    size_t N = 1;
    while (N < 2 * n) N <<= 1;
    double* re = (double*)calloc(N, sizeof(double));
    double* im = (double*)calloc(N, sizeof(double));
    memcpy(re, x, n * sizeof(double));
    fft_inplace(re, im, N, 0);
    for (size_t k = 0; k < N; k++) {
        re[k] = re[k] * re[k] + im[k] * im[k];
        im[k] = 0.0;
    }
    fft_inplace(re, im, N, 1);
    for (size_t k = 0; k < n; k++) r[k] = re[k] / (double)N;
    free(re);
    free(im);
}

/* ------------------------------------------------- reference moments */

static double* chunk4_prefix(const double* src, size_t n, int square) {
    double* d = (double*)calloc(n + 1, sizeof(double));
    size_t chunks = n / 4;
    for (size_t c = 0; c < chunks; c++) {
        size_t b = c * 4;
        double v0 = src[b], v1 = src[b + 1], v2 = src[b + 2], v3 = src[b + 3];
        if (square) { v0 *= v0; v1 *= v1; v2 *= v2; v3 *= v3; }
        d[b + 1] = v0;
        d[b + 2] = v0 + v1;
        d[b + 3] = v0 + v1 + v2;
        d[b + 4] = v0 + v1 + v2 + v3;
    }
    double carry = 0.0;
    double* carries = (double*)malloc((chunks ? chunks : 1) * sizeof(double));
    for (size_t c = 0; c < chunks; c++) {
        carries[c] = carry;
        carry += d[c * 4 + 4];
    }
    for (size_t c = 1; c < chunks; c++)
        for (size_t k = 1; k <= 4; k++) d[c * 4 + k] += carries[c];
    for (size_t i = chunks * 4; i < n; i++) d[i + 1] = d[i] + (square ? src[i] * src[i] : src[i]);
    free(carries);
    return d;
}

void orc_moments_reference(const float* g, uint32_t n, const uint32_t* ls, size_t nl,
                           double* mean, double* var) {
    double* gd = (double*)malloc((size_t)(n ? n : 1) * sizeof(double));
    for (uint32_t i = 0; i < n; i++) gd[i] = (double)g[i];
    double* d_star = chunk4_prefix(gd, n, 0);          /* n+1 */
    double* d_ss = chunk4_prefix(d_star, n + 1, 0);    /* n+2 */
    double* d2 = chunk4_prefix(d_star, n + 1, 1);      /* n+2 */
    double* r = (double*)malloc((size_t)(n + 1) * sizeof(double));
    orc_autocorr_fft(d_star, n + 1, r);
    for (size_t k = 0; k < nl; k++) {
        size_t l = ls[k];
        double n_w = (double)(n - l + 1);
        double sum_f = d_ss[n + 1] - d_ss[l] - d_ss[n - l + 1];
        double r_l = l < (size_t)n + 1 ? r[l] : 0.0;
        double sum_f2 = (d2[n + 1] - d2[l]) + d2[n - l + 1] - 2.0 * r_l;
        double m = sum_f / n_w, e2 = sum_f2 / n_w;
        double v = e2 - m * m;
        mean[k] = m;
        var[k] = v > 0.0 ? v : 0.0;
    }
    free(gd); free(d_star); free(d_ss); free(d2); free(r);
}

/* ---------------------------------------------------- exact moments */

void orc_moments_exact(const int64_t* G, uint32_t n, const uint32_t* ls, size_t nl,
                       double* mean, double* var, int64_t* s1_out, double* s2_out) {
    int64_t* P = (int64_t*)malloc(((size_t)n + 1) * sizeof(int64_t));
    P[0] = 0;
    for (uint32_t j = 0; j < n; j++) P[j + 1] = P[j] + G[j];
    for (size_t k = 0; k < nl; k++) {
        size_t l = ls[k];
        size_t nw = (size_t)n - l + 1;
        __int128 s1 = 0;
        unsigned __int128 s2 = 0;
        for (size_t i = 0; i < nw; i++) {
            int64_t w = P[i + l] - P[i];
            s1 += w;
            s2 += (unsigned __int128)((__int128)w * w);
        }
        unsigned __int128 num = s2 * (unsigned __int128)nw - (unsigned __int128)(s1 * s1);
        mean[k] = (double)s1 / (100.0 * (double)nw);
        var[k] = (double)num / (1e4 * (double)nw * (double)nw);
        if (s1_out) s1_out[k] = (int64_t)s1;
        if (s2_out) s2_out[k] = (double)s2;
    }
    free(P);
}

/* Exact moments through the autocorrelation identity (valid for any integer
 * signal, here H = G - c with c = round(mean G)):
 *   S2_H(L) = L*A[0] + 2*sum_{m=1}^{L-1} (L-m) A[m]
 *             - sum_{k=1}^{L-1} PH[k]^2 - sum_{k=1}^{L-1} (PH[n] - PH[n-k])^2,
 * A[m] = sum_k H[k] H[k+m], PH = prefix sums of H.  A comes from this file's
 * FP64 FFT rounded to integers; the rounding is accepted only when every value
 * is within 0.25 of an integer and A[0] equals sum H^2 computed exactly, else
 * the function returns 0 and the caller falls back to orc_moments_exact.
 * Everything after A is exact 128-bit integer arithmetic. */
int orc_moments_exact_fft(const int64_t* G, uint32_t n, const uint32_t* ls, size_t nl,
                          double* mean, double* var) {
    if (n == 0) return 1;
    __int128 tot = 0;
    for (uint32_t j = 0; j < n; j++) tot += G[j];
    double md = (double)tot / (double)n;
    int64_t c = (int64_t)llround(md);
    double* h = (double*)malloc((size_t)n * sizeof(double));
    __int128 a0 = 0;
    for (uint32_t j = 0; j < n; j++) {
        int64_t v = G[j] - c;
        h[j] = (double)v;
        a0 += (__int128)v * v;
    }
    double* r = (double*)malloc((size_t)n * sizeof(double));
    orc_autocorr_fft(h, n, r);
    int64_t* A = (int64_t*)malloc((size_t)n * sizeof(int64_t));
    int ok = 1;
    for (uint32_t m = 0; m < n; m++) {
        double rr = nearbyint(r[m]);
        if (fabs(r[m] - rr) > 0.25) ok = 0;
        A[m] = (int64_t)rr;
    }
    if ((__int128)A[0] != a0) ok = 0;
    if (!ok) {
        free(h); free(r); free(A);
        return 0;
    }
    int64_t* P = (int64_t*)malloc(((size_t)n + 1) * sizeof(int64_t));
    P[0] = 0;
    for (uint32_t j = 0; j < n; j++) P[j + 1] = P[j] + G[j];
    __int128* PP = (__int128*)malloc(((size_t)n + 2) * sizeof(__int128));
    PP[0] = 0;
    for (uint32_t j = 0; j <= n; j++) PP[j + 1] = PP[j] + P[j];
    /* running prefix sums over m / k, advanced to each L in ascending order */
    __int128 C1 = 0, C2 = 0, E1 = 0, E2 = 0;
    size_t cur = 1; /* sums cover indices 1..cur-1 */
    __int128 PHn = (__int128)P[n] - (__int128)c * n;
    for (size_t q = 0; q < nl; q++) {
        size_t L = ls[q];
        while (cur < L) {
            C1 += A[cur];
            C2 += (__int128)cur * A[cur];
            __int128 phk = (__int128)P[cur] - (__int128)c * cur;
            __int128 phm = PHn - ((__int128)P[n - cur] - (__int128)c * (n - cur));
            E1 += phk * phk;
            E2 += phm * phm;
            cur++;
        }
        size_t nw = (size_t)n - L + 1;
        __int128 T = (__int128)L * A[0] + 2 * ((__int128)L * C1 - C2);
        __int128 S2 = T - E1 - E2;
        __int128 S1G = PP[n + 1] - PP[L] - PP[n - L + 1];
        __int128 S1H = S1G - (__int128)c * L * nw;
        __int128 num = S2 * (__int128)nw - S1H * S1H;
        mean[q] = (double)S1G / (100.0 * (double)nw);
        var[q] = (double)num / (1e4 * (double)nw * (double)nw);
    }
    free(h); free(r); free(A); free(P); free(PP);
    return 1;
}
