// moments.hip -- sliding-window coverage moments of the DB files, computed on
// the GPU (replaces build_chrom_moments, src/fourier/moments.rs:129-254, and
// the depth signal of src/fourier/depth.rs:109-136, 206-215).
//
// Per (DB file, shard) with n = n_bins bins:
//   G[j] = bp coverage of bin j (integer; the reference's g = G/100), built
//          from the fractional impulse train of every interval, spikes
//          clipped at n*100;
//   H = G - c, c = round(mean G);  A[m] = sum_k H[k] H[k+m]  (FFT);
//   S1(L) = PP[n+1] - PP[L] - PP[n-L+1]                (P, PP prefix sums of G)
//   S2_H(L) = L A[0] + 2 (L C1[L] - C2[L]) - E1[L] - (Q[n] - Q[n-L+1])
//     C1[L] = sum_{1<=m<L} A[m], C2[L] = sum_{1<=m<L} m A[m],
//     E1[L] = sum_{1<=k<L} PH[k]^2, Q[i] = sum_{j<i} (PH[n] - PH[j])^2,
//     PH = prefix sums of H;
//   mean(L) = S1/(100 n_w),  var(L) = (S2_H n_w - S1_H^2) / (1e4 n_w^2),
//   S1_H = S1 - c L n_w, n_w = n - L + 1 (the variance is shift-invariant, so
//   centring removes its cancellation).
// The autocorrelation is a complex FFT of the zero-padded H of length
// N = next_pow2(2n - 1), |X|^2, and the inverse FFT; two signals of the same
// shard share each complex FFT (one as the real, one as the imaginary part).
// The FFTs are radix-8 passes (after one radix-2 or radix-4 pass) through a
// swizzled, bank-conflict-free LDS buffer; forward transforms are
// decimation in frequency (natural order in, bit-reversed positions out) and
// inverse ones either decimation in time from those positions or DIF again
// with the bit reversal folded into the output addresses, so no pass
// reorders data.  One block per signal pair when N fits in one block's LDS
// (8192 complex values where the device allows 128 KiB per block), otherwise
// the four-step split N = N1 * N2 into column FFTs (length N2), row FFTs with
// the four-step twiddles, the power spectrum and the inverse row FFTs fused
// (a row and its mirror row per block), and inverse column FFTs.  Signal
// pairs go through the three passes in groups whose workspace stays small
// (128 MiB by default).
// A is FP64; when every value is within 0.25 of an integer and A[0] equals
// sum H^2 it is rounded to the exact integers.  Prefix sums are double-double
// (two-sum compensated), so S2 keeps ~32 significant digits.  Tables are
// written for the reference's stored block sizes (src/fourier/moments.rs:225-247).
#include <algorithm>
#include <cmath>
#include <type_traits>

#include "engine.hpp"

namespace chk {
namespace {

constexpr int NT = 256;
constexpr int NTF = 1024;            // most threads of an FFT block (one radix-8 unit each)
constexpr uint32_t LOG_MAXM = 13;
constexpr uint32_t FFT_MAXM = 1u << LOG_MAXM;  // largest LDS-resident FFT (complex FP64)
constexpr uint32_t TW = FFT_MAXM;              // twiddle table period

// ------------------------------------------------------------ double-double
struct dd {
    double hi, lo;
};
__device__ inline dd two_sum(double a, double b) {
    double s = a + b, bb = s - a;
    return {s, (a - (s - bb)) + (b - bb)};
}
__device__ inline dd dd_add(dd a, dd b) {
    dd s = two_sum(a.hi, b.hi);
    double lo = s.lo + a.lo + b.lo;
    double hi = s.hi + lo;
    return {hi, lo - (hi - s.hi)};
}
__device__ inline dd dd_neg(dd a) { return {-a.hi, -a.lo}; }
__device__ inline dd dd_from_prod(double a, double b) {  // exact product
    double p = a * b;
    return {p, fma(a, b, -p)};
}
__device__ inline dd dd_mul_d(dd a, double b) {
    dd p = dd_from_prod(a.hi, b);
    double lo = p.lo + a.lo * b;
    double hi = p.hi + lo;
    return {hi, lo - (hi - p.hi)};
}
__device__ inline dd dd_mul(dd a, dd b) {
    dd p = dd_from_prod(a.hi, b.hi);
    double lo = p.lo + a.hi * b.lo + a.lo * b.hi;
    double hi = p.hi + lo;
    return {hi, lo - (hi - p.hi)};
}
__device__ inline dd dd_from_i128(__int128 v) {
    double hi = (double)v;
    __int128 r = v - (__int128)hi;
    return {hi, (double)r};
}
__device__ inline double dd_div_to_d(dd a, double b) { return (a.hi + a.lo) / b; }

__device__ inline uint32_t bitrev(uint32_t i, uint32_t logm) {
    return logm ? __brev(i) >> (32 - logm) : 0u;
}

__device__ inline double2 cmul(double2 a, double2 b) {
    return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}

// Twiddles W^e = exp(-2 pi i e / TW), e < TW / 2, from two 64-entry LDS tables
// (twa[j] = W^j, twb[i] = W^(64 i)): one complex product instead of a load
// from the TW-entry table in global memory in every pass.
__device__ inline double2 lds_twiddle(const double2* twa, const double2* twb, uint32_t e) {
    return cmul(twa[e & 63u], twb[e >> 6]);
}

// Fills the LDS twiddle tables from the global TW table (the caller's barrier
// orders them before use).
__device__ inline void load_lds_twiddles(const double2* tw, double2* twa, double2* twb) {
    for (uint32_t i = threadIdx.x; i < 128; i += blockDim.x) {
        if (i < 64) twa[i] = tw[i];
        else twb[i - 64] = tw[64 * (i - 64)];
    }
}

__device__ inline double2 cadd(double2 a, double2 b) { return {a.x + b.x, a.y + b.y}; }
__device__ inline double2 csub(double2 a, double2 b) { return {a.x - b.x, a.y - b.y}; }

// LDS layout of nfft stacked length-M FFTs, element e = f * M + n in 16-byte
// slots: slot(e) = e XOR (a linear function of e's bits 3..6 and of the
// FFT index f = e >> logM) in bits 0..3.  Each bit only takes in higher bits,
// so the map is a bijection that keeps every aligned 16-element group in
// place.  With it every access pattern of the passes below is free of bank
// conflicts, whether the LDS serves a wave's lanes 8 or 16 at a time: runs of
// consecutive elements, radix-8 units of span 8 (groups of 8 lanes 64
// elements apart), units of span 1 (8 consecutive elements per lane), and
// the first and last passes of column blocks, whose lanes walk the FFT index
// fastest.
__device__ inline uint32_t swz_mask(uint32_t e) {  // from bits 3..6 of e
    return ((e >> 2) & 2u) ^ (((e >> 4) & 1u) * 3u) ^ ((e >> 3) & 4u) ^ ((e >> 3) & 8u);
}
__device__ inline uint32_t swz_fmask(uint32_t f) {  // from bits 0..3 of the FFT index
    return ((f & 1u) << 2) ^ (((f >> 1) & 1u) * 11u) ^ ((f >> 1) & 2u) ^ ((f >> 3) & 1u);
}
__device__ inline uint32_t lds_slot(uint32_t e, uint32_t logM) {
    return e ^ swz_mask(e) ^ (logM >= 4 ? swz_fmask(e >> logM) : 0u);
}

// Element n of FFT f in the swizzled LDS buffer, as a load and a store.
struct LdsIO {
    double2* xs;
    uint32_t logM;
    __device__ double2 operator()(uint32_t f, uint32_t n) const { return xs[lds_slot((f << logM) + n, logM)]; }
    __device__ void operator()(uint32_t f, uint32_t n, double2 v) const {
        xs[lds_slot((f << logM) + n, logM)] = v;
    }
};

// v * W_4^1 (-i forward, +i inverse)
template <bool INV>
__device__ inline double2 rot4(double2 v) {
    return INV ? double2{-v.y, v.x} : double2{v.y, -v.x};
}

template <bool INV>
__device__ inline double2 tw_lds(const double2* twa, const double2* twb, uint32_t e) {
    double2 w = lds_twiddle(twa, twb, e);
    if (INV) w.y = -w.y;
    return w;
}

// Radix-R butterfly of one unit a[i] = x[base + i hp], k = base mod hp,
// hp = 2^loghp.  DIF (natural order in, bit-reversed out): radix-2 stages of
// span R/2 hp down to hp, each twiddling the difference.  DIT (bit-reversed in,
// natural out): spans hp up to R/2 hp, each twiddling the second input.
// Stage of span s uses W_(2s)^(k + j hp) for the unit's pair j; with
// w_s = W_(2s)^k that is w_s times W_4^1 or W_8^j.
template <int R, bool DIF, bool INV>
__device__ inline void butterfly(double2 (&a)[R], uint32_t k, uint32_t loghp, const double2* twa,
                                 const double2* twb) {
    constexpr double R2 = 0.70710678118654752440;  // 1 / sqrt(2)
    const double2 w1 = tw_lds<INV>(twa, twb, k << (LOG_MAXM - 1 - loghp));  // W_(2hp)^k
    if constexpr (R == 2) {
        if (DIF) {
            const double2 d = csub(a[0], a[1]);
            a[0] = cadd(a[0], a[1]);
            a[1] = cmul(d, w1);
        } else {
            const double2 t = cmul(a[1], w1);
            a[1] = csub(a[0], t);
            a[0] = cadd(a[0], t);
        }
        return;
    }
    const double2 w2 = tw_lds<INV>(twa, twb, k << (LOG_MAXM - 2 - loghp));  // W_(4hp)^k
    const double2 w2b = rot4<INV>(w2);                                      // W_(4hp)^(k+hp)
    if constexpr (R == 4) {
        if (DIF) {
            double2 d = csub(a[0], a[2]);
            a[0] = cadd(a[0], a[2]);
            a[2] = cmul(d, w2);
            d = csub(a[1], a[3]);
            a[1] = cadd(a[1], a[3]);
            a[3] = cmul(d, w2b);
#pragma unroll
            for (int i = 0; i < 4; i += 2) {
                d = csub(a[i], a[i + 1]);
                a[i] = cadd(a[i], a[i + 1]);
                a[i + 1] = cmul(d, w1);
            }
        } else {
#pragma unroll
            for (int i = 0; i < 4; i += 2) {
                const double2 t = cmul(a[i + 1], w1);
                a[i + 1] = csub(a[i], t);
                a[i] = cadd(a[i], t);
            }
            double2 t = cmul(a[2], w2);
            a[2] = csub(a[0], t);
            a[0] = cadd(a[0], t);
            t = cmul(a[3], w2b);
            a[3] = csub(a[1], t);
            a[1] = cadd(a[1], t);
        }
        return;
    }
    // R == 8: span 4hp uses W_(8hp)^k W_8^j, j < 4
    const double2 w3 = tw_lds<INV>(twa, twb, k << (LOG_MAXM - 3 - loghp));
    const double2 v1 = cmul(w3, INV ? double2{R2, R2} : double2{R2, -R2});
    const double2 v[4] = {w3, v1, rot4<INV>(w3), rot4<INV>(v1)};
    if (DIF) {
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const double2 d = csub(a[i], a[i + 4]);
            a[i] = cadd(a[i], a[i + 4]);
            a[i + 4] = cmul(d, v[i]);
        }
#pragma unroll
        for (int i = 0; i < 8; i += 4) {
            double2 d = csub(a[i], a[i + 2]);
            a[i] = cadd(a[i], a[i + 2]);
            a[i + 2] = cmul(d, w2);
            d = csub(a[i + 1], a[i + 3]);
            a[i + 1] = cadd(a[i + 1], a[i + 3]);
            a[i + 3] = cmul(d, w2b);
        }
#pragma unroll
        for (int i = 0; i < 8; i += 2) {
            const double2 d = csub(a[i], a[i + 1]);
            a[i] = cadd(a[i], a[i + 1]);
            a[i + 1] = cmul(d, w1);
        }
    } else {
#pragma unroll
        for (int i = 0; i < 8; i += 2) {
            const double2 t = cmul(a[i + 1], w1);
            a[i + 1] = csub(a[i], t);
            a[i] = cadd(a[i], t);
        }
#pragma unroll
        for (int i = 0; i < 8; i += 4) {
            double2 t = cmul(a[i + 2], w2);
            a[i + 2] = csub(a[i], t);
            a[i] = cadd(a[i], t);
            t = cmul(a[i + 3], w2b);
            a[i + 3] = csub(a[i + 1], t);
            a[i + 1] = cadd(a[i + 1], t);
        }
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const double2 t = cmul(a[i + 4], v[i]);
            a[i + 4] = csub(a[i], t);
            a[i] = cadd(a[i], t);
        }
    }
}

// One radix-R pass over nfft stacked length-M FFTs (span hp = 2^loghp):
// unit (f, k, group g) reads element g R hp + k + i hp, i < R, of FFT f via
// ld(f, n), and writes it back via st(f, n, v).  FMAJ: consecutive threads
// take consecutive FFTs (column blocks: their elements are adjacent in
// memory), else consecutive units of one FFT.
template <int R, bool DIF, bool INV, bool FMAJ, class Ld, class St>
__device__ inline void fft_pass(uint32_t logM, uint32_t lognfft, uint32_t loghp, const Ld& ld, const St& st,
                                const double2* twa, const double2* twb) {
    constexpr uint32_t LR = R == 2 ? 1 : (R == 4 ? 2 : 3);
    constexpr bool LDS_IN = std::is_same<Ld, LdsIO>::value, LDS_OUT = std::is_same<St, LdsIO>::value;
    const uint32_t logu = logM - LR;  // units per FFT
    const uint32_t nu = 1u << (logu + lognfft);
    // LDS slots of a unit: its first element E0 has no bits where i << loghp
    // has, and the swizzle is linear, so slot(E0 + (i << loghp)) =
    // slot(E0) ^ c_i with c_i = (i << loghp) ^ swz_mask(i << loghp), the same
    // for every unit of the pass
    uint32_t c[R];
#pragma unroll
    for (int i = 0; i < R; i++) c[i] = ((uint32_t)i << loghp) ^ swz_mask((uint32_t)i << loghp);
    for (uint32_t j = threadIdx.x; j < nu; j += blockDim.x) {
        const uint32_t f = FMAJ ? (j & ((1u << lognfft) - 1u)) : (j >> logu);
        const uint32_t jj = FMAJ ? (j >> lognfft) : (j & ((1u << logu) - 1u));
        const uint32_t k = jj & ((1u << loghp) - 1u);
        const uint32_t base = ((jj >> loghp) << (loghp + LR)) + k;
        const uint32_t s0 = lds_slot((f << logM) + base, logM);
        double2 a[R];
#pragma unroll
        for (int i = 0; i < R; i++) {
            if constexpr (LDS_IN) a[i] = ld.xs[s0 ^ c[i]];
            else a[i] = ld(f, base + ((uint32_t)i << loghp));
        }
        butterfly<R, DIF, INV>(a, k, loghp, twa, twb);
#pragma unroll
        for (int i = 0; i < R; i++) {
            if constexpr (LDS_OUT) st.xs[s0 ^ c[i]] = a[i];
            else st(f, base + ((uint32_t)i << loghp), a[i]);
        }
    }
}

// DIF FFTs of nfft = 2^lognfft stacked length-2^logM signals: natural order
// in through in(f, n), out(f, p, v) at bit-reversed positions p, the LDS in
// between.  The radix-2 or radix-4 remainder of log2 M (largest spans) comes
// first, then radix-8 passes down to span 1; the first pass reads its unit
// straight from `in` and the last writes straight to `out`.  FMAJ_IO gives
// those two passes the FFT-major thread order.
template <bool INV, bool FMAJ_IO, class In, class Out>
__device__ void fft_dif(double2* xs, uint32_t logM, uint32_t lognfft, const In& in, const Out& out,
                        const double2* twa, const double2* twb) {
    const LdsIO L{xs, logM};
    const uint32_t rem = logM % 3, lr0 = rem ? rem : 3u, lh0 = logM - lr0;
    if (lh0 == 0) {  // one pass
        if (lr0 == 1) fft_pass<2, true, INV, FMAJ_IO>(logM, lognfft, 0, in, out, twa, twb);
        else if (lr0 == 2) fft_pass<4, true, INV, FMAJ_IO>(logM, lognfft, 0, in, out, twa, twb);
        else fft_pass<8, true, INV, FMAJ_IO>(logM, lognfft, 0, in, out, twa, twb);
        return;
    }
    if (lr0 == 1) fft_pass<2, true, INV, FMAJ_IO>(logM, lognfft, lh0, in, L, twa, twb);
    else if (lr0 == 2) fft_pass<4, true, INV, FMAJ_IO>(logM, lognfft, lh0, in, L, twa, twb);
    else fft_pass<8, true, INV, FMAJ_IO>(logM, lognfft, lh0, in, L, twa, twb);
    __syncthreads();
    for (uint32_t lh = lh0 - 3; lh > 0; lh -= 3) {
        fft_pass<8, true, INV, false>(logM, lognfft, lh, L, L, twa, twb);
        __syncthreads();
    }
    fft_pass<8, true, INV, FMAJ_IO>(logM, lognfft, 0, L, out, twa, twb);
}

// DIT FFTs of the stacked signals held in the LDS at bit-reversed positions,
// written in natural order through out(f, n, v): radix-8 passes from span 1
// up, then the radix-2 or radix-4 remainder (largest spans).
template <bool INV, class Out>
__device__ void fft_dit(double2* xs, uint32_t logM, uint32_t lognfft, const Out& out, const double2* twa,
                        const double2* twb) {
    const LdsIO L{xs, logM};
    const uint32_t rem = logM % 3, n8 = logM / 3;
    for (uint32_t t = 0; t < n8; t++) {
        if (t + 1 == n8 && rem == 0) {
            fft_pass<8, false, INV, false>(logM, lognfft, 3 * t, L, out, twa, twb);
            return;
        }
        fft_pass<8, false, INV, false>(logM, lognfft, 3 * t, L, L, twa, twb);
        __syncthreads();
    }
    if (rem == 1) fft_pass<2, false, INV, false>(logM, lognfft, logM - 1, L, out, twa, twb);
    else fft_pass<4, false, INV, false>(logM, lognfft, logM - 2, L, out, twa, twb);
}

struct Rows {
    const int32_t* G;      // [nb][ldg] coverage per bin
    uint32_t ldg;
    const int64_t* cen;    // [nb] centring constant c
    double2* X;            // [(nb + 1) / 2][N] FFT workspace, one signal pair per row
    double* A;             // [nb][n] autocorrelation output
    uint32_t nb, n, N, logN;
    const double2* tw;     // exp(-2 pi i t / TW), t < TW
    const double2* twN;    // exp(-2 pi i t / N), t < FFT_MAXM (four-step only)
};

// exp(-+2 pi i e / N) for e < N <= 2^(2 LOG_MAXM): W_N^(e_hi 2^LOG_MAXM) from the
// TW table (it is W_(N / 2^LOG_MAXM)^e_hi) times W_N^e_lo from the per-N table.
__device__ inline double2 twiddle(const Rows& r, uint64_t e, bool inv) {
    const double2 hi = r.tw[(e >> LOG_MAXM) << (2 * LOG_MAXM - r.logN)];
    const double2 lo = r.twN[e & (FFT_MAXM - 1)];
    double2 w = cmul(hi, lo);
    if (inv) w.y = -w.y;
    return w;
}

// Two real signals share one complex FFT: z = h1 + i h2 with h = G - c of
// rows 2p and 2p + 1 (zero past n and for a missing odd row).  From a = Z[k],
// b = Z[N - k]: H1 = (a + conj b) / 2, H2 = (a - conj b) / 2i, and
// Y[k] = |H1|^2 + i |H2|^2 = Y[N - k]; the inverse FFT of Y is A1 + i A2,
// both autocorrelations being real.
__device__ inline double2 pair_power(double2 a, double2 b) {
    const double sr = a.x + b.x, si = a.y - b.y;
    const double dr = a.x - b.x, di = a.y + b.y;
    return {0.25 * (sr * sr + si * si), 0.25 * (dr * dr + di * di)};
}

struct PairIn {
    const int32_t* g0;
    const int32_t* g1;  // nullptr for a missing odd row
    double c0, c1;
    uint32_t n;
    __device__ PairIn(const Rows& r, uint32_t p) {
        g0 = r.G + (size_t)(2 * p) * r.ldg;
        c0 = (double)r.cen[2 * p];
        const bool two = 2 * p + 1 < r.nb;
        g1 = two ? g0 + r.ldg : nullptr;
        c1 = two ? (double)r.cen[2 * p + 1] : 0.0;
        n = r.n;
    }
    __device__ double2 operator()(uint64_t i) const {
        if (i >= n) return {0.0, 0.0};
        return {(double)g0[i] - c0, g1 ? (double)g1[i] - c1 : 0.0};
    }
};

struct PairOut {
    double* a0;
    double* a1;  // nullptr for a missing odd row
    double scale;
    __device__ PairOut(const Rows& r, uint32_t p) {
        a0 = r.A + (size_t)(2 * p) * r.n;
        a1 = 2 * p + 1 < r.nb ? a0 + r.n : nullptr;
        scale = 1.0 / (double)r.N;
    }
    __device__ void operator()(uint64_t i, double2 v) const {
        a0[i] = v.x * scale;
        if (a1) a1[i] = v.y * scale;
    }
};

// Whole autocorrelation of one signal pair in LDS (N <= FFT_MAXM): DIF
// forward FFT, Y = pair_power at the bit-reversed positions (Z[k] at bitrev(k)
// pairs with Z[N - k]), DIT inverse FFT back to natural order.
__global__ __launch_bounds__(NTF) void autocorr_small(Rows r) {
    extern __shared__ __align__(16) double2 xs[];
    __shared__ double2 twa[64], twb[64];
    load_lds_twiddles(r.tw, twa, twb);
    __syncthreads();
    const uint32_t p = blockIdx.x;
    const PairIn in(r, p);
    const LdsIO L{xs, r.logN};
    fft_dif<false, false>(xs, r.logN, 0, [&](uint32_t, uint32_t n) { return in(n); }, L, twa, twb);
    __syncthreads();
    for (uint32_t k = threadIdx.x; k <= r.N / 2; k += blockDim.x) {
        const uint32_t pa = bitrev(k, r.logN), pb = bitrev((r.N - k) & (r.N - 1), r.logN);
        const double2 y = pair_power(L(0, pa), L(0, pb));
        L(0, pa, y);
        L(0, pb, y);
    }
    __syncthreads();
    const PairOut out(r, p);
    fft_dit<true>(
        xs, r.logN, 0,
        [&](uint32_t, uint32_t n, double2 v) {
            if (n < r.n) out(n, v);
        },
        twa, twb);
}

// Column blocks read and write CB adjacent columns (a few dozen bytes per
// row), so neighbouring blocks share cache lines.  The hardware deals blocks
// round-robin over the 8 XCDs, each with its own L2: block b runs on XCD b % 8.
// Logical block (b % 8) * per-XCD count + b / 8 (the remainder spread over the
// first XCDs) gives each XCD a run of consecutive column groups, whose lines
// its L2 then fetches once.  The mapping is a bijection for any grid size.
__device__ inline uint32_t xcd_block(uint32_t b, uint32_t nb) {
    const uint32_t x = b & 7u, l = b >> 3, q = nb >> 3, r = nb & 7u;
    return (x < r ? x * (q + 1) : r * (q + 1) + (x - r) * q) + l;
}

// Four-step pass 1: CB = 2^logCB adjacent columns n1 of one signal pair,
// length-N2 DIF FFTs over n2 (element n1 + N1*n2); output position p holds
// k2 = bitrev(p), stored at n1 + N1*k2 (the twiddle W_N^(n1 k2) is applied
// by the row pass, where k2 is fixed).
__global__ __launch_bounds__(NTF) void fft_cols_fwd(Rows r, uint32_t N1, uint32_t logN2, uint32_t logCB) {
    extern __shared__ __align__(16) double2 xs[];
    __shared__ double2 twa[64], twb[64];
    load_lds_twiddles(r.tw, twa, twb);
    __syncthreads();
    const uint32_t ncb = N1 >> logCB, bl = xcd_block(blockIdx.x, gridDim.x);
    const uint32_t p = bl / ncb, c0 = (bl % ncb) << logCB;
    const PairIn in(r, p);
    double2* X = r.X + (size_t)p * r.N;
    fft_dif<false, true>(
        xs, logN2, logCB, [&](uint32_t f, uint32_t n2) { return in((uint64_t)(c0 + f) + (uint64_t)N1 * n2); },
        [&](uint32_t f, uint32_t pos, double2 v) { X[(c0 + f) + (size_t)N1 * bitrev(pos, logN2)] = v; },
        twa, twb);
}

// Block barrier that orders LDS traffic only: global loads issued for the
// next column group stay in flight across it (__syncthreads would wait for
// them).
__device__ inline void lds_barrier() { asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory"); }

constexpr int BUF_DWORD3 = 0x00020000;  // raw buffer, gfx9-family descriptor word 3

__device__ inline __amdgpu_buffer_rsrc_t make_rsrc(const void* base, uint32_t bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), (short)0, (int)bytes, BUF_DWORD3);
}

// Persistent, software-pipelined form of pass 1 for CB * N2 = 8 * blockDim
// (every thread takes 8 values in the first radix-2^LR pass, whatever N2):
// block b walks column groups b, b + gridDim, ... (gridDim a multiple of 8,
// so a group runs on the XCD xcd_block assigns it), and the coverage values
// of the next group's first pass are loaded into registers while the current
// group runs its LDS passes and stores.  Buffer loads past a row's n bins
// (or of a missing odd row) read 0 and need no branch, so the compiler
// leaves them in flight across the LDS-only barriers.
template <int LR>
__global__ __launch_bounds__(NTF) void fft_cols_fwd_p(Rows r, uint32_t N1, uint32_t logN2, uint32_t logCB,
                                                      uint32_t ngrp) {
    constexpr int R = 1 << LR, U = 8 / R;  // first-pass radix, units per thread
    extern __shared__ __align__(16) double2 xs[];
    __shared__ double2 twa[64], twb[64];
    load_lds_twiddles(r.tw, twa, twb);
    __syncthreads();
    const uint32_t logM = logN2, lognfft = logCB, loghp = logM - LR;
    const uint32_t ncb = N1 >> logCB, nfm = (1u << lognfft) - 1u, hpm = (1u << loghp) - 1u;
    // the thread's first-pass units: FFT index f (column c0 + f), k, base
    uint32_t uf[U], uk[U], ub[U];
#pragma unroll
    for (int u = 0; u < U; u++) {
        const uint32_t j = threadIdx.x + (uint32_t)u * blockDim.x;
        uf[u] = j & nfm;
        const uint32_t jj = j >> lognfft;
        uk[u] = jj & hpm;
        ub[u] = ((jj >> loghp) << (loghp + LR)) + uk[u];
    }
    uint32_t c[R];
#pragma unroll
    for (int i = 0; i < R; i++) c[i] = ((uint32_t)i << loghp) ^ swz_mask((uint32_t)i << loghp);
    // coverage of group g's first pass: rows 2p and 2p + 1 at n1 + N1 n2
    auto fetch = [&](uint32_t g, int32_t (&v0)[8], int32_t (&v1)[8]) {
        const uint32_t bl = xcd_block(g, ngrp), p = bl / ncb, c0 = (bl % ncb) << logCB;
        const int32_t* g0 = r.G + (size_t)(2 * p) * r.ldg;
        const bool two = 2 * p + 1 < r.nb;
        const auto q0 = make_rsrc(g0, r.n * 4u);
        const auto q1 = make_rsrc(two ? g0 + r.ldg : g0, two ? r.n * 4u : 0u);
#pragma unroll
        for (int u = 0; u < U; u++)
#pragma unroll
            for (int i = 0; i < R; i++) {
                const uint32_t idx = c0 + uf[u] + N1 * (ub[u] + ((uint32_t)i << loghp));
                v0[u * R + i] = (int32_t)__builtin_amdgcn_raw_buffer_load_b32(q0, idx * 4u, 0, 0);
                v1[u * R + i] = (int32_t)__builtin_amdgcn_raw_buffer_load_b32(q1, idx * 4u, 0, 0);
            }
    };
    const LdsIO L{xs, logM};
    int32_t a0[8], a1[8];
    uint32_t g = blockIdx.x;
    if (g < ngrp) fetch(g, a0, a1);
    for (; g < ngrp; g += gridDim.x) {
        const uint32_t bl = xcd_block(g, ngrp), p = bl / ncb, c0 = (bl % ncb) << logCB;
        const double cen0 = (double)r.cen[2 * p];
        const double cen1 = 2 * p + 1 < r.nb ? (double)r.cen[2 * p + 1] : 0.0;
        int32_t b0[8], b1[8];
        const uint32_t gn = g + gridDim.x;
        fetch(gn < ngrp ? gn : g, b0, b1);  // past the last group: reloaded, unused
        // pass 1 from the registers: H = G - c below n bins, 0 past them
#pragma unroll
        for (int u = 0; u < U; u++) {
            double2 x[R];
#pragma unroll
            for (int i = 0; i < R; i++) {
                const uint32_t idx = c0 + uf[u] + N1 * (ub[u] + ((uint32_t)i << loghp));
                const bool in = idx < r.n;
                x[i] = {in ? (double)a0[u * R + i] - cen0 : 0.0, in ? (double)a1[u * R + i] - cen1 : 0.0};
            }
            butterfly<R, true, false>(x, uk[u], loghp, twa, twb);
            const uint32_t s0 = lds_slot((uf[u] << logM) + ub[u], logM);
#pragma unroll
            for (int i = 0; i < R; i++) xs[s0 ^ c[i]] = x[i];
        }
        lds_barrier();
        for (uint32_t lh = loghp - 3; lh > 0; lh -= 3) {
            fft_pass<8, true, false, false>(logM, lognfft, lh, L, L, twa, twb);
            lds_barrier();
        }
        double2* X = r.X + (size_t)p * r.N;
        fft_pass<8, true, false, true>(
            logM, lognfft, 0, L,
            [&](uint32_t f, uint32_t pos, double2 v) { X[(c0 + f) + (size_t)N1 * bitrev(pos, logN2)] = v; },
            twa, twb);
        lds_barrier();  // the last pass's LDS reads are done before the next group's writes
#pragma unroll
        for (int k = 0; k < 8; k++) {
            a0[k] = b0[k];
            a1[k] = b1[k];
        }
    }
}

// Four-step pass 2: the rows k2 and N2 - k2 of one signal pair (row k2 holds
// Z[k2 + N2 k1], so Z[N - k] of its entry k1 is entry N1 - 1 - k1 of row
// N2 - k2, which after the DIF sits at the complementary bit-reversed
// position; rows 0 and N2/2 pair with themselves): twiddle W_N^(n1 k2) on
// load, forward FFTs, Y, inverse FFTs, twiddle W_N^(-n1 k2) on store, in
// place.  The row's twiddles come from two 64-entry LDS factors,
// W_N^(n1 k2) = W_N^(64 a k2) W_N^(b k2) for n1 = 64 a + b.
__global__ __launch_bounds__(NTF) void fft_rows(Rows r, uint32_t N1, uint32_t logN1, uint32_t N2) {
    extern __shared__ __align__(16) double2 xs[];
    __shared__ double2 twa[64], twb[64];
    __shared__ double2 rtw[2][128];  // per row: [a] = W_N^(64 a k2), [64 + b] = W_N^(b k2)
    load_lds_twiddles(r.tw, twa, twb);
    const uint32_t per = N2 / 2 + 1;
    const uint32_t p = blockIdx.x / per, q = blockIdx.x % per;
    const uint32_t nr = (q == 0 || q == N2 / 2) ? 1 : 2;
    for (uint32_t t = threadIdx.x; t < nr * 128; t += blockDim.x) {
        const uint32_t j = t >> 7, i = t & 127u;
        const uint64_t row = j ? N2 - q : q;
        if (i >= 64 || 64 * i < N1) rtw[j][i] = twiddle(r, (i < 64 ? 64ull * i : i - 64ull) * row, false);
    }
    __syncthreads();
    auto rowtw = [&](uint32_t f, uint32_t n) { return cmul(rtw[f][n >> 6], rtw[f][64 + (n & 63u)]); };
    double2* X = r.X + (size_t)p * r.N;
    const LdsIO L{xs, logN1};
    fft_dif<false, false>(
        xs, logN1, nr - 1,
        [&](uint32_t f, uint32_t n) { return cmul(X[(size_t)(f ? N2 - q : q) * N1 + n], rowtw(f, n)); }, L,
        twa, twb);
    __syncthreads();
    if (nr == 2) {
        for (uint32_t t = threadIdx.x; t < N1; t += blockDim.x) {
            const double2 y = pair_power(L(0, t), L(1, N1 - 1 - t));
            L(0, t, y);
            L(1, N1 - 1 - t, y);
        }
    } else if (q == 0) {  // k1 pairs with N1 - k1 (mod N1)
        for (uint32_t k = threadIdx.x; k <= N1 / 2; k += blockDim.x) {
            const uint32_t pa = bitrev(k, logN1), pb = bitrev((N1 - k) & (N1 - 1), logN1);
            const double2 y = pair_power(L(0, pa), L(0, pb));
            L(0, pa, y);
            L(0, pb, y);
        }
    } else {  // k1 pairs with N1 - 1 - k1
        for (uint32_t t = threadIdx.x; t < N1 / 2; t += blockDim.x) {
            const double2 y = pair_power(L(0, t), L(0, N1 - 1 - t));
            L(0, t, y);
            L(0, N1 - 1 - t, y);
        }
    }
    __syncthreads();
    fft_dit<true>(
        xs, logN1, nr - 1,
        [&](uint32_t f, uint32_t n1, double2 v) {
            double2 w = rowtw(f, n1);
            w.y = -w.y;
            X[(size_t)(f ? N2 - q : q) * N1 + n1] = cmul(v, w);
        },
        twa, twb);
}

// Four-step pass 3: CB columns, inverse length-N2 DIF FFTs over k2; position
// p holds n2 = bitrev(p): A[n1 + N1*n2] = Re / N (first row), Im / N (second
// row) below n.
__global__ __launch_bounds__(NTF) void fft_cols_inv(Rows r, uint32_t N1, uint32_t logN2, uint32_t logCB) {
    extern __shared__ __align__(16) double2 xs[];
    __shared__ double2 twa[64], twb[64];
    load_lds_twiddles(r.tw, twa, twb);
    __syncthreads();
    const uint32_t ncb = N1 >> logCB, bl = xcd_block(blockIdx.x, gridDim.x);
    const uint32_t p = bl / ncb, c0 = (bl % ncb) << logCB;
    const double2* X = r.X + (size_t)p * r.N;
    const PairOut out(r, p);
    fft_dif<true, true>(
        xs, logN2, logCB, [&](uint32_t f, uint32_t k2) { return X[(c0 + f) + (size_t)N1 * k2]; },
        [&](uint32_t f, uint32_t pos, double2 v) {
            const uint64_t idx = (uint64_t)(c0 + f) + (uint64_t)N1 * bitrev(pos, logN2);
            if (idx < r.n) out(idx, v);
        },
        twa, twb);
}

// ------------------------------------------------------------ binning

// +coverage impulses at each start and -impulses at each end of the
// intervals of one shard (layered table order), split over the two bins a
// coordinate straddles: bin x/100 gets 100 - x%100, the next bin x%100.
__global__ void impulses(const int32_t* start, const int32_t* end, const uint32_t* fid,
                         uint64_t cnt, const int32_t* slot_of_fid, uint32_t n, uint32_t ldg,
                         int32_t* D) {
    const uint32_t max_bp = n * 100u;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < cnt;
         i += (uint64_t)gridDim.x * blockDim.x) {
        const int32_t s = slot_of_fid[fid[i]];
        if (s < 0) continue;
        int32_t* row = D + (size_t)s * ldg;
#pragma unroll
        for (int e = 0; e < 2; e++) {
            uint32_t x = (uint32_t)(e ? end[i] : start[i]);
            const int32_t sign = e ? -1 : 1;
            x = x < max_bp ? x : max_bp;
            const uint32_t bin = x / 100u, rem = x % 100u;
            if (bin < n) atomicAdd(&row[bin], sign * (int32_t)(100u - rem));
            if (rem > 0 && bin + 1 < n) atomicAdd(&row[bin + 1], sign * (int32_t)rem);
        }
    }
}

template <typename T>
__device__ T block_incl_scan(T v, T* lds, T& total) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        T y = __shfl_up(v, d, 64);
        if (lane >= d) v += y;
    }
    if (lane == 63) lds[w] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        T run = 0;
        for (int i = 0; i < NT / 64; i++) {
            T t = lds[i];
            lds[i] = run;
            run += t;
        }
        lds[NT / 64] = run;
    }
    __syncthreads();
    v += lds[w];
    total = lds[NT / 64];
    __syncthreads();
    return v;
}

constexpr uint32_t EK = 16;       // consecutive positions per thread in the evaluation scan
constexpr uint32_t CH = NT * EK;  // positions per chunk
// A row's positions 0..n are scanned in segments of EVAL_SEG positions (a
// multiple of CH) by default, at most MAXSEG per row.
constexpr uint32_t EVAL_SEG = 65536;
constexpr uint32_t MAXSEG = 256;

// One block per row: G = inclusive prefix of the impulse row (in place), the
// centring constant c = round(sum G / n) and sum H^2 = sum (G - c)^2, exact
// (from sum G^2 in 128-bit integers), as a double-double.  The row goes by
// in chunks of NT * RK bins staged in LDS (the next chunk's loads in flight
// meanwhile); a thread scans its run of RK consecutive bins serially, so a
// chunk costs one block scan.  Also writes the sum of G over each segment of
// seg bins (seg a multiple of the chunk), gseg[row * nseg + k] for segment k.
__global__ __launch_bounds__(NT) void coverage_rows(int32_t* D, uint32_t ldg, uint32_t n,
                                                    int64_t* cen, int64_t* sumg, dd* sumh2,
                                                    uint32_t seg, uint32_t nseg, int64_t* gseg) {
    constexpr uint32_t RK = 16, RCH = NT * RK;
    static_assert(RCH == CH, "segments are whole chunks of both scans");
    __shared__ int64_t lds[NT / 64 + 1];
    __shared__ int32_t sv[RCH + RCH / RK];  // one pad word after every run
    __shared__ unsigned __int128 l2[NT];
    int32_t* row = D + (size_t)blockIdx.x * ldg;
    int64_t carry = 0, sg = 0, sgs = 0;
    unsigned __int128 sg2 = 0;
    int32_t pv[RK];
    auto fetch = [&](uint32_t base) {
#pragma unroll
        for (uint32_t m = 0; m < RK; m++) {
            const uint32_t i = base + threadIdx.x + m * NT;
            pv[m] = i < n ? row[i] : 0;
        }
    };
    fetch(0);
    for (uint32_t base = 0; base < n; base += RCH) {
        __syncthreads();  // the previous chunk's LDS reads are done
#pragma unroll
        for (uint32_t m = 0; m < RK; m++) {
            const uint32_t j = threadIdx.x + m * NT;
            sv[j + j / RK] = pv[m];
        }
        if (base + RCH < n) fetch(base + RCH);
        __syncthreads();
        const uint32_t r0 = threadIdx.x * (RK + 1);
        int64_t rs = 0;
#pragma unroll
        for (uint32_t k = 0; k < RK; k++) rs += sv[r0 + k];  // 0 past n
        int64_t tot;
        int64_t run = block_incl_scan<int64_t>(rs, lds, tot) - rs + carry;
#pragma unroll
        for (uint32_t k = 0; k < RK; k++) {
            run += sv[r0 + k];
            if (base + threadIdx.x * RK + k < n) {
                sv[r0 + k] = (int32_t)run;
                sgs += run;
                sg2 += (unsigned __int128)(uint64_t)(run * run);
            }
        }
        if ((base + RCH) % seg == 0 || base + RCH >= n) {  // the segment ends with this chunk
            int64_t segtot;
            (void)block_incl_scan<int64_t>(sgs, lds, segtot);
            if (threadIdx.x == 0) gseg[(size_t)blockIdx.x * nseg + base / seg] = segtot;
            sg += sgs;
            sgs = 0;
        }
        __syncthreads();
#pragma unroll
        for (uint32_t m = 0; m < RK; m++) {
            const uint32_t j = threadIdx.x + m * NT;
            if (base + j < n) row[base + j] = sv[j + j / RK];
        }
        carry += tot;
    }
    int64_t tot;
    (void)block_incl_scan<int64_t>(sg, lds, tot);
    l2[threadIdx.x] = sg2;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned __int128 s2 = 0;
        for (int t = 0; t < NT; t++) s2 += l2[t];
        const int64_t c = (int64_t)llround((double)tot / (double)n);
        sumg[blockIdx.x] = tot;
        cen[blockIdx.x] = c;
        sumh2[blockIdx.x] = dd_from_i128((__int128)s2 - 2 * (__int128)c * tot + (__int128)n * c * c);
    }
}

// ------------------------------------------------------------ evaluation
//
// The table entries need, at the stored block sizes L, prefix sums over the
// positions i = 0..n of a row: PP (sum of P[i] = sum_{j<i} G[j]), E1 (sum of
// ph^2, ph = P - c i), Q (sum of (PHn - ph)^2), C1 (sum of A[i]) and C2 (sum
// of A[i] i).  Q over i < m equals m PHn^2 - 2 PHn (PP - c m (m - 1) / 2) +
// E1 (ph = 0 at i = 0), so only PP and E1 are scanned on the G side, C1 and
// C2 on the A side: two scans of half the width run over the row's segments (seg
// positions each), each segment started from zero sums (and from P at its
// first position); eval_fix adds the earlier segments' totals.  When A is
// rounded every term is an integer or an exact product, so the sums do not
// depend on where the row is cut.

struct TermsG {  // sums of P, (P - c i)^2
    int64_t pp;
    dd e1;
};
struct TermsA {  // sums of A, A i
    dd c1, c2;
};

struct EvalArgs {
    const int32_t* G;
    uint32_t ldg, n, nl;
    const double* A;
    const int64_t* cen;
    const int64_t* sumg;
    const int32_t* pmapL;    // [n+2] position -> L index sampled at i = L, or -1
    const int32_t* pmapM;    // [n+2] position -> L index sampled at i = n-L+1, or -1
    const uint32_t* ls;      // [nl] stored block sizes
    const int64_t* out_off;  // [B] first table entry of the row's slot
    const uint8_t* runflag;  // [(n + 1) / EK + 1] bit 0: run r holds a position L, bit 1: n-L+1
    double* mean;
    double* var;
    // per row scratch of samples: [B][nl] each
    int64_t* sPP_L;
    int64_t* sPP_M;
    dd* sC1;
    dd* sC2;
    dd* sE1;
    dd* sE1M;                // E1 sampled at i = n-L+1
    uint32_t* exact;         // [B] 1 when A was rounded to integers
    const dd* sumh2;         // [B] sum H^2, exact
    uint32_t round_mode;     // 0 round within 0.25, 1 never round, 2 round only exact values
    uint32_t seg, nseg;      // positions per segment (a multiple of CH), segments per row
    const int64_t* gseg;     // [B][nseg] sum of G over each segment's positions
    TermsG* totg;            // [B][nseg] each segment's G-side totals
    TermsA* tota;            // [B][nseg] each segment's A-side totals
    uint32_t* segdev;        // [B][nseg] 1 when a rounded segment held a value off its integer by > tol
    uint32_t* redo;          // [B] 1 when the row's A is scanned again unrounded
};

// Whether a row's scan starts rounded: A[0] must round to sum H^2.
__device__ inline bool row_rounds(const EvalArgs& a, uint32_t b) {
    const dd sh2 = a.sumh2[b];
    const double A0r = a.A[(size_t)b * a.n];
    return a.round_mode != 1 && fabs((sh2.hi - rint(A0r)) + sh2.lo) <= 0.5;
}

__device__ inline dd shfl_up_dd(dd v, int d) {
    return {__shfl_up(v.hi, d, 64), __shfl_up(v.lo, d, 64)};
}
__device__ inline TermsG tadd(const TermsG& a, const TermsG& b) {
    return {a.pp + b.pp, dd_add(a.e1, b.e1)};
}
__device__ inline TermsA tadd(const TermsA& a, const TermsA& b) {
    return {dd_add(a.c1, b.c1), dd_add(a.c2, b.c2)};
}
__device__ inline TermsG tshfl(const TermsG& v, int d) {
    return {__shfl_up(v.pp, d, 64), shfl_up_dd(v.e1, d)};
}
__device__ inline TermsA tshfl(const TermsA& v, int d) {
    return {shfl_up_dd(v.c1, d), shfl_up_dd(v.c2, d)};
}

// Exclusive block scan of v; total receives the block's sum.
template <class T>
__device__ T block_excl_scan_t(T v, T* lds, T& total) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    T x = v;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        T y = tshfl(x, d);
        if (lane >= d) x = tadd(x, y);
    }
    if (lane == 63) lds[w] = x;
    __syncthreads();
    if (threadIdx.x == 0) {
        T run{};
        for (int i = 0; i < NT / 64; i++) {
            T t = lds[i];
            lds[i] = run;
            run = tadd(run, t);
        }
        lds[NT / 64] = run;
    }
    __syncthreads();
    T wex = tshfl(x, 1);  // wave-exclusive prefix
    if (lane == 0) wex = T{};
    T res = tadd(wex, lds[w]);
    total = lds[NT / 64];
    __syncthreads();
    return res;
}

// The G side, one block per (segment s, row b): positions [s seg, min((s + 1)
// seg, n + 1)), P started at the earlier segments' sum of G.  The segment
// goes by in chunks of CH positions: the chunk's G is staged in LDS with
// coalesced loads (one pad word after every run of EK, so that a thread's
// run is read without bank conflicts); a serial pass over the thread's run,
// block scans of P and of the run totals, and a second serial pass, only
// over runs that hold a sampled position, that emits the exclusive prefixes
// at the samples.  No registers are held for the next chunk's loads: the
// kernel fits 5 blocks per CU, whose loads cover each other's scans.
__global__ __launch_bounds__(NT) void eval_seg_g(EvalArgs a) {
    __shared__ TermsG lds[NT / 64 + 1];
    __shared__ int64_t ldsi[NT / 64 + 1];
    __shared__ int32_t sG[CH + CH / EK];
    const uint32_t s = blockIdx.x, b = blockIdx.y, n = a.n;
    const int32_t* G = a.G + (size_t)b * a.ldg;
    const int64_t c = a.cen[b];
    const uint64_t base0 = (uint64_t)s * a.seg;
    const uint64_t end = min(base0 + a.seg, (uint64_t)n + 1);
    int64_t pcarry = 0;
    for (uint32_t k = 0; k < s; k++) pcarry += a.gseg[(size_t)b * a.nseg + k];
    // 0 <= P <= sum G and 0 <= c i <= c n: when sum G and c n are both below
    // 2^29, |ph| < 2^29 and a run's sum of 16 squares (each < 2^58) is exact
    // in int64
    const bool fast = a.sumg[b] < (1ll << 29) && c * (int64_t)n < (1ll << 29);
    TermsG carry{};
    auto term = [&](uint32_t i, int64_t P) {
        TermsG t{};
        t.pp = P;
        const double ph = (double)(P - c * (int64_t)i);
        if (i >= 1 && i < n) t.e1 = dd_from_prod(ph, ph);
        return t;
    };
    int32_t pg[EK];
    auto fetch = [&](uint64_t base) {
#pragma unroll
        for (uint32_t m = 0; m < EK; m++) {
            const uint64_t i = base + threadIdx.x + m * NT;
            pg[m] = i < n ? G[i] : 0;
        }
    };
    for (uint64_t base = base0; base < end; base += CH) {
        fetch(base);
        __syncthreads();  // the previous chunk's LDS reads are done
#pragma unroll
        for (uint32_t m = 0; m < EK; m++) {
            const uint32_t j = threadIdx.x + m * NT;
            sG[j + j / EK] = pg[m];
        }
        __syncthreads();
        const uint32_t r0 = threadIdx.x * (EK + 1);
        const uint64_t i0 = base + (uint64_t)threadIdx.x * EK;
        int64_t gsum = 0;
#pragma unroll
        for (uint32_t k = 0; k < EK; k++) gsum += (int64_t)sG[r0 + k];  // 0 past n
        int64_t gtot;
        const int64_t gex = block_incl_scan<int64_t>(gsum, ldsi, gtot) - gsum;
        TermsG loc{};
        int64_t P = pcarry + gex;
        if (fast) {
            // run totals in int64: ph = P - c i steps by G - c
            int64_t spp = 0, se1 = 0, ph = P - c * (int64_t)i0;
#pragma unroll 8
            for (uint32_t k = 0; k < EK; k++) {
                const uint64_t i = i0 + k;
                if (i >= end) break;
                spp += P;
                if (i >= 1 && i < n) se1 += ph * ph;
                const int64_t g = sG[r0 + k];
                P += g;
                ph += g - c;
            }
            loc.pp = spp;
            loc.e1 = dd_from_i128(se1);
        } else {
            for (uint32_t k = 0; k < EK; k++) {
                const uint64_t i = i0 + k;
                if (i >= end) break;
                loc = tadd(loc, term((uint32_t)i, P));
                P += (int64_t)sG[r0 + k];
            }
        }
        TermsG tot;
        TermsG run = tadd(block_excl_scan_t(loc, lds, tot), carry);
        if (i0 < end && a.runflag[i0 / EK]) {
            P = pcarry + gex;
            const size_t rb = (size_t)b * a.nl;
            for (uint32_t k = 0; k < EK; k++) {
                const uint64_t i = i0 + k;
                if (i >= end) break;
                const int32_t jl = a.pmapL[i], jm = a.pmapM[i];
                if (jl >= 0) {
                    a.sPP_L[rb + jl] = run.pp;
                    a.sE1[rb + jl] = run.e1;
                }
                if (jm >= 0) {
                    a.sPP_M[rb + jm] = run.pp;
                    a.sE1M[rb + jm] = run.e1;
                }
                run = tadd(run, term((uint32_t)i, P));
                P += (int64_t)sG[r0 + k];
            }
        }
        carry = tadd(carry, tot);
        pcarry += gtot;
    }
    if (threadIdx.x == 0) a.totg[(size_t)b * a.nseg + s] = carry;
}

// The A side, one block per (segment s, row b), chunks as in eval_seg_g (4
// blocks per CU).
// Pass 0 rounds A to integers when row_rounds holds and flags a segment
// that holds a value further than tol from its integer (not seen in
// practice); pass 1 scans, unrounded, the rows eval_fix marked for redo.
__global__ __launch_bounds__(NT) void eval_seg_a(EvalArgs a, uint32_t pass) {
    __shared__ TermsA lds[NT / 64 + 1];
    __shared__ double sA[CH + CH / EK];
    __shared__ uint32_t sh_dev;
    const uint32_t s = blockIdx.x, b = blockIdx.y, n = a.n;
    if (pass && !a.redo[b]) return;
    const double* A = a.A + (size_t)b * n;
    const bool exact = !pass && row_rounds(a, b);
    // rounded, every |A[i]| <= A[0] (an autocorrelation peaks at lag 0); below
    // 2^45 a run's sums of A and of A k (k < 16) stay below 2^53, exact in FP64
    const bool fast = exact && rint(A[0]) < 0x1p45;
    const double tol = a.round_mode == 2 ? 0.0 : 0.25;
    const uint64_t base0 = (uint64_t)s * a.seg;
    const uint64_t end = min(base0 + a.seg, (uint64_t)n + 1);
    TermsA carry{};
    double dev = 0.0;
    if (exact && s == 0) dev = fabs(A[0] - rint(A[0]));
    // A[i] is taken as 0 outside 1 <= i < n, where C1 and C2 gain nothing
    auto term = [&](uint32_t i, double av) { return TermsA{{av, 0.0}, dd_from_prod(av, (double)i)}; };
    double pa[EK];
    auto fetch = [&](uint64_t base) {
#pragma unroll
        for (uint32_t m = 0; m < EK; m++) {
            const uint64_t i = base + threadIdx.x + m * NT;
            pa[m] = (i >= 1 && i < n) ? A[i] : 0.0;
        }
    };
    for (uint64_t base = base0; base < end; base += CH) {
        fetch(base);
        __syncthreads();  // the previous chunk's LDS reads are done
#pragma unroll
        for (uint32_t m = 0; m < EK; m++) {
            const uint32_t j = threadIdx.x + m * NT;
            double av = pa[m];
            if (exact) {
                const double ar = rint(av);
                dev = fmax(dev, fabs(av - ar));
                av = ar;
            }
            sA[j + j / EK] = av;
        }
        __syncthreads();
        const uint32_t r0 = threadIdx.x * (EK + 1);
        const uint64_t i0 = base + (uint64_t)threadIdx.x * EK;
        TermsA loc{};
        if (fast) {
            // run totals C1 = sum A, C2 = i0 sum A + sum A k, the sums in FP64
            double s1 = 0.0, s2 = 0.0;
#pragma unroll 8
            for (uint32_t k = 0; k < EK; k++) {  // 0 past n; chunks end at the segment's end
                const double av = sA[r0 + k];
                s1 += av;
                s2 = fma(av, (double)k, s2);
            }
            loc.c1 = {s1, 0.0};
            loc.c2 = dd_add(dd_from_prod((double)i0, s1), dd{s2, 0.0});
        } else {
            for (uint32_t k = 0; k < EK; k++) {
                const uint64_t i = i0 + k;
                if (i >= end) break;
                loc = tadd(loc, term((uint32_t)i, sA[r0 + k]));
            }
        }
        TermsA tot;
        TermsA run = tadd(block_excl_scan_t(loc, lds, tot), carry);
        if (i0 < end && (a.runflag[i0 / EK] & 1)) {
            const size_t rb = (size_t)b * a.nl;
            for (uint32_t k = 0; k < EK; k++) {
                const uint64_t i = i0 + k;
                if (i >= end) break;
                const int32_t jl = a.pmapL[i];
                if (jl >= 0) {
                    a.sC1[rb + jl] = run.c1;
                    a.sC2[rb + jl] = run.c2;
                }
                run = tadd(run, term((uint32_t)i, sA[r0 + k]));
            }
        }
        carry = tadd(carry, tot);
    }
    if (threadIdx.x == 0) sh_dev = 0;
    __syncthreads();
    if (dev > tol) sh_dev = 1;
    __syncthreads();
    if (threadIdx.x == 0) {
        const size_t k = (size_t)b * a.nseg + s;
        a.tota[k] = carry;
        a.segdev[k] = exact && sh_dev ? 1u : 0u;
    }
}

// One block per row, after eval_seg_g and the same pass of eval_seg_a: the
// exclusive prefixes of the segments' totals, the samples made row-global,
// then the table entries.  In pass 0 a row with a flagged segment is marked
// for redo and left for pass 1; pass 1 handles only the marked rows.
__global__ __launch_bounds__(NT) void eval_fix(EvalArgs a, uint32_t pass) {
    __shared__ TermsG preg[MAXSEG + 1];
    __shared__ TermsA prea[MAXSEG + 1];
    __shared__ uint32_t sh_redo;
    const uint32_t b = blockIdx.x, n = a.n;
    if (pass && !a.redo[b]) return;
    const bool exact = !pass && row_rounds(a, b);
    const size_t sb = (size_t)b * a.nseg;
    if (threadIdx.x == 0) {
        TermsG rg{};
        TermsA ra{};
        uint32_t bad = 0;
        for (uint32_t k = 0; k < a.nseg; k++) {
            preg[k] = rg;
            prea[k] = ra;
            rg = tadd(rg, a.totg[sb + k]);
            ra = tadd(ra, a.tota[sb + k]);
            bad |= a.segdev[sb + k];
        }
        preg[a.nseg] = rg;
        prea[a.nseg] = ra;
        sh_redo = bad;
        a.redo[b] = bad;
        a.exact[b] = exact && !bad ? 1u : 0u;
    }
    __syncthreads();
    if (sh_redo) return;
    // totals over positions 0..n: PP[n+1] = PP[n] + P[n] (P[n] = sum G) and
    // E1[n+1] = E1[n]
    const double A0r = a.A[(size_t)b * n];
    const double A0 = exact ? rint(A0r) : A0r;
    const int64_t c = a.cen[b];
    const int64_t PHn = a.sumg[b] - c * (int64_t)n;
    const size_t rb = (size_t)b * a.nl;
    const int64_t PPend = preg[a.nseg].pp;
    const __int128 PPn = (__int128)PPend - a.sumg[b];
    const dd E1n = preg[a.nseg].e1;
    for (uint32_t j = threadIdx.x; j < a.nl; j += NT) {
        const uint64_t L = a.ls[j];
        const uint64_t nw = (uint64_t)n - L + 1;
        const TermsG& gl = preg[L / a.seg];
        const TermsG& gm = preg[nw / a.seg];
        const TermsA& al = prea[L / a.seg];
        const int64_t ppL = a.sPP_L[rb + j] + gl.pp, ppM = a.sPP_M[rb + j] + gm.pp;
        const dd C1 = dd_add(a.sC1[rb + j], al.c1), C2 = dd_add(a.sC2[rb + j], al.c2);
        const dd E1 = dd_add(a.sE1[rb + j], gl.e1), E1M = dd_add(a.sE1M[rb + j], gm.e1);
        const __int128 S1 = (__int128)PPend - ppL - ppM;
        const __int128 S1H = S1 - (__int128)c * (__int128)L * (__int128)nw;
        // T = L A0 + 2 (L C1 - C2)
        dd T = dd_add(dd_from_prod(A0, (double)L), dd_mul_d(dd_add(dd_mul_d(C1, (double)L), dd_neg(C2)), 2.0));
        // E = E1[L] + Q[n] - Q[M], M = nw: Q[n] - Q[M] = (n - M) PHn^2
        //   - 2 PHn (PP[n] - PP[M] - c (n (n - 1) - M (M - 1)) / 2) + E1[n] - E1[M]
        const __int128 dtri = ((__int128)n * (n - 1) - (__int128)nw * (nw - 1)) / 2;
        const __int128 K = (__int128)(n - nw) * PHn * PHn - 2 * (__int128)PHn * ((PPn - ppM) - (__int128)c * dtri);
        dd E = dd_add(dd_add(E1, dd_add(E1n, dd_neg(E1M))), dd_from_i128(K));
        dd S2 = dd_add(T, dd_neg(E));
        dd s1h = dd_from_i128(S1H);
        dd num = dd_add(dd_mul_d(S2, (double)nw), dd_neg(dd_mul(s1h, s1h)));
        const double nwd = (double)nw;
        double v = dd_div_to_d(num, 1e4 * nwd * nwd);
        const int64_t o = a.out_off[b] + j;
        a.mean[o] = (double)S1 / (100.0 * nwd);
        a.var[o] = v > 0.0 ? v : 0.0;
    }
}

inline uint32_t ilog2(uint64_t v) {
    uint32_t l = 0;
    while ((1ull << l) < v) l++;
    return l;
}

}  // namespace

void moments_build(Index& ix) {
    Moments& M = ix.mom;
    CHK_REQUIRE(M.genome, CHK_ERR_INVALID, "set the genome table before building moments");
    const uint32_t S = ix.nshards, F = ix.nfiles;
    hipStream_t st = ix.stream;
    // slots: (file, shard) with intervals on a genome shard
    std::vector<int64_t> slot_off((size_t)F * S, -1);
    std::vector<uint32_t> slot_len((size_t)F * S, 0);
    std::vector<std::vector<uint32_t>> lists(S);
    uint64_t total = 0;
    for (uint32_t s = 0; s < S; s++) compact_ls(M.nbins[s], &lists[s]);
    for (uint32_t f = 0; f < F; f++)
        for (uint32_t s = 0; s < S; s++) {
            if (M.nbins[s] == 0 || ix.fs_count[(size_t)f * S + s] == 0) continue;
            slot_off[(size_t)f * S + s] = (int64_t)total;
            slot_len[(size_t)f * S + s] = (uint32_t)lists[s].size();
            total += lists[s].size();
        }
    M.mean.reserve(total);
    M.var.reserve(total);
    // twiddles exp(-2 pi i t / TW), t < TW, from the host C library in FP64
    std::vector<double2> twh(TW), twnh(FFT_MAXM);
    for (uint32_t t = 0; t < TW; t++) {
        const double ang = -2.0 * M_PI * (double)t / (double)TW;
        twh[t] = make_double2(std::cos(ang), std::sin(ang));
    }
    DevBuf<double2> tw, twN;
    tw.reserve(TW);
    twN.reserve(FFT_MAXM);
    CHK_HIP(hipMemcpyAsync(tw.p, twh.data(), TW * sizeof(double2), hipMemcpyHostToDevice, st));
    DevBuf<int32_t> slot_of_fid, pmapL, pmapM, D;
    DevBuf<uint8_t> runflag;
    DevBuf<int64_t> cen, sumg, out_off, sPP_L, sPP_M;
    DevBuf<dd> sC1, sC2, sE1, sE1M, sumh2;
    DevBuf<uint32_t> d_ls, d_exact, segdev, redo;
    DevBuf<int64_t> gseg;
    DevBuf<TermsG> totg;
    DevBuf<TermsA> tota;
    DevBuf<double2> X;
    DevBuf<double> A;
    slot_of_fid.reserve(F);
    std::vector<int32_t> sof(F, -1);
    int nexact = 0, nrows = 0;
    // FFT blocks hold up to maxm complex FP64 values in LDS, one radix-8 unit
    // per thread: 8192 (128 KiB, one block per CU) where the device allows
    // that much LDS per block, else 4096.  Row blocks (a row and its mirror,
    // 2 * N1 values) hold mrow values: 4096 by default, so that two blocks
    // share a CU and one's loads overlap the other's passes; column blocks use
    // maxm, which gives them more adjacent columns per block.
    int max_lds = 0;
    CHK_HIP(hipDeviceGetAttribute(&max_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, ix.device));
    const uint32_t maxm = max_lds >= (int)(FFT_MAXM * sizeof(double2)) + 8192 ? FFT_MAXM : FFT_MAXM / 2;
    const uint32_t mrow = ix.tune.fft_block == FFT_MAXM ? maxm : FFT_MAXM / 2;
    const int lds_need = (int)(maxm * sizeof(double2));
    if (lds_need > 64 * 1024) {
        CHK_HIP(hipFuncSetAttribute((const void*)autocorr_small, hipFuncAttributeMaxDynamicSharedMemorySize, lds_need));
        CHK_HIP(hipFuncSetAttribute((const void*)fft_cols_fwd, hipFuncAttributeMaxDynamicSharedMemorySize, lds_need));
        CHK_HIP(hipFuncSetAttribute((const void*)fft_cols_fwd_p<1>, hipFuncAttributeMaxDynamicSharedMemorySize, lds_need));
        CHK_HIP(hipFuncSetAttribute((const void*)fft_cols_fwd_p<2>, hipFuncAttributeMaxDynamicSharedMemorySize, lds_need));
        CHK_HIP(hipFuncSetAttribute((const void*)fft_cols_fwd_p<3>, hipFuncAttributeMaxDynamicSharedMemorySize, lds_need));
        CHK_HIP(hipFuncSetAttribute((const void*)fft_rows, hipFuncAttributeMaxDynamicSharedMemorySize, lds_need));
        CHK_HIP(hipFuncSetAttribute((const void*)fft_cols_inv, hipFuncAttributeMaxDynamicSharedMemorySize, lds_need));
    }
    for (uint32_t s = 0; s < S; s++) {
        const uint32_t n = M.nbins[s];
        if (n < 2) continue;  // empty tables (n <= 1 stores nothing)
        std::vector<uint32_t> files;
        for (uint32_t f = 0; f < F; f++)
            if (ix.fs_count[(size_t)f * S + s]) files.push_back(f);
        if (files.empty()) continue;
        const auto& ls = lists[s];
        const uint32_t nl = (uint32_t)ls.size();
        const uint64_t N = 1ull << ilog2(2ull * n - 1);
        CHK_REQUIRE(N <= (uint64_t)mrow * maxm / 2, CHK_ERR_INVALID,
                    "shard too long for the moment FFT");
        const uint32_t logN = ilog2(N);
        if (N > maxm) {  // four-step twiddles exp(-2 pi i t / N), t < FFT_MAXM
            CHK_HIP(hipStreamSynchronize(st));  // the previous shard's table is no longer read
            for (uint32_t t = 0; t < FFT_MAXM; t++) {
                const double ang = -2.0 * M_PI * (double)t / (double)N;
                twnh[t] = make_double2(std::cos(ang), std::sin(ang));
            }
            CHK_HIP(hipMemcpyAsync(twN.p, twnh.data(), FFT_MAXM * sizeof(double2), hipMemcpyHostToDevice, st));
        }
        // sample maps: position L -> index of L; position n-L+1 -> index of L
        // and the runs of EK positions that hold either
        std::vector<int32_t> pl(n + 2, -1), pm(n + 2, -1);
        const size_t nrun = (size_t)n / EK + 1;
        std::vector<uint8_t> rf(nrun, 0);
        for (uint32_t j = 0; j < nl; j++) {
            pl[ls[j]] = (int32_t)j;
            pm[n - ls[j] + 1] = (int32_t)j;
            rf[ls[j] / EK] |= 1;
            rf[(n - ls[j] + 1) / EK] |= 2;
        }
        pmapL.reserve(n + 2);
        pmapM.reserve(n + 2);
        runflag.reserve(nrun);
        d_ls.reserve(nl);
        CHK_HIP(hipMemcpyAsync(pmapL.p, pl.data(), (n + 2) * 4, hipMemcpyHostToDevice, st));
        CHK_HIP(hipMemcpyAsync(pmapM.p, pm.data(), (n + 2) * 4, hipMemcpyHostToDevice, st));
        CHK_HIP(hipMemcpyAsync(runflag.p, rf.data(), nrun, hipMemcpyHostToDevice, st));
        CHK_HIP(hipMemcpyAsync(d_ls.p, ls.data(), nl * 4, hipMemcpyHostToDevice, st));
        // rows per batch: as many as a third of the free HBM holds (FFT
        // workspace, autocorrelation, coverage), so long chromosomes still
        // give every CU rows to work on
        size_t free_b = 0, total_b = 0;
        CHK_HIP(hipMemGetInfo(&free_b, &total_b));
        // evaluation segments: seg positions each (a multiple of CH), at
        // most MAXSEG per row
        const uint32_t seg_min = (uint32_t)(((uint64_t)n + 1 + MAXSEG - 1) / MAXSEG);
        const uint32_t seg = std::max<uint32_t>(ix.tune.eval_seg ? ix.tune.eval_seg : EVAL_SEG,
                                                (seg_min + CH - 1) / CH * CH);
        const uint32_t nseg = n / seg + 1;  // positions 0..n
        const uint64_t row_bytes = (uint64_t)n * 12 + 64 * (uint64_t)nl +
                                   (uint64_t)nseg * (sizeof(TermsG) + sizeof(TermsA) + 12) +
                                   (N <= maxm ? 0 : N * sizeof(double2) / 2);
        const uint32_t B = (uint32_t)std::max<uint64_t>(
            1, std::min<uint64_t>({files.size(), (uint64_t)(free_b / 3) / row_bytes, 4096}));
        const uint32_t ldg = n + 1;
        // four-step FFTs run over groups of signal pairs whose workspace fits
        // the budget (by default well inside the 256 MiB last-level cache, so
        // the three passes re-read it from there rather than from HBM)
        const uint64_t fft_budget = (uint64_t)(ix.tune.fft_mib ? ix.tune.fft_mib : 128) << 20;
        const uint32_t PG = (uint32_t)std::max<uint64_t>(
            1, std::min<uint64_t>((B + 1) / 2, fft_budget / (N * sizeof(double2))));
        D.reserve((size_t)B * ldg);
        if (N > maxm) X.reserve((size_t)PG * N);
        A.reserve((size_t)B * n);
        cen.reserve(B);
        sumg.reserve(B);
        sumh2.reserve(B);
        out_off.reserve(B);
        d_exact.reserve(B);
        sPP_L.reserve((size_t)B * nl);
        sPP_M.reserve((size_t)B * nl);
        sC1.reserve((size_t)B * nl);
        sC2.reserve((size_t)B * nl);
        sE1.reserve((size_t)B * nl);
        sE1M.reserve((size_t)B * nl);
        gseg.reserve((size_t)B * nseg);
        totg.reserve((size_t)B * nseg);
        tota.reserve((size_t)B * nseg);
        segdev.reserve((size_t)B * nseg);
        redo.reserve(B);
        const size_t NSUB = (size_t)ix.ngroups * ix.nlayers;  // segments of one shard
        const LayerSeg& Ll = ix.lseg[(size_t)s * NSUB + NSUB - 1];
        const uint64_t soff = ix.lseg[(size_t)s * NSUB].off, scnt = Ll.off + Ll.n - soff;
        for (size_t f0 = 0; f0 < files.size(); f0 += B) {
            const uint32_t nb = (uint32_t)std::min<size_t>(B, files.size() - f0);
            std::vector<int64_t> oo(nb);
            for (uint32_t k = 0; k < nb; k++) {
                sof[files[f0 + k]] = (int32_t)k;
                oo[k] = slot_off[(size_t)files[f0 + k] * S + s];
            }
            CHK_HIP(hipMemcpyAsync(slot_of_fid.p, sof.data(), F * 4, hipMemcpyHostToDevice, st));
            CHK_HIP(hipMemcpyAsync(out_off.p, oo.data(), nb * 8, hipMemcpyHostToDevice, st));
            CHK_HIP(hipMemsetAsync(D.p, 0, (size_t)nb * ldg * 4, st));
            const unsigned grid = (unsigned)std::min<uint64_t>((scnt + NT - 1) / NT, (uint64_t)ix.ncu * 16);
            if (grid)
                hipLaunchKernelGGL(impulses, dim3(grid), dim3(NT), 0, st, ix.l_start.p + soff,
                                   ix.l_end.p + soff, ix.l_fid.p + soff, scnt, slot_of_fid.p, n,
                                   ldg, D.p);
            hipLaunchKernelGGL(coverage_rows, dim3(nb), dim3(NT), 0, st, D.p, ldg, n, cen.p, sumg.p,
                               sumh2.p, seg, nseg, gseg.p);
            CHK_HIP(hipGetLastError());
            Rows r{D.p, ldg, cen.p, X.p, A.p, nb, n, (uint32_t)N, logN, tw.p, twN.p};
            const uint32_t np = (nb + 1) / 2;  // signal pairs
            hipEvent_t t0 = ix.prof.begin("moments_fft", st);
            if (N <= maxm) {
                hipLaunchKernelGGL(autocorr_small, dim3(np), dim3(std::max<uint32_t>(64, (uint32_t)N / 8)),
                                   N * sizeof(double2), st, r);
            } else {
                // rows of N1 = mrow / 2 (two per block), columns of N2 = N / N1 (CB per block)
                const uint32_t N1 = mrow / 2, N2 = (uint32_t)(N / N1);
                const uint32_t CB = std::min<uint32_t>(N1, maxm / N2);
                const size_t lds_col = (size_t)CB * N2 * sizeof(double2);
                const uint32_t tcol = CB * N2 / 8, trow = N1 / 4;
                for (uint32_t p0 = 0; p0 < np; p0 += PG) {
                    const uint32_t pn = std::min(PG, np - p0);
                    Rows rg = r;  // pairs p0 .. p0 + pn - 1, workspace from X[0]
                    rg.G = D.p + (size_t)2 * p0 * ldg;
                    rg.cen = cen.p + 2 * p0;
                    rg.A = A.p + (size_t)2 * p0 * n;
                    rg.nb = std::min(2 * pn, nb - 2 * p0);
                    const uint32_t ngrp = pn * (N1 / CB), logN2 = ilog2(N2);
                    if (ix.tune.fft_pipe == 0 && logN2 >= 4 && CB * N2 == 8 * tcol) {
                        // persistent blocks, one per CU (a multiple of 8 blocks
                        // when the groups fill the device)
                        const uint32_t pg = std::min<uint32_t>(ngrp, (uint32_t)ix.ncu);
                        const uint32_t lr = logN2 % 3 ? logN2 % 3 : 3;
                        auto k = lr == 1 ? fft_cols_fwd_p<1> : (lr == 2 ? fft_cols_fwd_p<2> : fft_cols_fwd_p<3>);
                        hipLaunchKernelGGL(k, dim3(pg), dim3(tcol), lds_col, st, rg, N1, logN2, ilog2(CB), ngrp);
                    } else {
                        hipLaunchKernelGGL(fft_cols_fwd, dim3(ngrp), dim3(tcol), lds_col, st, rg, N1, logN2,
                                           ilog2(CB));
                    }
                    hipLaunchKernelGGL(fft_rows, dim3(pn * (N2 / 2 + 1)), dim3(trow),
                                       2 * N1 * sizeof(double2), st, rg, N1, ilog2(N1), N2);
                    hipLaunchKernelGGL(fft_cols_inv, dim3(pn * (N1 / CB)), dim3(tcol), lds_col, st, rg, N1,
                                       ilog2(N2), ilog2(CB));
                }
            }
            CHK_HIP(hipGetLastError());
            ix.prof.end("moments_fft", t0, st);
            EvalArgs ea{D.p, ldg, n, nl, A.p, cen.p, sumg.p, pmapL.p, pmapM.p, d_ls.p, out_off.p,
                        runflag.p, M.mean.p, M.var.p, sPP_L.p, sPP_M.p, sC1.p, sC2.p, sE1.p, sE1M.p, d_exact.p,
                        sumh2.p, ix.tune.mom_round, seg, nseg, gseg.p, totg.p, tota.p, segdev.p, redo.p};
            t0 = ix.prof.begin("moments_eval", st);
            // the G side once; the A side in pass 0 over every row, in pass 1
            // again, unrounded, for the rows whose rounded scan was not exact
            // (its blocks return at once for the others)
            hipLaunchKernelGGL(eval_seg_g, dim3(nseg, nb), dim3(NT), 0, st, ea);
            for (uint32_t pass = 0; pass < 2; pass++) {
                hipLaunchKernelGGL(eval_seg_a, dim3(nseg, nb), dim3(NT), 0, st, ea, pass);
                hipLaunchKernelGGL(eval_fix, dim3(nb), dim3(NT), 0, st, ea, pass);
            }
            CHK_HIP(hipGetLastError());
            ix.prof.end("moments_eval", t0, st);
            std::vector<uint32_t> ex(nb);
            CHK_HIP(hipMemcpyAsync(ex.data(), d_exact.p, nb * 4, hipMemcpyDeviceToHost, st));
            CHK_HIP(hipStreamSynchronize(st));
            for (uint32_t k = 0; k < nb; k++) {
                nexact += ex[k] ? 1 : 0;
                sof[files[f0 + k]] = -1;
            }
            nrows += nb;
        }
    }
    M.stat_exact_rows = nexact;
    M.stat_rows = nrows;
    // install slot tables (entries written above)
    const size_t ns = (size_t)F * S;
    M.slot_off = slot_off;
    M.slot_len = slot_len;
    std::vector<uint8_t> has(F, 0);
    for (size_t k = 0; k < ns; k++)
        if (slot_off[k] >= 0) has[k / S] = 1;
    M.total = total;
    M.d_slot_off.reserve(ns);
    M.d_slot_len.reserve(ns);
    M.d_has_entry.reserve(F);
    CHK_HIP(hipMemcpyAsync(M.d_slot_off.p, slot_off.data(), ns * 8, hipMemcpyHostToDevice, st));
    CHK_HIP(hipMemcpyAsync(M.d_slot_len.p, slot_len.data(), ns * 4, hipMemcpyHostToDevice, st));
    CHK_HIP(hipMemcpyAsync(M.d_has_entry.p, has.data(), F, hipMemcpyHostToDevice, st));
    CHK_HIP(hipStreamSynchronize(st));
    M.present = true;
}

void moments_download(Index& ix, int64_t* slot_off, uint32_t* slot_len, double* mean, double* var) {
    Moments& M = ix.mom;
    CHK_REQUIRE(M.present, CHK_ERR_NOT_FOUND, "the index has no moment tables");
    const size_t ns = (size_t)ix.nfiles * ix.nshards;
    if (slot_off) std::copy(M.slot_off.begin(), M.slot_off.end(), slot_off);
    if (slot_len) std::copy(M.slot_len.begin(), M.slot_len.end(), slot_len);
    (void)ns;
    if (mean && M.total)
        CHK_HIP(hipMemcpy(mean, M.mean.p, M.total * 8, hipMemcpyDeviceToHost));
    if (var && M.total)
        CHK_HIP(hipMemcpy(var, M.var.p, M.total * 8, hipMemcpyDeviceToHost));
}

// ------------------------------------------------------------ FP32 lag-limited route
//
// The C3 precision form: moments at a few block sizes L <= LB_LAGS (the
// configurations' 16 widths of 1..1000 bins) need only the lags m < L of the
// autocorrelation, so A[m] for m < LB_LAGS comes from blocked FP32 FFTs
// instead of one FP64 transform of the whole chromosome:
//   A[m] = sum_b sum_{t < LB_HOP} H[b HOP + t] H[b HOP + t + m],
// each block b one complex FFT of LB_N = 4096 points of z = x + i y with
// x = H[b HOP .. +HOP) (zero beyond HOP) and y = H[b HOP .. +N), since
// c[m] = sum_t x[t] y[t + m] never wraps for m < N - HOP.  Two blocks share
// each inverse FFT (conj(X1) Y1 + i conj(X2) Y2 gives c1 + i c2).  The FFTs
// are three Stockham radix-16 passes through LDS (16 points per thread,
// natural order in and out); a block's lags are added into FP32 two-sum
// (compensated) accumulators, the work group's partial sums are combined in
// the same two-sum arithmetic in a fixed order, and only the variance step
// runs in FP64:
//   S2_full(L) = L A[0] + 2 sum_{1<=d<L} (L - d) A[d]   (zero-padded H)
//   S2_H(L)    = S2_full - sum_{1<=m<L} PH[m]^2 - sum_{1<=k<L} SUF[k]^2
//   S1_H(L)    = sum_{j=n-L+1..n} PH[j] - sum_{j<L} PH[j]
// with PH the prefix sums of H and SUF[k] the sum of its last k values, both
// exact integers over the first and last L bins; mean = (S1_H / n_w + c L) /
// 100, var = (S2_H / n_w - (S1_H / n_w)^2) / 1e4, n_w = n - L + 1.
namespace {

constexpr uint32_t LB_N = 4096, LB_HOP = 3072, LB_LAGS = LB_N - LB_HOP;
constexpr uint32_t LB_PAIRS = 8;  // block pairs per work group (49,152 bins)

__device__ inline float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ inline float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ inline float2 cmul(float2 a, float2 b) {
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ inline float2 cconj(float2 a) { return make_float2(a.x, -a.y); }

// forward 4-point DFT in place
__device__ inline void dft4(float2& a0, float2& a1, float2& a2, float2& a3) {
    const float2 s02 = cadd(a0, a2), d02 = csub(a0, a2), s13 = cadd(a1, a3), d13 = csub(a1, a3);
    a0 = cadd(s02, s13);
    a2 = csub(s02, s13);
    a1 = make_float2(d02.x + d13.y, d02.y - d13.x);  // d02 - i d13
    a3 = make_float2(d02.x - d13.y, d02.y + d13.x);  // d02 + i d13
}

// Forward 8-point DFT in place, as 2 x 4 (n = 4 n1 + n2) with the internal
// twiddles exp(-2 pi i n2 k1 / 8): input v[n] in natural order, output
// V[k1 + 2 k2] in register 4 k1 + k2 (digit-transposed; the callers' store
// addresses undo it, so no register moves are needed).
__device__ inline void dft8(float2 (&v)[8]) {
    constexpr float R2 = 0.70710678118654752f;
#pragma unroll
    for (int n2 = 0; n2 < 4; n2++) {
        const float2 x = v[n2], y = v[4 + n2];
        v[n2] = cadd(x, y);
        v[4 + n2] = csub(x, y);
    }
    v[5] = cmul(v[5], make_float2(R2, -R2));   // W8^1
    v[6] = make_float2(v[6].y, -v[6].x);       // W8^2 = -i
    v[7] = cmul(v[7], make_float2(-R2, -R2));  // W8^3
    dft4(v[0], v[1], v[2], v[3]);
    dft4(v[4], v[5], v[6], v[7]);
}

// frequency held by register r after dft8: k1 + 2 k2 with r = 4 k1 + k2
__device__ constexpr int dig(int r) { return (r >> 2) + 2 * (r & 3); }

constexpr int LB_T = 512;  // threads of a lag FFT work group (8 points each)

#ifndef LB_PADSHIFT
#define LB_PADSHIFT 5
#endif
// one padding slot after every 2^LB_PADSHIFT values (1 in 32 measured best:
// 101.6 ms for the C3 FFTs against 109 ms at 1 in 8, 113 ms at 1 in 64, 142 ms
// with none)
__device__ inline uint32_t lpad8(uint32_t i) { return i + (i >> LB_PADSHIFT); }

// Twiddle tables of the three twiddled passes, laid out so that a wave reads
// consecutive entries: pass S (8, 64, 512) holds W^(l k 4096 / (8 S)) at
// tw_off(S) + (k - 1) S + l for l < S, k = 1..7 (4,088 entries in all).
__device__ constexpr uint32_t tw_off(uint32_t S) { return S == 8 ? 0u : (S == 64 ? 56u : 504u); }
constexpr uint32_t LB_TW = 7 * (8 + 64 + 512);

// One radix-8 Stockham pass of stride S over the LDS (S > 1): read the
// thread's 8 inputs in[j + 512 k], twiddle, DFT, and write
// out[(j / S) 8 S + j % S + S k].
template <uint32_t S>
__device__ inline void pass8(float2 (&v)[8], float2* lds, const float2* T, bool write) {
    const uint32_t j = threadIdx.x;
#pragma unroll
    for (int k = 0; k < 8; k++) v[k] = lds[lpad8(j + LB_T * k)];
    __syncthreads();
#pragma unroll
    for (int k = 1; k < 8; k++) v[k] = cmul(v[k], T[tw_off(S) + (k - 1) * S + (j % S)]);
    dft8(v);
    if (write) {
        const uint32_t base = (j / S) * 8 * S + (j % S);
#pragma unroll
        for (int r = 0; r < 8; r++) lds[lpad8(base + S * dig(r))] = v[r];
        __syncthreads();
    }
}

// Forward 4096-point FFT in place (four radix-8 Stockham passes): on entry
// thread j holds v[k] = z[j + 512 k]; on exit register r holds
// Z[j + 512 dig(r)] and, when `publish`, the LDS holds Z in natural order.
// T = the twiddled passes' tables (tw_off layout).
__device__ inline void fft4096(float2 (&v)[8], float2* lds, const float2* T, bool publish) {
    const uint32_t j = threadIdx.x;
    dft8(v);  // pass 0: stride 1, no twiddles
#pragma unroll
    for (int r = 0; r < 8; r++) lds[lpad8(j * 8 + dig(r))] = v[r];
    __syncthreads();
    pass8<8>(v, lds, T, true);
    pass8<64>(v, lds, T, true);
    pass8<512>(v, lds, T, publish);  // output index j + 512 k: natural order
}

// FP32 two-sum: s + v = t + e exactly; the error goes into c
__device__ inline void two_sum_add(float& s, float& c, float v) {
    const float t = s + v;
    const float bp = t - s;
    const float e = (s - (t - bp)) + (v - bp);
    s = t;
    c += e;
}

struct LagArgs {
    const int32_t* G;       // coverage rows, row stride ldg
    uint32_t ldg, n;
    const int64_t* cen;     // centring constant per row
    const float2* T;        // the passes' twiddle tables (tw_off layout)
    uint32_t npairs, nchunk;
    float* part;            // [row][chunk][LB_LAGS][2]: two-sum (sum, compensation)
    uint32_t* inexact;      // per row: set when some |H| >= 2^24 (not exact in FP32)
};

// Block b's autocorrelation spectrum conj(X) Y at the thread's 8
// frequencies k = j + 512 q, from Z = FFT(x + i y) (own values in the
// registers, partners N - k in the LDS).  W[q] = conj(X) Y for the first
// block of a pair; the second adds i conj(X) Y.
template <bool SECOND>
__device__ inline void lag_spectrum(const float2 (&v)[8], const float2* lds, float2 (&W)[8]) {
    const uint32_t j = threadIdx.x;
#pragma unroll
    for (int r = 0; r < 8; r++) {
        const int q = dig(r);
        const uint32_t k = j + LB_T * q;
        const float2 a = v[r];
        const float2 bb = cconj(lds[lpad8((LB_N - k) & (LB_N - 1))]);
        // conj(X) Y = conj(a + bb) (a - bb) (-i) / 4
        const float2 pr = cmul(cconj(cadd(a, bb)), csub(a, bb));
        const float2 P = make_float2(0.25f * pr.y, -0.25f * pr.x);
        if (SECOND) W[q] = make_float2(W[q].x - P.y, W[q].y + P.x);  // + i P
        else W[q] = P;
    }
}

// The coverage of block blk (4096 bins from blk * HOP; past the row's end
// and past the last block: c, i.e. H = 0), loaded ahead of its FFT.
__device__ inline void lag_fetch(const LagArgs& a, const int32_t* row, int32_t c, uint32_t blk,
                                 uint32_t nblk, int32_t (&g)[8]) {
    const uint32_t j = threadIdx.x;
    const uint64_t t0 = (uint64_t)blk * LB_HOP;
#pragma unroll
    for (int k = 0; k < 8; k++) {
        const uint64_t gi = t0 + j + LB_T * k;
        g[k] = (blk < nblk && gi < a.n) ? row[gi] : c;
    }
}

// z = x + i y of a fetched block: y = H, x = H on the first HOP points
__device__ inline void lag_convert(const int32_t (&g)[8], int32_t c, float2 (&v)[8], uint32_t& bad) {
    const uint32_t j = threadIdx.x;
#pragma unroll
    for (int k = 0; k < 8; k++) {
        const int32_t hv = g[k] - c;
        bad |= (hv >= (1 << 24) || hv <= -(1 << 24)) ? 1u : 0u;
        const float h = (float)hv;
        v[k] = make_float2(j + LB_T * k < LB_HOP ? h : 0.f, h);
    }
}

__global__ __launch_bounds__(LB_T) __attribute__((amdgpu_waves_per_eu(4))) void lag_fft32(LagArgs a) {
    __shared__ float2 lds[LB_N + (LB_N >> LB_PADSHIFT)];
    __shared__ float2 tw[LB_TW];  // the passes' twiddles, read 21 times per FFT
    const uint32_t r = blockIdx.y, chunk = blockIdx.x, j = threadIdx.x;
    for (uint32_t m = j; m < LB_TW; m += LB_T) tw[m] = a.T[m];
    __syncthreads();
    const int32_t* row = a.G + (size_t)r * a.ldg;
    const int32_t c = (int32_t)a.cen[r];
    const uint32_t nblk = (a.n + LB_HOP - 1) / LB_HOP;
    float s[2] = {0.f, 0.f}, cc[2] = {0.f, 0.f};
    uint32_t bad = 0;
    const uint32_t p0 = chunk * LB_PAIRS, p1 = min(a.npairs, p0 + LB_PAIRS);
    // software pipeline: the next block's coverage is in flight during each FFT
    int32_t g[8];
    lag_fetch(a, row, c, 2 * p0, nblk, g);
    for (uint32_t p = p0; p < p1; p++) {
        float2 v[8], W[8];
        lag_convert(g, c, v, bad);
        lag_fetch(a, row, c, 2 * p + 1, nblk, g);
        fft4096(v, lds, tw, true);
        lag_spectrum<false>(v, lds, W);
        __syncthreads();  // the partner reads are done before the LDS is reused
        lag_convert(g, c, v, bad);  // a block past the last one is all zero
        lag_fetch(a, row, c, 2 * p + 2, nblk, g);
        if (2 * p + 1 < nblk) {
            fft4096(v, lds, tw, true);
            lag_spectrum<true>(v, lds, W);
            __syncthreads();
        }
        // inverse of W = P1 + i P2 as conj(FFT(conj(W))): c1 + i c2, times 4096
#pragma unroll
        for (int q = 0; q < 8; q++) v[q] = cconj(W[q]);
        fft4096(v, lds, tw, false);
        // lags j and j + 512 sit in registers 0 and 4 (dig 0 and 1):
        // c1[m] = Re / N, c2[m] = -Im / N (the conjugate)
#pragma unroll
        for (int k = 0; k < 2; k++) {
            two_sum_add(s[k], cc[k], v[4 * k].x * (1.0f / LB_N));
            two_sum_add(s[k], cc[k], -v[4 * k].y * (1.0f / LB_N));
        }
    }
    float* out = a.part + ((size_t)r * a.nchunk + chunk) * LB_LAGS * 2;
#pragma unroll
    for (int q = 0; q < 2; q++) {
        out[2 * (j + LB_T * q)] = s[q];
        out[2 * (j + LB_T * q) + 1] = cc[q];
    }
    if (bad) atomicOr(&a.inexact[r], 1u);
}

struct Lag32Eval {
    const int32_t* G;
    uint32_t ldg, n;
    const int64_t* cen;
    const int64_t* sumg;
    const float* part;
    uint32_t nchunk;
    const uint32_t* widths;
    uint32_t nw;
    const int64_t* out_slot;  // per row: first output entry
    double* mean;
    double* var;
    double* A_out;            // optional [row][LB_LAGS]: the FP64 lags (tests), or null
    uint32_t* redo;           // [row][nw]: 1 when the width is evaluated exactly instead
};

// Rows this short (or widths whose FP32 error bound is not small against
// S2_H) are evaluated by exact window sums instead (lag_exact).
constexpr uint32_t LB_EXACT_N = 8192;
// FP32 FFT error of the weighted lag sum: relative to A[0] L^1.5 it stays
// below 1.7e-7 on white-noise tracks (measured); 2.5e-7 with margin.
constexpr double LB_ERR = 2.5e-7;
constexpr double LB_TOL = 5.0e-7;

// One block per row: lags from the work groups' two-sum partials (fixed
// order), exact edge prefix sums, then mean / variance per width.
__global__ __launch_bounds__(NT) void lag_eval32(Lag32Eval a) {
    __shared__ double A[LB_LAGS];
    __shared__ int64_t PHl[LB_LAGS + 1], SUF[LB_LAGS + 1];
    const uint32_t r = blockIdx.x, j = threadIdx.x, n = a.n;
    const float* pr = a.part + (size_t)r * a.nchunk * LB_LAGS * 2;
    for (uint32_t m = j; m < LB_LAGS; m += NT) {
        float S = 0.f, Cp = 0.f;
        for (uint32_t k = 0; k < a.nchunk; k++) {
            two_sum_add(S, Cp, pr[((size_t)k * LB_LAGS + m) * 2]);
            Cp += pr[((size_t)k * LB_LAGS + m) * 2 + 1];
        }
        A[m] = (double)S + (double)Cp;
        if (a.A_out) a.A_out[(size_t)r * LB_LAGS + m] = A[m];
    }
    const int32_t* row = a.G + (size_t)r * a.ldg;
    const int64_t c = a.cen[r];
    const uint32_t E = min(LB_LAGS, n);  // edge bins needed (L <= n - 1 < E + 1)
    if (j == 0) {
        int64_t acc = 0;
        PHl[0] = 0;
        for (uint32_t m = 1; m <= E; m++) {
            acc += (int64_t)row[m - 1] - c;
            PHl[m] = acc;
        }
        acc = 0;
        SUF[0] = 0;
        for (uint32_t k = 1; k <= E; k++) {
            acc += (int64_t)row[n - k] - c;
            SUF[k] = acc;
        }
    }
    __syncthreads();
    const int64_t PHn = a.sumg[r] - c * (int64_t)n;
    for (uint32_t wi = j; wi < a.nw; wi += NT) {
        const uint32_t L = a.widths[wi];
        const int64_t o = a.out_slot[r] + wi;
        if (L == 0 || L >= n || L > LB_LAGS) {
            a.mean[o] = NAN;
            a.var[o] = NAN;
            continue;
        }
        // S1_H = L PH[n] - sum_{k<L} SUF[k] - sum_{m<L} PH[m]
        int64_t s1 = (int64_t)L * PHn;
        __int128 def = 0;
        for (uint32_t k = 0; k < L; k++) {
            s1 -= SUF[k] + PHl[k];
            def += (__int128)PHl[k] * PHl[k] + (__int128)SUF[k] * SUF[k];
        }
        double s2 = (double)L * A[0];
        for (uint32_t d = 1; d < L; d++) s2 += 2.0 * (double)(L - d) * A[d];
        s2 -= (double)def;
        const double nwin = (double)(n - L + 1);
        const double m1 = (double)s1 / nwin;
        const double v = s2 / nwin - m1 * m1;
        a.mean[o] = (m1 + (double)c * (double)L) / 100.0;
        a.var[o] = (v > 0.0 ? v : 0.0) / 1.0e4;
        // error bound of the FP32 lags in S2_H, against the variance sum it feeds
        const double err = LB_ERR * A[0] * (double)L * sqrt((double)L);
        const double vs = (v > 0.0 ? v : 0.0) * nwin;
        a.redo[(size_t)r * a.nw + wi] = (n <= LB_EXACT_N || !(err <= LB_TOL * vs)) ? 1u : 0u;
    }
}

struct LagExact {
    const int32_t* G;
    uint32_t ldg, n;
    const int64_t* cen;
    const uint32_t* widths;
    uint32_t nw;
    const uint32_t* redo;     // [row][nw]
    const uint32_t* rows;     // rows with at least one width to redo
    const int64_t* out_slot;
    double* mean;
    double* var;
};

// Exact S1_H, S2_H by window sums for the widths flagged in `redo`: thread t
// walks the windows k of its run [k0, k1), w_{k+1} = w_k + H[k+L] - H[k],
// from w_{k0} summed directly (L <= 1024 values); S2 in 128-bit integers.
__global__ __launch_bounds__(NT) void lag_exact(LagExact a) {
    __shared__ __int128 l2[NT];
    __shared__ int64_t l1[NT];
    const uint32_t r = a.rows[blockIdx.x], j = threadIdx.x, n = a.n;
    const int32_t* row = a.G + (size_t)r * a.ldg;
    const int64_t c = a.cen[r];
    const uint32_t run = (n + NT - 1) / NT;
    const uint32_t b0 = min(n, j * run), b1 = min(n, b0 + run);
    for (uint32_t wi = 0; wi < a.nw; wi++) {
        if (!a.redo[(size_t)r * a.nw + wi]) continue;
        const uint32_t L = a.widths[wi];
        const int64_t o = a.out_slot[r] + wi;
        if (L == 0 || L >= n) continue;  // NaN already written
        const uint32_t K = n - L + 1;    // windows k = 0..n-L
        const uint32_t k0 = min(K, b0), k1 = min(K, b1);
        int64_t s1 = 0;
        __int128 s2 = 0;
        if (k0 < k1) {
            int64_t w = 0;
            for (uint32_t t = k0; t < k0 + L; t++) w += (int64_t)row[t] - c;
            for (uint32_t k = k0; k < k1; k++) {
                s1 += w;
                s2 += (__int128)w * w;
                if (k + 1 < k1) w += ((int64_t)row[k + L] - c) - ((int64_t)row[k] - c);
            }
        }
        l1[j] = s1;
        l2[j] = s2;
        __syncthreads();
        if (j == 0) {
            int64_t S1 = 0;
            __int128 S2 = 0;
            for (uint32_t t = 0; t < NT; t++) {
                S1 += l1[t];
                S2 += l2[t];
            }
            const double nwin = (double)K;
            const double m1 = (double)S1 / nwin;
            const double v = (double)S2 / nwin - m1 * m1;
            a.mean[o] = (m1 + (double)c * (double)L) / 100.0;
            a.var[o] = (v > 0.0 ? v : 0.0) / 1.0e4;
        }
        __syncthreads();
    }
}

}  // namespace

void moments_widths(Index& ix, const uint32_t* widths, uint32_t nw, double* mean, double* var,
                    uint32_t* inexact_rows, uint64_t* bins_out, double* lags_out, uint64_t* exact_widths) {
    Moments& M = ix.mom;
    CHK_REQUIRE(M.genome, CHK_ERR_INVALID, "set the genome table before computing moments");
    CHK_REQUIRE(nw > 0 && widths && mean && var, CHK_ERR_INVALID, "widths and outputs are required");
    for (uint32_t w = 0; w < nw; w++)
        CHK_REQUIRE(widths[w] >= 1 && widths[w] <= LB_LAGS, CHK_ERR_INVALID,
                    "widths of the FP32 route are 1.." + std::to_string(LB_LAGS) + " bins");
    const uint32_t S = ix.nshards, F = ix.nfiles;
    hipStream_t st = ix.stream;
    const size_t nout = (size_t)F * S * nw;
    // the twiddled passes' tables, exp(-2 pi i l k / (8 S)) in FP64 rounded to FP32
    std::vector<float2> th(LB_TW);
    for (uint32_t S : {8u, 64u, 512u})
        for (uint32_t k = 1; k < 8; k++)
            for (uint32_t l = 0; l < S; l++) {
                const double ang = -2.0 * M_PI * (double)(l * k) / (double)(8 * S);
                th[(S == 8 ? 0 : (S == 64 ? 56 : 504)) + (k - 1) * S + l] =
                    make_float2((float)std::cos(ang), (float)std::sin(ang));
            }
    DevBuf<float2> T;
    DevBuf<uint32_t> d_w, d_bad;
    DevBuf<double> d_mean, d_var, d_lags;
    DevBuf<int32_t> slot_of_fid, D;
    DevBuf<int64_t> cen, sumg, gseg, out_slot;
    DevBuf<dd> sumh2;
    DevBuf<float> part;
    DevBuf<uint32_t> redo, redo_rows;
    uint64_t nexact_w = 0;
    T.reserve(LB_TW);
    d_w.reserve(nw);
    d_mean.reserve(nout);
    d_var.reserve(nout);
    if (lags_out) d_lags.reserve((size_t)F * S * LB_LAGS);
    slot_of_fid.reserve(F);
    CHK_HIP(hipMemcpyAsync(T.p, th.data(), LB_TW * sizeof(float2), hipMemcpyHostToDevice, st));
    CHK_HIP(hipMemcpyAsync(d_w.p, widths, nw * 4, hipMemcpyHostToDevice, st));
    // slots without a table stay NaN
    std::vector<double> nanv(nout, NAN);
    CHK_HIP(hipMemcpyAsync(d_mean.p, nanv.data(), nout * 8, hipMemcpyHostToDevice, st));
    CHK_HIP(hipMemcpyAsync(d_var.p, nanv.data(), nout * 8, hipMemcpyHostToDevice, st));
    std::vector<int32_t> sof(F, -1);
    uint32_t nbad_rows = 0;
    uint64_t bins = 0;
    for (uint32_t s = 0; s < S; s++) {
        const uint32_t n = M.nbins[s];
        if (n < 2) continue;
        std::vector<uint32_t> files;
        for (uint32_t f = 0; f < F; f++)
            if (ix.fs_count[(size_t)f * S + s]) files.push_back(f);
        if (files.empty()) continue;
        const uint32_t ldg = n + 1;
        const uint32_t nblk = (n + LB_HOP - 1) / LB_HOP, npairs = (nblk + 1) / 2;
        const uint32_t nchunk = (npairs + LB_PAIRS - 1) / LB_PAIRS;
        size_t free_b = 0, total_b = 0;
        CHK_HIP(hipMemGetInfo(&free_b, &total_b));
        const uint64_t row_bytes = (uint64_t)ldg * 4 + (uint64_t)nchunk * LB_LAGS * 8 + 64;
        const uint32_t B = (uint32_t)std::max<uint64_t>(
            1, std::min<uint64_t>({files.size(), (uint64_t)(free_b / 4) / row_bytes, 2048}));
        D.reserve((size_t)B * ldg);
        part.reserve((size_t)B * nchunk * LB_LAGS * 2);
        cen.reserve(B);
        sumg.reserve(B);
        sumh2.reserve(B);
        gseg.reserve(B);
        out_slot.reserve(B);
        d_bad.reserve(B);
        redo.reserve((size_t)B * nw);
        const uint32_t seg = (ldg + CH - 1) / CH * CH;  // one segment: no evaluation split here
        const size_t NSUB = (size_t)ix.ngroups * ix.nlayers;
        const LayerSeg& Ll = ix.lseg[(size_t)s * NSUB + NSUB - 1];
        const uint64_t soff = ix.lseg[(size_t)s * NSUB].off, scnt = Ll.off + Ll.n - soff;
        for (size_t f0 = 0; f0 < files.size(); f0 += B) {
            const uint32_t nb = (uint32_t)std::min<size_t>(B, files.size() - f0);
            std::vector<int64_t> oo(nb);
            for (uint32_t k = 0; k < nb; k++) {
                sof[files[f0 + k]] = (int32_t)k;
                oo[k] = ((int64_t)files[f0 + k] * S + s) * nw;
            }
            CHK_HIP(hipMemcpyAsync(slot_of_fid.p, sof.data(), F * 4, hipMemcpyHostToDevice, st));
            CHK_HIP(hipMemcpyAsync(out_slot.p, oo.data(), nb * 8, hipMemcpyHostToDevice, st));
            CHK_HIP(hipMemsetAsync(D.p, 0, (size_t)nb * ldg * 4, st));
            CHK_HIP(hipMemsetAsync(d_bad.p, 0, (size_t)nb * 4, st));
            hipEvent_t t0 = ix.prof.begin("m32_coverage", st);
            const unsigned grid = (unsigned)std::min<uint64_t>((scnt + NT - 1) / NT, (uint64_t)ix.ncu * 16);
            if (grid)
                hipLaunchKernelGGL(impulses, dim3(grid), dim3(NT), 0, st, ix.l_start.p + soff,
                                   ix.l_end.p + soff, ix.l_fid.p + soff, scnt, slot_of_fid.p, n, ldg, D.p);
            hipLaunchKernelGGL(coverage_rows, dim3(nb), dim3(NT), 0, st, D.p, ldg, n, cen.p, sumg.p,
                               sumh2.p, seg, 1u, gseg.p);
            CHK_HIP(hipGetLastError());
            ix.prof.end("m32_coverage", t0, st);
            LagArgs la{D.p, ldg, n, cen.p, T.p, npairs, nchunk, part.p, d_bad.p};
            t0 = ix.prof.begin("m32_fft", st);
            hipLaunchKernelGGL(lag_fft32, dim3(nchunk, nb), dim3(LB_T), 0, st, la);
            CHK_HIP(hipGetLastError());
            ix.prof.end("m32_fft", t0, st);
            Lag32Eval ea{D.p, ldg, n, cen.p, sumg.p, part.p, nchunk, d_w.p, nw, out_slot.p,
                         d_mean.p, d_var.p, nullptr, redo.p};
            DevBuf<double> lrow;
            if (lags_out) {
                lrow.reserve((size_t)nb * LB_LAGS);
                ea.A_out = lrow.p;
            }
            t0 = ix.prof.begin("m32_eval", st);
            hipLaunchKernelGGL(lag_eval32, dim3(nb), dim3(NT), 0, st, ea);
            CHK_HIP(hipGetLastError());
            ix.prof.end("m32_eval", t0, st);
            // widths whose FP32 error bound is not small (and short rows):
            // exact window sums
            std::vector<uint32_t> rf((size_t)nb * nw), rl;
            CHK_HIP(hipMemcpyAsync(rf.data(), redo.p, rf.size() * 4, hipMemcpyDeviceToHost, st));
            CHK_HIP(hipStreamSynchronize(st));
            for (uint32_t k = 0; k < nb; k++) {
                bool any = false;
                for (uint32_t w = 0; w < nw; w++) {
                    any |= rf[(size_t)k * nw + w] != 0;
                    nexact_w += rf[(size_t)k * nw + w] != 0;
                }
                if (any) rl.push_back(k);
            }
            if (!rl.empty()) {
                redo_rows.reserve(rl.size());
                CHK_HIP(hipMemcpyAsync(redo_rows.p, rl.data(), rl.size() * 4, hipMemcpyHostToDevice, st));
                LagExact xa{D.p, ldg, n, cen.p, d_w.p, nw, redo.p, redo_rows.p, out_slot.p, d_mean.p, d_var.p};
                t0 = ix.prof.begin("m32_exact", st);
                hipLaunchKernelGGL(lag_exact, dim3((uint32_t)rl.size()), dim3(NT), 0, st, xa);
                CHK_HIP(hipGetLastError());
                ix.prof.end("m32_exact", t0, st);
            }
            std::vector<uint32_t> bad(nb);
            CHK_HIP(hipMemcpyAsync(bad.data(), d_bad.p, nb * 4, hipMemcpyDeviceToHost, st));
            if (lags_out) {
                for (uint32_t k = 0; k < nb; k++)
                    CHK_HIP(hipMemcpyAsync(lags_out + ((size_t)files[f0 + k] * S + s) * LB_LAGS,
                                           lrow.p + (size_t)k * LB_LAGS, LB_LAGS * 8,
                                           hipMemcpyDeviceToHost, st));
            }
            CHK_HIP(hipStreamSynchronize(st));
            for (uint32_t k = 0; k < nb; k++) {
                nbad_rows += bad[k] ? 1 : 0;
                sof[files[f0 + k]] = -1;
            }
            bins += (uint64_t)nb * n;
        }
    }
    CHK_HIP(hipMemcpyAsync(mean, d_mean.p, nout * 8, hipMemcpyDeviceToHost, st));
    CHK_HIP(hipMemcpyAsync(var, d_var.p, nout * 8, hipMemcpyDeviceToHost, st));
    CHK_HIP(hipStreamSynchronize(st));
    if (inexact_rows) *inexact_rows = nbad_rows;
    if (bins_out) *bins_out = bins;
    if (exact_widths) *exact_widths = nexact_w;
}

}  // namespace chk
