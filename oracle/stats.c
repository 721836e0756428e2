/*
 * stats.c -- CPU restatement of the analytic p-value path
 * (TEST INFRASTRUCTURE ONLY; see oracle.h).
 *
 *  - orc_group_lengths: build_query_chrom_data's grouping
 *    (src/fourier/moments.rs:138-177): l = (e-s)/100 in FP64, sorted
 *    ascending, grouped by l_int = max(1, ceil(l)); per group
 *    sum_scale = sum l/l_int, sum_scale2 = sum (l/l_int)^2, count.
 *  - orc_analytic_stats: compute_analytic_stats (src/fourier/stats.rs:34-88)
 *    with the moment-store lookup rule of src/io/momentstore.rs:66-79
 *    (None when l_int == 0, l_int >= n_bins, or the compact slot is past the
 *    stored table).  Each group carries the DB-side chromosome slot its
 *    query chromosome resolves to (-1 = not stored for this DB source).
 *  - orc_nb_llr / orc_poisson_llr / orc_erfc_as: stats.rs:93-126, the erfc
 *    being Abramowitz & Stegun 7.1.26 evaluated in FP64.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>

double orc_erfc_as(double x) {
    if (x < 0.0) return 2.0 - orc_erfc_as(-x);
    double t = 1.0 / (1.0 + 0.3275911 * x);
    double p = t * (0.254829592 +
                    t * (-0.284496736 + t * (1.421413741 + t * (-1.453152027 + t * 1.061405429))));
    return p * exp(-x * x);
}

int orc_poisson_llr(double mu, double obs, double* llr) {
    if (mu <= 0.0 || obs <= 0.0) return 0;
    *llr = obs * log(obs / mu) - (obs - mu);
    return 1;
}

int orc_nb_llr(double mu, double var, double obs, double* llr) {
    if (mu <= 0.0 || var <= mu) return 0;
    double r = mu * mu / (var - mu);
    double p = mu / var;
    double p_o = r / (r + obs);
    if (!(0.0 < p && p < 1.0 && 0.0 < p_o && p_o < 1.0)) return 0;
    double theta = log((1.0 - p_o) / (1.0 - p));
    *llr = obs * theta + r * log(p_o / p);
    return 1;
}

static int cmp_f64(const void* a, const void* b) {
    double x = *(const double*)a, y = *(const double*)b;
    return x < y ? -1 : x > y;
}

size_t orc_group_lengths(size_t n, const uint32_t* len_bp, uint32_t* l_int, double* sum_scale,
                         double* sum_scale2, uint64_t* count) {
    double* l = (double*)malloc((n ? n : 1) * sizeof(double));
    for (size_t i = 0; i < n; i++) l[i] = (double)len_bp[i] / (double)ORC_BIN_SIZE;
    qsort(l, n, sizeof(double), cmp_f64);
    size_t ng = 0, i = 0;
    while (i < n) {
        double c = ceil(l[i]);
        uint32_t li = c < 1.0 ? 1u : (uint32_t)c;
        double s0 = l[i] / (double)li;
        double ss = s0, ss2 = s0 * s0;
        uint64_t cnt = 1;
        i++;
        while (i < n) {
            double c2 = ceil(l[i]);
            uint32_t li2 = c2 < 1.0 ? 1u : (uint32_t)c2;
            if (li2 != li) break;
            double s = l[i] / (double)li;
            ss += s;
            ss2 += s * s;
            cnt++;
            i++;
        }
        l_int[ng] = li;
        sum_scale[ng] = ss;
        sum_scale2[ng] = ss2;
        count[ng] = cnt;
        ng++;
    }
    free(l);
    return ng;
}

int orc_analytic_stats(size_t ngroups, const uint32_t* l_int, const double* sum_scale,
                       const double* sum_scale2, const uint64_t* count,
                       const int32_t* chrom_slot, const uint32_t* slot_n_bins,
                       const uint64_t* slot_off, const uint64_t* slot_len,
                       const double* tab_mean, const double* tab_var, double observed,
                       double* p, double* llr, int* has_llr) {
    double mu = 0.0, var = 0.0, mu_raw = 0.0;
    int any = 0;
    for (size_t g = 0; g < ngroups; g++) {
        int32_t s = chrom_slot[g];
        if (s < 0) continue;
        uint32_t li = l_int[g];
        if (li == 0 || li >= slot_n_bins[s]) continue;
        size_t ci = orc_compact_index(li);
        if (ci >= slot_len[s]) continue;
        double m = tab_mean[slot_off[s] + ci], v = tab_var[slot_off[s] + ci];
        mu += m * sum_scale[g];
        var += v * sum_scale2[g];
        mu_raw += m / (double)li * (double)count[g];
        any = 1;
    }
    if (!any) return 0;
    double l = 0.0;
    int ok;
    if (var > mu) {
        ok = orc_nb_llr(mu, var, observed, &l);
    } else {
        double eff = mu_raw > 0.0 ? mu / mu_raw : 1.0;
        ok = orc_poisson_llr(mu_raw, observed / eff, &l);
    }
    *has_llr = ok;
    *llr = ok ? l : 0.0;
    if (ok && l > 0.0) {
        double tail = orc_erfc_as(sqrt(l)) * 0.5;
        *p = observed >= mu ? tail : 1.0 - tail;
    } else {
        *p = 1.0;
    }
    return 1;
}
