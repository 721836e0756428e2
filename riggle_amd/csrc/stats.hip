// stats.hip -- analytic NB / Poisson p-values per (query file, DB file) cell.
//
// Restates compute_analytic_stats (src/fourier/stats.rs:34-88) for every cell
// with a nonzero overlap count (src/tasks/query/pvalues.rs:29-35, 58-65):
//   mu = sum mean(l) * sum_scale,  var = sum var(l) * sum_scale2,
//   mu_raw = sum mean(l) / l * count  over the query's length groups whose
// shard has a stored moment table for the DB file (src/io/momentstore.rs:66-79:
// a group is skipped when l >= n_bins or its compact slot is past the table);
// NB LLR when var > mu, else Poisson on observed / (mu / mu_raw); p =
// erfc(sqrt(LLR)) / 2 (enrichment) or 1 - that (depletion), with the
// Abramowitz & Stegun 7.1.26 erfc of stats.rs:116-126, all in FP64.  observed
// is the exact overlap in bins, bp / 100.
#include <algorithm>
#include <cmath>

#include "engine.hpp"

namespace chk {

// src/fourier/moments.rs:58-66, evaluated with the C library's log exactly
// as the reference's f64::ln.
size_t compact_index(size_t l) {
    if (l <= 100) return l - 1;
    size_t k = (size_t)(std::log((double)l / 100.0) / std::log(1.01));
    return 100 + k - 1;
}

// Stored block sizes of a compact table (src/fourier/moments.rs:225-247).
size_t compact_ls(uint32_t n_bins, std::vector<uint32_t>* ls) {
    if (ls) ls->clear();
    if (n_bins <= 1) return 0;
    size_t max_l = n_bins - 1, k = 0;
    for (size_t l = 1; l <= std::min<size_t>(100, max_l); l++, k++)
        if (ls) ls->push_back((uint32_t)l);
    double l_f = 100.0;
    size_t prev = 100;
    for (;;) {
        l_f *= 1.01;
        size_t l = (size_t)std::ceil(l_f);
        if (l > max_l) break;
        if (l > prev) {
            if (ls) ls->push_back((uint32_t)l);
            k++;
            prev = l;
        }
    }
    return k;
}

// Length groups of every (query file, shard): l_int = max(1, ceil(len/100))
// = (len + 99) / 100; sum_scale = sum(len) / (100 l_int) and sum_scale2 =
// sum(len^2) / (100 l_int)^2 from exact integer sums.
void query_groups(Query& q, const uint32_t* shard, const uint32_t* start, const uint32_t* end,
                  const uint32_t* qfile) {
    Index& ix = *q.idx;
    q.has_groups = false;
    if (!ix.mom.genome) return;
    const uint32_t S = ix.nshards;
    const size_t nseg = (size_t)q.nfiles * S;
    std::vector<uint32_t> cnt(nseg + 1, 0);
    for (uint64_t i = 0; i < q.n_in; i++) {
        uint32_t c = shard[i];
        if (c >= S || ix.mom.nbins[c] == 0) continue;
        cnt[(size_t)(qfile ? qfile[i] : 0) * S + c + 1]++;
    }
    for (size_t g = 0; g < nseg; g++) cnt[g + 1] += cnt[g];
    std::vector<uint32_t> lint(cnt[nseg]);
    std::vector<uint32_t> len(cnt[nseg]);
    {
        std::vector<uint32_t> fill(cnt.begin(), cnt.end() - 1);
        for (uint64_t i = 0; i < q.n_in; i++) {
            uint32_t c = shard[i];
            if (c >= S || ix.mom.nbins[c] == 0) continue;
            size_t g = (size_t)(qfile ? qfile[i] : 0) * S + c;
            uint32_t L = end[i] - start[i];
            len[fill[g]] = L;
            lint[fill[g]++] = (uint32_t)(((uint64_t)L + 99) / 100);
        }
    }
    std::vector<uint32_t> off(nseg + 1, 0), gl, gci, gsh;
    std::vector<double> gss, gss2, gcnt;
    std::vector<uint64_t> order;
    for (size_t g = 0; g < nseg; g++) {
        off[g] = (uint32_t)gl.size();
        size_t a = cnt[g], z = cnt[g + 1];
        if (z == a) continue;
        order.resize(z - a);
        for (size_t k = a; k < z; k++) order[k - a] = ((uint64_t)lint[k] << 32) | len[k];
        std::sort(order.begin(), order.end());
        size_t k = 0;
        while (k < order.size()) {
            uint32_t li = (uint32_t)(order[k] >> 32);
            unsigned __int128 s1 = 0, s2 = 0;
            uint64_t c = 0;
            while (k < order.size() && (uint32_t)(order[k] >> 32) == li) {
                uint64_t L = (uint32_t)order[k];
                s1 += L;
                s2 += (unsigned __int128)L * L;
                c++;
                k++;
            }
            double d = 100.0 * (double)li;
            gl.push_back(li);
            gci.push_back((uint32_t)compact_index(li));
            gsh.push_back((uint32_t)(g % S));
            gss.push_back((double)s1 / d);
            gss2.push_back((double)s2 / (d * d));
            gcnt.push_back((double)c);
        }
    }
    off[nseg] = (uint32_t)gl.size();
    q.ngroups = gl.size();
    hipStream_t st = q.stream;
    q.g_off.reserve(nseg + 1);
    CHK_HIP(hipMemcpyAsync(q.g_off.p, off.data(), (nseg + 1) * 4, hipMemcpyHostToDevice, st));
    if (q.ngroups) {
        q.g_lint.reserve(q.ngroups);
        q.g_ci.reserve(q.ngroups);
        q.g_shard.reserve(q.ngroups);
        q.g_ss.reserve(q.ngroups);
        q.g_ss2.reserve(q.ngroups);
        q.g_cnt.reserve(q.ngroups);
        CHK_HIP(hipMemcpyAsync(q.g_lint.p, gl.data(), q.ngroups * 4, hipMemcpyHostToDevice, st));
        CHK_HIP(hipMemcpyAsync(q.g_ci.p, gci.data(), q.ngroups * 4, hipMemcpyHostToDevice, st));
        CHK_HIP(hipMemcpyAsync(q.g_shard.p, gsh.data(), q.ngroups * 4, hipMemcpyHostToDevice, st));
        CHK_HIP(hipMemcpyAsync(q.g_ss.p, gss.data(), q.ngroups * 8, hipMemcpyHostToDevice, st));
        CHK_HIP(hipMemcpyAsync(q.g_ss2.p, gss2.data(), q.ngroups * 8, hipMemcpyHostToDevice, st));
        CHK_HIP(hipMemcpyAsync(q.g_cnt.p, gcnt.data(), q.ngroups * 8, hipMemcpyHostToDevice, st));
    }
    CHK_HIP(hipStreamSynchronize(st));
    q.has_groups = true;
}

void moments_upload(Index& ix, const int64_t* slot_off, const uint32_t* slot_len, uint64_t total,
                    const double* mean, const double* var) {
    CHK_REQUIRE(ix.mom.genome, CHK_ERR_INVALID, "set the genome table before the moments");
    const size_t ns = (size_t)ix.nfiles * ix.nshards;
    Moments& M = ix.mom;
    M.slot_off.assign(slot_off, slot_off + ns);
    M.slot_len.assign(slot_len, slot_len + ns);
    std::vector<uint8_t> has(ix.nfiles, 0);
    for (size_t s = 0; s < ns; s++) {
        if (slot_off[s] < 0) continue;
        CHK_REQUIRE((uint64_t)slot_off[s] + slot_len[s] <= total, CHK_ERR_INVALID,
                    "moment slot exceeds the table");
        has[s / ix.nshards] = 1;
    }
    M.total = total;
    hipStream_t st = ix.stream;
    M.d_slot_off.reserve(ns);
    M.d_slot_len.reserve(ns);
    M.d_has_entry.reserve(ix.nfiles);
    M.mean.reserve(total);
    M.var.reserve(total);
    CHK_HIP(hipMemcpyAsync(M.d_slot_off.p, slot_off, ns * 8, hipMemcpyHostToDevice, st));
    CHK_HIP(hipMemcpyAsync(M.d_slot_len.p, slot_len, ns * 4, hipMemcpyHostToDevice, st));
    CHK_HIP(hipMemcpyAsync(M.d_has_entry.p, has.data(), ix.nfiles, hipMemcpyHostToDevice, st));
    if (total) {
        CHK_HIP(hipMemcpyAsync(M.mean.p, mean, total * 8, hipMemcpyHostToDevice, st));
        CHK_HIP(hipMemcpyAsync(M.var.p, var, total * 8, hipMemcpyHostToDevice, st));
    }
    CHK_HIP(hipStreamSynchronize(st));
    M.present = true;
}

namespace {

__device__ inline double erfc_as(double x) {
    // Abramowitz & Stegun 7.1.26 (src/fourier/stats.rs:116-126)
    const bool neg = x < 0.0;
    const double ax = neg ? -x : x;
    const double t = 1.0 / (1.0 + 0.3275911 * ax);
    const double p = t * (0.254829592 +
                          t * (-0.284496736 + t * (1.421413741 + t * (-1.453152027 + t * 1.061405429))));
    const double r = p * exp(-ax * ax);
    return neg ? 2.0 - r : r;
}

struct StatsArgs {
    uint32_t nq, F, S;
    const uint64_t* counts;
    const uint64_t* bp;
    const uint32_t* g_off;
    const uint32_t* g_lint;
    const uint32_t* g_ci;
    const uint32_t* g_shard;
    const double* g_ss;
    const double* g_ss2;
    const double* g_cnt;
    const uint32_t* nbins;
    const int64_t* slot_off;
    const uint32_t* slot_len;
    const uint8_t* has_entry;
    const double* mean;
    const double* var;
    double* p;
    double* llr;
    uint8_t* flags;
};

// WPC waves per cell, one cell per block; the cell's lanes stride over the
// query file's length groups.  Every early exit below is uniform over the
// block, so the barriers are reached by all of its threads.  Measured on one
// 1M-interval query file (8,383 length groups) against 1,000 DB files: 4, 8
// and 16 waves per cell take the same time.
constexpr uint32_t SCAP = 256;  // shards whose slot data a block keeps in LDS
constexpr int PV_WPC = 4;       // waves per cell

template <int WPC>
__global__ __launch_bounds__(WPC * 64) void pvalue_kernel(StatsArgs a) {
    constexpr uint32_t CPB = 1;  // one cell per block: the LDS slot table is the cell's
    __shared__ double red[WPC][3];
    __shared__ uint32_t red_any[WPC];
    __shared__ int64_t s_so[SCAP];
    __shared__ uint32_t s_sl[SCAP], s_nb[SCAP];
    const uint32_t wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const uint32_t sub = wv % WPC;  // wave within the cell
    const uint32_t cell = blockIdx.x * CPB + wv / WPC;
    if (cell >= a.nq * a.F) return;
    const uint32_t r = cell / a.F, f = cell % a.F;
    const uint64_t count = a.counts[cell];
    if (count == 0 || !a.has_entry[f]) {  // no record: p = 1 (docs/stats.md:214-216), llr 0
        if (sub == 0 && lane == 0) {
            a.p[cell] = 1.0;
            a.llr[cell] = 0.0;
            a.flags[cell] = 0;
        }
        return;
    }
    const uint32_t g0 = a.g_off[(size_t)r * a.S], g1 = a.g_off[(size_t)(r + 1) * a.S];
    // the cell's per-shard slot data (slot offset, length, table size) for the
    // first SCAP shards in LDS, filled while the first groups load
    for (uint32_t c = threadIdx.x; c < a.S && c < SCAP; c += WPC * 64) {
        s_so[c] = a.slot_off[(size_t)f * a.S + c];
        s_sl[c] = a.slot_len[(size_t)f * a.S + c];
        s_nb[c] = a.nbins[c];
    }
    __syncthreads();
    double mu = 0.0, vr = 0.0, mu_raw = 0.0;
    uint32_t any = 0;
    // four groups per lane and round: each level of the group -> slot ->
    // table chain is loaded for all four at once (at clamped indices, so the
    // loads are unconditional), then the groups that have a table entry add
    // in group order.  A group counts only when the DB file stores moments
    // for its shard (slot_off >= 0) and the table reaches its block size.
    constexpr uint32_t U = 4, GS = WPC * 64;
    for (uint32_t g = g0 + sub * 64 + lane; g < g1; g += U * GS) {
        uint32_t c[U], li[U], ci[U];
        double ss[U], ss2[U], cn[U];
        bool ok[U];
#pragma unroll
        for (uint32_t u = 0; u < U; u++) {
            const uint32_t gu = g + u * GS;
            ok[u] = gu < g1;
            const uint32_t gi = ok[u] ? gu : g0;
            c[u] = a.g_shard[gi];
            li[u] = a.g_lint[gi];
            ci[u] = a.g_ci[gi];
            ss[u] = a.g_ss[gi];
            ss2[u] = a.g_ss2[gi];
            cn[u] = a.g_cnt[gi];
        }
        int64_t so[U];
        uint32_t sl[U], nb[U];
#pragma unroll
        for (uint32_t u = 0; u < U; u++) {
            if (c[u] < SCAP) {
                so[u] = s_so[c[u]];
                sl[u] = s_sl[c[u]];
                nb[u] = s_nb[c[u]];
            } else {
                so[u] = a.slot_off[(size_t)f * a.S + c[u]];
                sl[u] = a.slot_len[(size_t)f * a.S + c[u]];
                nb[u] = a.nbins[c[u]];
            }
        }
        double m[U], v[U];
#pragma unroll
        for (uint32_t u = 0; u < U; u++) {
            ok[u] = ok[u] && so[u] >= 0 && li[u] < nb[u] && ci[u] < sl[u];
            const int64_t ti = ok[u] ? so[u] + ci[u] : 0;  // the file has a table: entry 0 exists
            m[u] = a.mean[ti];
            v[u] = a.var[ti];
        }
#pragma unroll
        for (uint32_t u = 0; u < U; u++) {
            if (ok[u]) {
                mu += m[u] * ss[u];
                vr += v[u] * ss2[u];
                mu_raw += m[u] / (double)li[u] * cn[u];
                any = 1;
            }
        }
    }
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) {
        mu += __shfl_xor(mu, d, 64);
        vr += __shfl_xor(vr, d, 64);
        mu_raw += __shfl_xor(mu_raw, d, 64);
        any |= __shfl_xor(any, d, 64);
    }
    if (WPC > 1) {
        if (lane == 0) {
            red[wv][0] = mu;
            red[wv][1] = vr;
            red[wv][2] = mu_raw;
            red_any[wv] = any;
        }
        __syncthreads();
        if (sub == 0 && lane == 0) {
            for (uint32_t w = wv + 1; w < wv + WPC; w++) {
                mu += red[w][0];
                vr += red[w][1];
                mu_raw += red[w][2];
                any |= red_any[w];
            }
        }
    }
    if (sub != 0 || lane != 0) return;
    if (!any) {  // no shared chromosome: no record (stats.rs:61-63)
        a.p[cell] = 1.0;
        a.llr[cell] = 0.0;
        a.flags[cell] = 0;
        return;
    }
    const double obs = (double)a.bp[cell] / 100.0;
    double l = 0.0;
    bool ok = false;
    if (vr > mu) {  // NB (stats.rs:102-113)
        if (mu > 0.0 && vr > mu) {
            const double rr = mu * mu / (vr - mu), pp = mu / vr, po = rr / (rr + obs);
            if (0.0 < pp && pp < 1.0 && 0.0 < po && po < 1.0) {
                l = obs * log((1.0 - po) / (1.0 - pp)) + rr * log(po / pp);
                ok = true;
            }
        }
    } else {  // Poisson on count units (stats.rs:70-77, 93-99)
        const double eff = mu_raw > 0.0 ? mu / mu_raw : 1.0;
        const double oc = obs / eff;
        if (mu_raw > 0.0 && oc > 0.0) {
            l = oc * log(oc / mu_raw) - (oc - mu_raw);
            ok = true;
        }
    }
    double p = 1.0;
    if (ok && l > 0.0) {
        const double tail = erfc_as(sqrt(l)) * 0.5;
        p = obs >= mu ? tail : 1.0 - tail;
    }
    a.p[cell] = p;
    a.llr[cell] = ok ? l : 0.0;
    a.flags[cell] = (uint8_t)(1u | (ok ? 2u : 0u));
}

}  // namespace

void pvalues_run(Query& q, const uint64_t* d_counts, const uint64_t* d_bp, double* d_p,
                 double* d_llr, uint8_t* d_flags, hipStream_t st) {
    Index& ix = *q.idx;
    CHK_REQUIRE(ix.mom.present, CHK_ERR_NOT_FOUND, "the index has no moment tables");
    CHK_REQUIRE(q.has_groups, CHK_ERR_INVALID,
                "query created before the index had a genome table: no length groups");
    CHK_REQUIRE((uint64_t)q.nfiles * ix.nfiles < (1ull << 32), CHK_ERR_INVALID,
                "query files x DB files must stay below 2^32 cells");
    const uint32_t cells = q.nfiles * ix.nfiles;
    StatsArgs a{q.nfiles, ix.nfiles, ix.nshards, d_counts, d_bp, q.g_off.p, q.g_lint.p, q.g_ci.p,
                q.g_shard.p, q.g_ss.p, q.g_ss2.p, q.g_cnt.p, ix.mom.d_nbins.p, ix.mom.d_slot_off.p,
                ix.mom.d_slot_len.p, ix.mom.d_has_entry.p, ix.mom.mean.p, ix.mom.var.p, d_p, d_llr,
                d_flags};
    hipEvent_t t0 = ix.prof.begin("pvalues", st);
    // four waves per cell whatever the batch size: a single query file's
    // cells still fill the device, and a cell's FP64 sums are taken in the
    // same order in every batch, so its p-value never depends on the batch
    hipLaunchKernelGGL(pvalue_kernel<PV_WPC>, dim3(cells), dim3(PV_WPC * 64), 0, st, a);
    CHK_HIP(hipGetLastError());
    ix.prof.end("pvalues", t0, st);
}

}  // namespace chk
