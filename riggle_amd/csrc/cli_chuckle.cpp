// cli_chuckle.cpp -- the `chuckle` command line (add | query | info |
// montecarlo) over the MI355X library, a drop-in for the reference's CLI
// (src/main.rs:17-78) on the reference's on-disk database format:
//   DB/meta.json                       sid_map, layer_config, shards, num_layers
//                                      (src/io/meta.rs:101-110)
//   DB/shards/<shard>/layer_K.pos      sorted (start, end) u32 LE pairs
//   DB/shards/<shard>/layer_K.sid      parallel u32 LE D_SIDs
//   DB/shards/<shard>/layer_K.idx      u64 LE jump table, tile = max_size(K)/2
//                                      (src/io/layer.rs:34-199, src/io/jump.rs:19-169)
// Moment tables are kept in this build's own file, DB/momentmap.chk (the
// reference's rkyv store is not read); `query --stats` rebuilds them on the
// GPU when that file is missing or stale.
//
// Every step goes through the library's C ABI: chk_db_add (parse over host
// threads, bucket sort on the device, layer writing), chk_index_open,
// chk_query_open (parallel BED parsing), chk_query_stats / chk_sweep and
// chk_write_results (StatsOutput JSON, src/stats/mod.rs:31-60, sorted as
// src/main.rs:146-211).
//
// `chuckle montecarlo` is the reference's Monte-Carlo shift permutation
// driver (src/bin/montecarlo.rs) on the batched sweep: trials are generated in
// memory and each batch of them is one multi-file sweep on the device.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <random>
#include <set>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <utility>
#include <vector>

#include "chuckle_hip.h"
#include "dbfile.hpp"

namespace {

using chk::db::Iv;
using chk::db::Meta;
using chk::db::exists;
using chk::db::join;

struct Fail : std::runtime_error {
    using std::runtime_error::runtime_error;
};

void chk(int code) {
    if (code != CHK_OK) throw Fail(chk_last_error());
}

struct IndexHandle {
    chk_index* p = nullptr;
    ~IndexHandle() {
        if (p) chk_index_free(p);
    }
};

// Intervals of one BED file per chromosome, chromosomes in name order.
std::map<std::string, std::vector<Iv>> bed_by_name(const std::string& path) {
    const chk::db::BedFile f = chk::db::parse_bed(path, chk::db::host_threads());
    std::map<std::string, std::vector<Iv>> out;
    for (size_t s = 0; s < f.shards.size(); s++) out[f.shards[s]] = f.ivs[s];
    return out;
}

// ------------------------------------------------------------------ add

void run_add(const std::string& input, const std::string& db, size_t batch_size, bool stats,
             const std::vector<std::string>& chroms, bool sorted_sids) {
    std::vector<const char*> wl;
    for (auto& c : chroms) wl.push_back(c.c_str());
    chk(chk_db_add(input.c_str(), db.c_str(), batch_size, stats ? 1 : 0, wl.data(), (uint32_t)wl.size(),
                   sorted_sids ? 1 : 0, 0));
}

// ------------------------------------------------------------------ query

void run_query(const std::string& db, const std::string& qpath, const std::string& out, bool stats,
               bool intervals, size_t batch_size) {
    if (intervals)
        throw Fail("--intervals (pair listing) is not supported by this build; use counts or --stats");
    if (!exists(join(db, "meta.json"))) throw Fail("Database not found at " + db);
    // query files sorted by path (src/tasks/query/files.rs:26-39)
    std::vector<std::string> files;
    if (chk::db::is_dir(qpath)) {
        for (auto& p : chk::db::list_dir(qpath))
            if (chk::db::is_bed_file(p)) files.push_back(p);
        std::sort(files.begin(), files.end());
    } else {
        files.push_back(qpath);
    }
    IndexHandle ix;
    chk(chk_index_open(db.c_str(), 0, &ix.p));
    uint32_t S = 0, F = 0;
    chk(chk_index_info(ix.p, &S, &F, nullptr, nullptr));
    std::vector<std::string> dnames(F);
    std::vector<const char*> dptr(F);
    for (uint32_t d = 0; d < F; d++) {
        char buf[4096];
        chk(chk_index_file_name(ix.p, d, buf, sizeof buf, nullptr));
        dnames[d] = buf;
        dptr[d] = dnames[d].c_str();
    }
    if (stats && chk_moments_download(ix.p, nullptr, nullptr, nullptr, nullptr, nullptr) != CHK_OK) {
        int32_t rows = 0, exact = 0;
        chk(chk_moments_build(ix.p, &rows, &exact));
    }
    std::vector<std::string> qnames;
    std::vector<uint64_t> cnt, bp;
    std::vector<double> p, llr;
    std::vector<uint8_t> fl;
    const size_t bs = std::max<size_t>(1, batch_size ? batch_size : files.size());
    for (size_t b0 = 0; b0 < files.size(); b0 += bs) {
        std::vector<const char*> paths;
        for (size_t i = b0; i < std::min(files.size(), b0 + bs); i++) paths.push_back(files[i].c_str());
        std::vector<uint32_t> row(paths.size());
        chk_query* q = nullptr;
        chk(chk_query_open(ix.p, (uint32_t)paths.size(), paths.data(), row.data(), &q));
        if (!q) continue;  // every file of the batch is empty (files.rs:79-101)
        uint32_t R = 0;
        for (size_t k = 0; k < paths.size(); k++)
            if (row[k] != 0xFFFFFFFFu) {
                qnames.push_back(chk::db::basename_of(paths[k]));
                R++;
            }
        const size_t cells = (size_t)R * F, o = cnt.size();
        cnt.resize(o + cells);
        bp.resize(o + cells);
        p.resize(o + cells);
        llr.resize(o + cells);
        fl.resize(o + cells);
        const int rc = stats ? chk_query_stats(ix.p, q, cnt.data() + o, bp.data() + o, p.data() + o,
                                               llr.data() + o, fl.data() + o)
                             : chk_sweep(ix.p, q, cnt.data() + o, bp.data() + o);
        chk_query_free(q);
        chk(rc);
    }
    std::vector<const char*> qptr;
    for (auto& n : qnames) qptr.push_back(n.c_str());
    chk(chk_write_results(out.c_str(), (uint32_t)qnames.size(), qptr.data(), F, dptr.data(), cnt.data(),
                          bp.data(), p.data(), llr.data(), fl.data(), stats ? 1 : 0));
    const Meta meta = Meta::load(db);
    std::printf("Query complete: %zu query files \xc3\x97 %zu database sources \xe2\x86\x92 %s\n",
                qnames.size(), meta.sid_map.size(), out.c_str());
}

// ------------------------------------------------------------------ info

void run_info(const std::string& db) {
    const Meta m = Meta::load(db);
    std::printf("Database: %s\n", db.c_str());
    std::printf("Layer config: min_size=%u, growth_factor=%u\n", m.lc.min_size, m.lc.growth_factor);
    std::printf("Layers used: %u\n", m.num_layers);
    std::string sh;
    for (size_t k = 0; k < m.shards.size(); k++) sh += (k ? ", " : "") + m.shards[k];
    std::printf("Shards (%zu): %s\n", m.shards.size(), sh.c_str());
    std::printf("\nSources (%zu):\n", m.sid_map.size());
    for (auto& kv : m.sid_map) std::printf("  SID %4u: %s\n", kv.first, kv.second.c_str());
}

// ------------------------------------------------------------------ montecarlo

// rand 0.8.5 StdRng as src/bin/montecarlo.rs seeds and draws it.  Restated
// from the published algorithms of rand_core 0.6.4 / rand_chacha 0.3.1 / rand
// 0.8.5 (none is vendored with the reference):
//   - ChaCha12 keystream (D. J. Bernstein's ChaCha, 12 rounds), 64-bit block
//     counter in words 12-13 and a zero 64-bit stream id in words 14-15, read
//     as little-endian u32 words block after block; a u64 is two consecutive
//     words, low first;
//   - seed_from_u64: the 32-byte key is eight PCG32 (XSH-RR) outputs;
//   - gen_range(lo..=hi) on u32 / i64 (UniformInt::sample_single_inclusive):
//     range = hi - lo + 1, zone = (range << clz(range)) - 1, draw v of the
//     type's width, widening multiply v * range; accept when the low half
//     <= zone and return lo + the high half.
class StdRng {
    uint32_t key_[8];
    uint64_t ctr_ = 0;
    uint32_t buf_[16];
    int idx_ = 16;

    static uint32_t rotl(uint32_t x, int k) { return (x << k) | (x >> (32 - k)); }
    static void quarter(uint32_t& a, uint32_t& b, uint32_t& c, uint32_t& d) {
        a += b; d = rotl(d ^ a, 16);
        c += d; b = rotl(b ^ c, 12);
        a += b; d = rotl(d ^ a, 8);
        c += d; b = rotl(b ^ c, 7);
    }
    void block() {
        const uint32_t in[16] = {0x61707865u, 0x3320646eu, 0x79622d32u, 0x6b206574u,
                                 key_[0], key_[1], key_[2], key_[3], key_[4], key_[5], key_[6], key_[7],
                                 (uint32_t)ctr_, (uint32_t)(ctr_ >> 32), 0u, 0u};
        uint32_t x[16];
        std::memcpy(x, in, sizeof x);
        for (int r = 0; r < 12; r += 2) {
            quarter(x[0], x[4], x[8], x[12]);
            quarter(x[1], x[5], x[9], x[13]);
            quarter(x[2], x[6], x[10], x[14]);
            quarter(x[3], x[7], x[11], x[15]);
            quarter(x[0], x[5], x[10], x[15]);
            quarter(x[1], x[6], x[11], x[12]);
            quarter(x[2], x[7], x[8], x[13]);
            quarter(x[3], x[4], x[9], x[14]);
        }
        for (int i = 0; i < 16; i++) buf_[i] = x[i] + in[i];
        ctr_++;
        idx_ = 0;
    }

  public:
    explicit StdRng(uint64_t state) {
        for (int i = 0; i < 8; i++) {
            state = state * 6364136223846793005ull + 11634580027462260723ull;
            const uint32_t xs = (uint32_t)(((state >> 18) ^ state) >> 27);
            const uint32_t rot = (uint32_t)(state >> 59);
            key_[i] = (xs >> rot) | (xs << ((32 - rot) & 31));
        }
    }
    uint32_t next_u32() {
        if (idx_ >= 16) block();
        return buf_[idx_++];
    }
    uint64_t next_u64() {
        const uint64_t lo = next_u32();
        return lo | ((uint64_t)next_u32() << 32);
    }
    uint32_t range_u32(uint32_t lo, uint32_t hi) {
        const uint32_t range = hi - lo + 1u;
        if (range == 0) return next_u32();
        const uint32_t zone = (range << __builtin_clz(range)) - 1u;
        for (;;) {
            const uint64_t m = (uint64_t)next_u32() * range;
            if ((uint32_t)m <= zone) return lo + (uint32_t)(m >> 32);
        }
    }
    int64_t range_i64(int64_t lo, int64_t hi) {
        const uint64_t range = (uint64_t)hi - (uint64_t)lo + 1u;
        if (range == 0) return (int64_t)next_u64();
        const uint64_t zone = (range << __builtin_clzll(range)) - 1u;
        for (;;) {
            const unsigned __int128 m = (unsigned __int128)next_u64() * range;
            if ((uint64_t)m <= zone) return (int64_t)((uint64_t)lo + (uint64_t)(m >> 64));
        }
    }
};

// Parts of [s, e) inside the whitelist intervals of one chromosome
// (src/bin/montecarlo.rs:71-77): from the first interval whose end is past s,
// while its start is before e.  The whitelist is expected sorted and
// non-overlapping, as the reference's own comment says.
template <typename Emit>
void intersect_whitelist(uint32_t s, uint32_t e, const std::vector<Iv>& wl, Emit emit) {
    size_t lo = 0, hi = wl.size();
    while (lo < hi) {
        const size_t mid = (lo + hi) / 2;
        if (wl[mid].end <= s) lo = mid + 1; else hi = mid;
    }
    for (size_t i = lo; i < wl.size() && wl[i].start < e; i++)
        emit(std::max(wl[i].start, s), std::min(wl[i].end, e));
}

// Rust's {:.4e} for a finite f64: mantissa with 4 decimals, then 'e' and the
// exponent without sign for positive powers or leading zeros ("5.0000e-1").
std::string rust_exp4(double v) {
    char b[64];
    std::snprintf(b, sizeof b, "%.4e", v);
    std::string m(b);
    const size_t ep = m.find('e');
    const std::string mant = m.substr(0, ep);
    int ex = std::atoi(m.c_str() + ep + 1);
    return mant + "e" + std::to_string(ex);
}

// src/bin/montecarlo.rs:187-356.  Observed counts come from a sweep of the
// query file; trials are drawn in the reference's order (per trial: the shift
// or, per interval in file order within chromosomes in name order, a start)
// and swept in batches of batch_size trials, one device sweep per batch.
void run_montecarlo(const std::string& db, const std::string& qpath, std::vector<uint64_t> milestones,
                    size_t batch_size, bool have_seed, uint64_t seed, const std::string& wl_path,
                    bool shuffle, const std::string& keep_dir, bool dry_run) {
    if (milestones.empty()) throw Fail("specify at least one trial milestone with --trials");
    if (dry_run && keep_dir.empty()) throw Fail("--dry-run needs --keep-trials");
    std::sort(milestones.begin(), milestones.end());
    milestones.erase(std::unique(milestones.begin(), milestones.end()), milestones.end());
    const uint64_t target = milestones.back();
    batch_size = std::max<size_t>(1, batch_size);
    const auto qbed = bed_by_name(qpath);
    size_t n_ivs = 0;
    for (auto& kv : qbed) n_ivs += kv.second.size();
    std::fprintf(stderr, "Query: %zu intervals across %zu chromosomes\n", n_ivs, qbed.size());
    std::map<std::string, std::vector<Iv>> wl;
    const bool have_wl = !wl_path.empty();
    if (have_wl) {
        wl = bed_by_name(wl_path);
        std::fprintf(stderr, "Whitelist: active \xe2\x80\x94 shifts clipped to annotated regions\n");
    }
    std::fprintf(stderr, "Mode: %s\n", shuffle ? "shuffle (per-interval independent placement)"
                                                : "circular (global shift)");
    std::map<std::string, uint32_t> chrom_size;
    int64_t max_shift = 0;
    {
        // the hg38 table (src/fourier/genome.rs:7-34) by name
        std::vector<std::string> names;
        for (auto& kv : qbed) names.push_back(kv.first);
        const auto g = chk::db::genome_for(names, nullptr);
        for (size_t k = 0; k < names.size(); k++)
            if (g[k]) chrom_size[names[k]] = (uint32_t)g[k];
        const auto all = chk::db::genome_for({"chr1"}, nullptr);
        max_shift = (int64_t)all[0];  // chr1 is the longest hg38 chromosome
    }
    // --dry-run: trials only (written by --keep-trials), no database or device
    Meta meta;
    if (!dry_run) {
        if (!exists(join(db, "meta.json"))) throw Fail("Database not found at " + db);
        meta = Meta::load(db);
    }
    const uint32_t S = (uint32_t)meta.shards.size();
    const uint32_t F = dry_run ? 0u : meta.next_sid();
    std::unordered_map<std::string, uint32_t> shard_of;
    for (uint32_t s = 0; s < S; s++) shard_of[meta.shards[s]] = s;
    IndexHandle ix;
    if (!dry_run) {
        chk(chk_index_open(db.c_str(), 0, &ix.p));
    }
    // count rows of one batch of query files given as (chromosome -> intervals)
    auto sweep_rows = [&](const std::vector<std::vector<std::pair<std::string, Iv>>>& files) {
        std::vector<uint32_t> sh, st, en, qf;
        for (uint32_t r = 0; r < files.size(); r++)
            for (auto& ci : files[r]) {
                auto it = shard_of.find(ci.first);
                sh.push_back(it == shard_of.end() ? S : it->second);
                st.push_back(ci.second.start);
                en.push_back(ci.second.end);
                qf.push_back(r);
            }
        std::vector<uint64_t> cnt(files.size() * (size_t)F), bp(cnt.size());
        chk_query* q = nullptr;
        chk(chk_query_create(ix.p, (uint32_t)files.size(), st.size(), sh.data(), st.data(), en.data(),
                             qf.data(), &q));
        const int rc = chk_sweep(ix.p, q, cnt.data(), bp.data());
        chk_query_free(q);
        chk(rc);
        return cnt;
    };
    // observed counts: the query file alone (an empty file has no row)
    std::vector<uint64_t> observed(F, 0);
    {
        std::vector<std::pair<std::string, Iv>> one;
        for (auto& kv : qbed)
            for (auto& iv : kv.second) one.push_back({kv.first, iv});
        if (!one.empty() && !dry_run) observed = sweep_rows({one});
    }
    std::vector<uint32_t> sids;
    for (auto& kv : meta.sid_map) sids.push_back(kv.first);
    std::sort(sids.begin(), sids.end());
    size_t nz = 0;
    for (uint32_t d : sids) nz += d < F && observed[d] > 0;
    std::fprintf(stderr, "Database: %zu sources; %zu have non-zero overlap with query\n", sids.size(), nz);
    if (!have_seed) seed = ((uint64_t)std::random_device{}() << 32) ^ std::random_device{}();
    StdRng rng(seed);
    std::vector<uint64_t> exceed(F, 0);
    uint64_t total = 0;
    size_t mi = 0, nbatch = 0;
    while (total < target) {
        const uint64_t to_run = std::min<uint64_t>(milestones[mi] - total, batch_size);
        std::vector<std::vector<std::pair<std::string, Iv>>> trials;
        for (uint64_t t = 0; t < to_run; t++) {
            std::vector<std::pair<std::string, Iv>> out;
            const int64_t shift = shuffle ? 0 : rng.range_i64(-max_shift, max_shift);
            for (auto& kv : qbed) {  // chromosomes in name order
                auto cs = chrom_size.find(kv.first);
                if (cs == chrom_size.end()) continue;
                const uint32_t size = cs->second;
                const std::vector<Iv>* w = nullptr;
                if (have_wl) {
                    auto wi = wl.find(kv.first);
                    if (wi != wl.end()) w = &wi->second;
                }
                auto put = [&](uint32_t a, uint32_t b) {
                    if (w) intersect_whitelist(a, b, *w, [&](uint32_t x, uint32_t y) { out.push_back({kv.first, {x, y}}); });
                    else out.push_back({kv.first, {a, b}});
                };
                for (auto& iv : kv.second) {
                    if (shuffle) {  // src/bin/montecarlo.rs:124-163
                        const uint32_t len = iv.end - iv.start;
                        if (len == 0 || len > size) continue;
                        const uint32_t ns = rng.range_u32(0, size - len);
                        put(ns, ns + len);
                    } else {  // src/bin/montecarlo.rs:79-122
                        const int64_t a = (int64_t)iv.start + shift, b = (int64_t)iv.end + shift;
                        if (b <= 0 || a >= (int64_t)size) continue;
                        const uint32_t ta = (uint32_t)std::max<int64_t>(a, 0);
                        const uint32_t tb = (uint32_t)std::min<int64_t>(b, size);
                        if (ta >= tb) continue;
                        put(ta, tb);
                    }
                }
            }
            // a trial that wrote nothing is not queried (its file is removed)
            if (!out.empty()) trials.push_back(std::move(out));
        }
        if (!keep_dir.empty()) {
            const std::string bd = join(keep_dir, "batch_" + std::to_string(nbatch));
            chk::db::mkdirs(bd);
            for (size_t k = 0; k < trials.size(); k++) {
                char nm[32];
                std::snprintf(nm, sizeof nm, "trial_%08zu.bed", k);
                std::string txt;
                for (auto& ci : trials[k])
                    txt += ci.first + "\t" + std::to_string(ci.second.start) + "\t" + std::to_string(ci.second.end) + "\n";
                chk::db::write_file(join(bd, nm), txt.data(), txt.size());
            }
        }
        for (uint32_t d = 0; d < F; d++)
            if (observed[d] == 0) exceed[d] += to_run;  // 0 >= 0 in every trial
        if (!trials.empty() && !dry_run) {
            const auto cnt = sweep_rows(trials);
            for (size_t r = 0; r < trials.size(); r++)
                for (uint32_t d = 0; d < F; d++)
                    if (observed[d] > 0 && cnt[r * F + d] >= observed[d]) exceed[d]++;
        }
        total += to_run;
        nbatch++;
        while (mi < milestones.size() && total >= milestones[mi] && dry_run) mi++;
        while (mi < milestones.size() && total >= milestones[mi]) {
            std::fprintf(stderr, "Reached milestone %llu after %llu trials\n", (unsigned long long)milestones[mi],
                         (unsigned long long)total);
            std::printf("# milestone=%llu trials=%llu\n", (unsigned long long)milestones[mi],
                        (unsigned long long)total);
            std::printf("db_name\tobserved\tp_value\n");
            for (uint32_t d : sids) {
                const uint64_t obs = d < F ? observed[d] : 0;
                const double pv = obs == 0 ? 1.0 : (double)exceed[d] / (double)total;
                std::printf("%s\t%llu\t%s\n", meta.sid_map.at(d).c_str(), (unsigned long long)obs,
                            rust_exp4(pv).c_str());
            }
            std::printf("\n");
            mi++;
        }
        if (mi >= milestones.size()) break;
    }
    std::fflush(stdout);
}

const char* USAGE =
    "Usage: chuckle <COMMAND>\n\n"
    "Commands:\n"
    "  add    Add BED files to the database (creates the database if it does not exist)\n"
    "         -i, --input <DIR|FILE> -d, --db <DB> [--batch-size N] [--stats] [--chroms c...]\n"
    "         [--sorted-sids: D_SIDs in file-name order instead of directory order]\n"
    "  query  Query a database with BED file(s)\n"
    "         -d, --db <DB> -q, --query <FILE|DIR> -o, --output <JSON> [--stats | --intervals] [--batch-size N]\n"
    "  info   Print database summary (shards, sources, layers)\n"
    "         -d, --db <DB>\n"
    "  montecarlo  Monte Carlo shift permutation p-values of overlap counts\n"
    "         -d, --db <DB> -q, --query <FILE> -t, --trials <N,N,...> [-b, --batch-size N] [-s, --seed S]\n"
    "         [-w, --whitelist <BED>] [--mode circular|shuffle] [--keep-trials <DIR> [--dry-run]]\n";

}  // namespace

int main(int argc, char** argv) {
    try {
        if (argc < 2 || !std::strcmp(argv[1], "-h") || !std::strcmp(argv[1], "--help")) {
            std::fputs(USAGE, stdout);
            return argc < 2 ? 2 : 0;
        }
        const std::string cmd = argv[1];
        std::string input, db, query, output, whitelist, mode = "circular", keep;
        size_t batch = 0;
        bool stats = false, intervals = false, have_seed = false, have_batch = false, dry_run = false;
        bool sorted_sids = false;
        uint64_t seed = 0;
        std::vector<uint64_t> trials;
        std::vector<std::string> chroms;
        for (int i = 2; i < argc; i++) {
            const std::string a = argv[i];
            auto val = [&]() -> std::string {
                if (i + 1 >= argc) throw Fail("missing value for " + a);
                return argv[++i];
            };
            if (a == "-i" || a == "--input") input = val();
            else if (a == "-d" || a == "--db") db = val();
            else if (a == "-q" || a == "--query") query = val();
            else if (a == "-o" || a == "--output") output = val();
            else if (a == "--batch-size" || a == "-b") {
                batch = std::strtoull(val().c_str(), nullptr, 10);
                have_batch = true;
            } else if (a == "--dry-run") dry_run = true;
            else if (a == "--sorted-sids") sorted_sids = true;
            else if (a == "-t" || a == "--trials") {
                const std::string v = val();
                for (size_t p0 = 0; p0 <= v.size();) {
                    size_t p1 = v.find(',', p0);
                    if (p1 == std::string::npos) p1 = v.size();
                    if (p1 > p0) trials.push_back(std::strtoull(v.substr(p0, p1 - p0).c_str(), nullptr, 10));
                    p0 = p1 + 1;
                }
            } else if (a == "-s" || a == "--seed") {
                seed = std::strtoull(val().c_str(), nullptr, 10);
                have_seed = true;
            } else if (a == "-w" || a == "--whitelist") whitelist = val();
            else if (a == "--mode") mode = val();
            else if (a == "--keep-trials") keep = val();
            else if (a == "--stats") stats = true;
            else if (a == "--intervals") intervals = true;
            else if (a == "--chroms") {
                while (i + 1 < argc && argv[i + 1][0] != '-') chroms.push_back(argv[++i]);
            } else throw Fail("unexpected argument '" + a + "'");
        }
        if (cmd == "add") {
            if (input.empty() || db.empty()) throw Fail("add needs --input and --db");
            run_add(input, db, batch, stats, chroms, sorted_sids);
        } else if (cmd == "query") {
            if (db.empty() || query.empty() || output.empty()) throw Fail("query needs --db, --query and --output");
            if (stats && intervals) throw Fail("--stats cannot be used with --intervals");
            run_query(db, query, output, stats, intervals, batch);
        } else if (cmd == "info") {
            if (db.empty()) throw Fail("info needs --db");
            run_info(db);
        } else if (cmd == "montecarlo") {
            if (db.empty() || query.empty()) throw Fail("montecarlo needs --db and --query");
            if (mode != "circular" && mode != "shuffle") throw Fail("--mode must be circular or shuffle");
            run_montecarlo(db, query, trials, have_batch ? batch : 100, have_seed, seed, whitelist,
                           mode == "shuffle", keep, dry_run);
        } else {
            throw Fail("unknown command '" + cmd + "'");
        }
    } catch (const std::exception& e) {
        std::fprintf(stderr, "Error: %s\n", e.what());
        return 1;
    }
    return 0;
}
