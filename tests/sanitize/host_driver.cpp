// host_driver.cpp -- exercises the library's host-side format code (BED
// parsing over host threads, meta.json, jump tables, result JSON; dbfile.cpp)
// in a CPU build with AddressSanitizer and UndefinedBehaviorSanitizer.
// Usage: host_driver <work dir> <bed file>...
//   every file is parsed with 1 and with 8 threads (the results must agree);
//   a meta.json with the files' shards is saved, loaded and compared; a jump
//   table is checked against its definition; a result matrix is written as
//   JSON in both modes.  Prints one line per file and "ok"; exits 1 on a
//   mismatch (parse errors are printed and count as handled).
#include <algorithm>
#include <cstdio>
#include <string>
#include <vector>

#include "dbfile.hpp"

namespace chk {
// the moment-table file functions are not exercised here
void moments_download(Index&, int64_t*, uint32_t*, double*, double*) { throw Error(CHK_ERR_INTERNAL, "unused"); }
void moments_upload(Index&, const int64_t*, const uint32_t*, uint64_t, const double*, const double*) {
    throw Error(CHK_ERR_INTERNAL, "unused");
}
void index_init(Index&, int, uint32_t, uint32_t, uint64_t) { throw Error(CHK_ERR_INTERNAL, "unused"); }
}  // namespace chk

using namespace chk::db;

static int fail(const std::string& m) {
    std::fprintf(stderr, "mismatch: %s\n", m.c_str());
    return 1;
}

int main(int argc, char** argv) {
    if (argc < 3) return 2;
    const std::string dir = argv[1];
    Meta meta;
    for (int k = 2; k < argc; k++) {
        try {
            const BedFile a = parse_bed(argv[k], 1), b = parse_bed(argv[k], 8);
            if (a.n != b.n || a.shards != b.shards) return fail(std::string("threads differ: ") + argv[k]);
            for (size_t s = 0; s < a.shards.size(); s++) {
                if (a.ivs[s].size() != b.ivs[s].size()) return fail("shard sizes");
                for (size_t i = 0; i < a.ivs[s].size(); i++)
                    if (a.ivs[s][i].start != b.ivs[s][i].start || a.ivs[s][i].end != b.ivs[s][i].end)
                        return fail("intervals");
                bool seen = false;
                for (auto& x : meta.shards) seen |= x == a.shards[s];
                if (!seen) meta.shards.push_back(a.shards[s]);
                // jump table of the shard's starts (sorted) against its definition
                std::vector<Pos> p;
                for (auto& iv : a.ivs[s]) p.push_back({iv.start, iv.end});
                std::sort(p.begin(), p.end(), [](const Pos& x, const Pos& y) { return x.start < y.start; });
                const uint32_t tile = 1000;
                const auto jt = jump_table(p.data(), p.size(), tile);
                size_t c = 0;  // #starts below t * tile, t ascending
                for (size_t t = 0; t < jt.size(); t++) {
                    while (c < p.size() && p[c].start < t * (uint64_t)tile) c++;
                    if (jt[t] != c) return fail("jump table");
                }
                if (jt.size() != (p.empty() ? 1 : p.back().start / tile + 1)) return fail("jump table size");
            }
            meta.sid_map[meta.next_sid()] = basename_of(argv[k]);
            std::printf("%s: %llu intervals, %zu shards\n", basename_of(argv[k]).c_str(),
                        (unsigned long long)a.n, a.shards.size());
        } catch (const chk::Error& e) {
            std::printf("%s: error: %s\n", basename_of(argv[k]).c_str(), e.what());
        }
    }
    meta.num_layers = 3;
    mkdirs(dir);
    meta.save(dir);
    const Meta m2 = Meta::load(dir);
    if (m2.sid_map != meta.sid_map || m2.shards != meta.shards || m2.num_layers != 3) return fail("meta");
    const std::vector<std::string> q = {"q\"1.bed", "q2.bed"}, d = {"a.bed", "b\\x.bed", "c.bed"};
    const uint64_t cnt[6] = {3, 0, 7, 1, 2, 0}, bp[6] = {150, 0, 3500, 1, 20, 0};
    const double p[6] = {0.5, 1, 1e-300, 0.25, 0.75, 1}, llr[6] = {1, 0, 700, 2, 0.1, 0};
    const uint8_t fl[6] = {3, 0, 3, 1, 3, 0};
    write_results(join(dir, "counts.json"), q, d, cnt, bp, p, llr, fl, false);
    write_results(join(dir, "stats.json"), q, d, cnt, bp, p, llr, fl, true);
    const JVal j1 = parse_json(read_file(join(dir, "stats.json")), "stats.json");
    if (!j1.get("results") || j1.get("results")->arr.size() != 4) return fail("stats json");
    std::printf("ok\n");
    return 0;
}
