// dbfile.hpp -- the reference's on-disk database and BED formats on the
// host side of the library: meta.json, layer files and jump tables, BED(.gz)
// parsing over host threads, the moment-table file of this build and the
// StatsOutput JSON writer.
#pragma once

#include <cstdint>
#include <functional>
#include <map>
#include <set>
#include <string>
#include <vector>

#include "common.hpp"

namespace chk {
struct Index;

namespace db {

// ------------------------------------------------------------------ files
bool exists(const std::string& p);
bool is_dir(const std::string& p);
void mkdirs(const std::string& p);
std::string join(const std::string& a, const std::string& b);
std::string basename_of(const std::string& p);
// file name ending in .bed or .bed.gz (src/io/parse.rs:16-19)
bool is_bed_file(const std::string& p);
// Entries of a directory in readdir order (the reference's read_dir order).
std::vector<std::string> list_dir(const std::string& d);
std::string read_file(const std::string& p);
void write_file(const std::string& p, const void* data, size_t n);
// Host threads for parallel host work: hardware threads, at most 16.
unsigned host_threads();
// Runs f(t) for every t in [0, T) on the library's persistent host workers
// (host_threads() - 1 of them, started on first use) and the calling thread.
// The caller also runs every task no worker has taken yet, so concurrent
// calls from several threads, and calls from inside a task, always finish.
// After a task throws, tasks not yet started are skipped and the first
// exception is rethrown to the caller.
void host_run(unsigned T, const std::function<void(unsigned)>& f);

// ------------------------------------------------------------------ BED
struct Iv {
    uint32_t start, end;
};

// One parsed BED file: intervals per shard name in first-seen shard order,
// each shard's intervals in file order (src/io/parse.rs:53-141: '#' and
// blank lines skipped, first non-integer column = shard ("default" when
// none), first two integer columns = start, end, start < end).
struct BedFile {
    std::vector<std::string> shards;
    std::vector<std::vector<Iv>> ivs;
    uint64_t n = 0;
};
BedFile parse_bed(const std::string& path, unsigned threads);
// Many files over host threads (one file per thread, large files split
// across threads); results in the order of `paths`.
std::vector<BedFile> parse_beds(const std::vector<std::string>& paths, unsigned threads);

// ------------------------------------------------------------------ JSON
struct JVal {
    enum Kind { Null, Bool, Num, Str, Arr, Obj } kind = Null;
    double num = 0;
    std::string str;
    std::vector<JVal> arr;
    std::vector<std::pair<std::string, JVal>> obj;
    const JVal* get(const std::string& k) const;
};
JVal parse_json(const std::string& text, const std::string& what);
std::string json_escape(const std::string& s);
// Shortest round-trip decimal in serde_json's (ryu's) layout; null if not finite.
std::string json_f64(double v);

// ------------------------------------------------------------------ meta
struct LayerConfig {
    uint32_t min_size = 128, growth_factor = 4;
    uint32_t layer_for_size(uint32_t size) const;  // src/io/meta.rs:33-61
    uint32_t layer_max_size(uint32_t layer) const;  // src/io/meta.rs:66-72
    uint32_t tile_size(uint32_t layer) const;       // src/io/meta.rs:78-80
};

// DB/meta.json (src/io/meta.rs:101-145).
struct Meta {
    std::map<uint32_t, std::string> sid_map;
    LayerConfig lc;
    std::vector<std::string> shards;
    uint32_t num_layers = 0;
    static Meta load(const std::string& db);
    void save(const std::string& db) const;  // serde_json pretty layout
    uint32_t next_sid() const { return sid_map.empty() ? 0 : sid_map.rbegin()->first + 1; }
};

struct Pos {
    uint32_t start, end;
};
std::string layer_path(const std::string& db, const std::string& shard, uint32_t k, const char* ext);
// src/io/jump.rs:66-113: table[t] = #{start < t * tile}, t = 0..=max_start/tile;
// an empty layer gets [0].
std::vector<uint64_t> jump_table(const Pos* p, size_t n, uint32_t tile);

// Every interval of a database as flat arrays (shard index in meta.shards
// order, start, end, D_SID); layer files read over host threads.
struct Flat {
    std::vector<uint32_t> shard, start, end, sid;
};
Flat load_intervals(const std::string& db, const Meta& m);

// hg38 lengths of the shards (src/fourier/genome.rs:7-34), 0 for shards
// outside the table or outside the whitelist (when given).
std::vector<uint64_t> genome_for(const std::vector<std::string>& shards,
                                 const std::set<std::string>* whitelist);

// DB/momentmap.chk (this build's moment store).
void save_moments(const std::string& db, Index& ix, const Meta& m);
// Loads the tables when the file matches the index (file, shard and interval
// counts); returns false when it is absent or stale.
bool load_moments(const std::string& db, Index& ix, const Meta& m);

// src/tasks/build.rs:86-276 (add_to_database): parse each batch of BED
// files, bucket by (shard, size layer), sort every bucket by start on the
// device, merge into the existing layer files (existing intervals first
// among equal starts, src/io/layer.rs:135-199), write the jump tables and
// meta.json.  With sorted_sids the files take D_SIDs in path order instead
// of the directory's order.  With stats, the moment tables of every file are
// rebuilt on the device and saved.
void add(const std::string& input, const std::string& db, uint64_t batch_size, bool stats,
         const std::set<std::string>& chroms, bool sorted_sids, int device);

// StatsOutput JSON (src/stats/mod.rs:31-60) of a result matrix, sorted as
// src/main.rs:194-201 does: by p ascending in stats mode (records only for
// cells with flags bit 0), else by count descending (nonzero cells).
void write_results(const std::string& path, const std::vector<std::string>& qnames,
                   const std::vector<std::string>& dnames, const uint64_t* counts, const uint64_t* bp,
                   const double* p, const double* llr, const uint8_t* flags, bool stats);

}  // namespace db
}  // namespace chk
