// cli_chuckle.cpp -- the `chuckle` command line (add | query | info) over the
// MI355X library, a drop-in for the reference's CLI (src/main.rs:17-78) on
// the reference's on-disk database format:
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
// Output JSON follows StatsOutput (src/stats/mod.rs:31-60) pretty-printed
// with two-space indentation, sorted as src/main.rs:146-211 does.
//
// `chuckle montecarlo` is the reference's Monte-Carlo shift permutation
// driver (src/bin/montecarlo.rs) on the batched sweep: trials are generated in
// memory and each batch of them is one multi-file sweep on the device.
#include <dirent.h>
#include <sys/stat.h>
#include <zlib.h>

#include <algorithm>
#include <cerrno>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <random>
#include <set>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <utility>
#include <vector>

#include "chuckle_hip.h"

namespace {

struct Fail : std::runtime_error {
    using std::runtime_error::runtime_error;
};

void chk(int code) {
    if (code != CHK_OK) throw Fail(chk_last_error());
}

// ------------------------------------------------------------------ files

bool exists(const std::string& p) {
    struct stat st;
    return ::stat(p.c_str(), &st) == 0;
}

bool is_dir(const std::string& p) {
    struct stat st;
    return ::stat(p.c_str(), &st) == 0 && S_ISDIR(st.st_mode);
}

void mkdirs(const std::string& p) {
    std::string cur;
    for (size_t i = 0; i <= p.size(); i++) {
        if (i == p.size() || p[i] == '/') {
            if (!cur.empty() && !exists(cur) && ::mkdir(cur.c_str(), 0755) != 0 && errno != EEXIST)
                throw Fail("cannot create directory " + cur + ": " + std::strerror(errno));
        }
        if (i < p.size()) cur += p[i];
    }
}

std::string join(const std::string& a, const std::string& b) {
    return a.empty() || a.back() == '/' ? a + b : a + "/" + b;
}

std::string basename_of(const std::string& p) {
    const size_t k = p.find_last_of('/');
    return k == std::string::npos ? p : p.substr(k + 1);
}

bool ends_with(const std::string& s, const char* suf) {
    const size_t n = std::strlen(suf);
    return s.size() >= n && s.compare(s.size() - n, n, suf) == 0;
}

// src/io/parse.rs:16-19
bool is_bed_file(const std::string& p) {
    const std::string n = basename_of(p);
    return ends_with(n, ".bed") || ends_with(n, ".bed.gz");
}

// Directory entries in readdir order (the reference's read_dir order).
std::vector<std::string> list_dir(const std::string& d) {
    std::vector<std::string> out;
    DIR* dp = ::opendir(d.c_str());
    if (!dp) throw Fail("cannot read directory " + d + ": " + std::strerror(errno));
    while (struct dirent* e = ::readdir(dp)) {
        if (!std::strcmp(e->d_name, ".") || !std::strcmp(e->d_name, "..")) continue;
        out.push_back(join(d, e->d_name));
    }
    ::closedir(dp);
    return out;
}

std::string read_text(const std::string& p) {
    FILE* f = std::fopen(p.c_str(), "rb");
    if (!f) throw Fail("IO error: cannot open " + p + ": " + std::strerror(errno));
    std::string s;
    char buf[1 << 16];
    size_t k;
    while ((k = std::fread(buf, 1, sizeof buf, f)) > 0) s.append(buf, k);
    std::fclose(f);
    return s;
}

void write_bytes(const std::string& p, const void* data, size_t n) {
    FILE* f = std::fopen(p.c_str(), "wb");
    if (!f) throw Fail("IO error: cannot write " + p + ": " + std::strerror(errno));
    if (n && std::fwrite(data, 1, n, f) != n) {
        std::fclose(f);
        throw Fail("IO error: short write to " + p);
    }
    std::fclose(f);
}

template <typename T>
std::vector<T> read_array(const std::string& p) {
    if (!exists(p)) return {};
    const std::string s = read_text(p);
    if (s.size() % sizeof(T)) throw Fail(p + " has invalid size: " + std::to_string(s.size()) + " bytes");
    std::vector<T> v(s.size() / sizeof(T));
    if (!v.empty()) std::memcpy(v.data(), s.data(), s.size());
    return v;
}

// ------------------------------------------------------------------ BED

struct Iv {
    uint32_t start, end, sid;
};

// A token parses as u32 when it is an optional '+' then decimal digits whose
// value fits (Rust's u32::from_str).
bool parse_u32(const char* b, const char* e, uint32_t* out) {
    if (b < e && *b == '+') b++;
    if (b == e) return false;
    uint64_t v = 0;
    for (const char* p = b; p < e; p++) {
        if (*p < '0' || *p > '9') return false;
        v = v * 10 + (uint64_t)(*p - '0');
        if (v > 0xFFFFFFFFull) return false;
    }
    *out = (uint32_t)v;
    return true;
}

// src/io/parse.rs:53-141: first non-integer column = shard, first two integer
// columns = start, end; '#' and blank lines skipped; "default" shard when no
// non-integer column; start < end required.
std::map<std::string, std::vector<Iv>> parse_bed_file(const std::string& path, uint32_t sid) {
    gzFile gz = gzopen(path.c_str(), "rb");  // reads plain files as they are
    if (!gz) throw Fail("IO error: cannot open " + path);
    std::map<std::string, std::vector<Iv>> shards;
    std::string line;
    std::vector<char> buf(1 << 16);
    size_t line_num = 0;
    bool more = true;
    while (more) {
        line.clear();
        for (;;) {
            if (!gzgets(gz, buf.data(), (int)buf.size())) {
                more = false;
                break;
            }
            line += buf.data();
            if (!line.empty() && line.back() == '\n') break;
        }
        if (line.empty() && !more) break;
        line_num++;
        while (!line.empty() && (line.back() == '\n' || line.back() == '\r')) line.pop_back();
        const char* p = line.c_str();
        const char* e = p + line.size();
        while (p < e && std::isspace((unsigned char)*p)) p++;
        while (e > p && std::isspace((unsigned char)e[-1])) e--;
        if (p == e || *p == '#') continue;
        std::string shard;
        bool have_shard = false, have_s = false, have_e = false;
        uint32_t s = 0, en = 0;
        int cols = 0;
        while (p < e) {
            while (p < e && std::isspace((unsigned char)*p)) p++;
            if (p == e) break;
            const char* t = p;
            while (p < e && !std::isspace((unsigned char)*p)) p++;
            cols++;
            uint32_t v;
            if (parse_u32(t, p, &v)) {
                if (!have_s) { s = v; have_s = true; }
                else if (!have_e) { en = v; have_e = true; }
            } else if (!have_shard) {
                shard.assign(t, p);
                have_shard = true;
            }
            if (have_shard && have_e) break;
        }
        if (cols < 2 || !have_s || !have_e) {
            gzclose(gz);
            throw Fail("Line " + std::to_string(line_num) + ": invalid format - expected at least 2 integer columns");
        }
        if (s >= en) {
            gzclose(gz);
            throw Fail("Line " + std::to_string(line_num) + ": start (" + std::to_string(s) +
                       ") must be less than end (" + std::to_string(en) + ")");
        }
        shards[have_shard ? shard : std::string("default")].push_back({s, en, sid});
    }
    gzclose(gz);
    return shards;
}

// ------------------------------------------------------------------ JSON

struct JVal {
    enum Kind { Null, Bool, Num, Str, Arr, Obj } kind = Null;
    double num = 0;
    std::string str;
    std::vector<JVal> arr;
    std::vector<std::pair<std::string, JVal>> obj;
    const JVal* get(const std::string& k) const {
        for (auto& kv : obj)
            if (kv.first == k) return &kv.second;
        return nullptr;
    }
};

struct JParser {
    const std::string& s;
    size_t i = 0;
    explicit JParser(const std::string& x) : s(x) {}
    void ws() {
        while (i < s.size() && std::isspace((unsigned char)s[i])) i++;
    }
    [[noreturn]] void bad() { throw Fail("JSON error: malformed meta.json at byte " + std::to_string(i)); }
    std::string string() {
        if (s[i] != '"') bad();
        i++;
        std::string out;
        while (i < s.size() && s[i] != '"') {
            if (s[i] == '\\') {
                i++;
                if (i >= s.size()) bad();
                const char c = s[i++];
                switch (c) {
                    case 'n': out += '\n'; break;
                    case 't': out += '\t'; break;
                    case 'r': out += '\r'; break;
                    case 'b': out += '\b'; break;
                    case 'f': out += '\f'; break;
                    case 'u': {
                        if (i + 4 > s.size()) bad();
                        unsigned cp = (unsigned)std::strtoul(s.substr(i, 4).c_str(), nullptr, 16);
                        i += 4;
                        if (cp >= 0xD800 && cp < 0xDC00 && i + 6 <= s.size() && s[i] == '\\' && s[i + 1] == 'u') {
                            unsigned lo = (unsigned)std::strtoul(s.substr(i + 2, 4).c_str(), nullptr, 16);
                            i += 6;
                            cp = 0x10000 + ((cp - 0xD800) << 10) + (lo - 0xDC00);
                        }
                        if (cp < 0x80) out += (char)cp;
                        else if (cp < 0x800) { out += (char)(0xC0 | (cp >> 6)); out += (char)(0x80 | (cp & 63)); }
                        else if (cp < 0x10000) { out += (char)(0xE0 | (cp >> 12)); out += (char)(0x80 | ((cp >> 6) & 63)); out += (char)(0x80 | (cp & 63)); }
                        else { out += (char)(0xF0 | (cp >> 18)); out += (char)(0x80 | ((cp >> 12) & 63)); out += (char)(0x80 | ((cp >> 6) & 63)); out += (char)(0x80 | (cp & 63)); }
                        break;
                    }
                    default: out += c;
                }
            } else {
                out += s[i++];
            }
        }
        if (i >= s.size()) bad();
        i++;
        return out;
    }
    JVal value() {
        ws();
        if (i >= s.size()) bad();
        JVal v;
        const char c = s[i];
        if (c == '{') {
            v.kind = JVal::Obj;
            i++;
            ws();
            if (s[i] == '}') { i++; return v; }
            for (;;) {
                ws();
                std::string k = string();
                ws();
                if (s[i] != ':') bad();
                i++;
                v.obj.emplace_back(k, value());
                ws();
                if (s[i] == ',') { i++; continue; }
                if (s[i] == '}') { i++; break; }
                bad();
            }
        } else if (c == '[') {
            v.kind = JVal::Arr;
            i++;
            ws();
            if (s[i] == ']') { i++; return v; }
            for (;;) {
                v.arr.push_back(value());
                ws();
                if (s[i] == ',') { i++; continue; }
                if (s[i] == ']') { i++; break; }
                bad();
            }
        } else if (c == '"') {
            v.kind = JVal::Str;
            v.str = string();
        } else if (!s.compare(i, 4, "true") || !s.compare(i, 5, "false")) {
            v.kind = JVal::Bool;
            v.num = s[i] == 't';
            i += s[i] == 't' ? 4 : 5;
        } else if (!s.compare(i, 4, "null")) {
            i += 4;
        } else {
            char* end = nullptr;
            v.kind = JVal::Num;
            v.num = std::strtod(s.c_str() + i, &end);
            if (end == s.c_str() + i) bad();
            i = (size_t)(end - s.c_str());
        }
        return v;
    }
};

std::string json_escape(const std::string& s) {
    std::string o = "\"";
    for (unsigned char c : s) {
        switch (c) {
            case '"': o += "\\\""; break;
            case '\\': o += "\\\\"; break;
            case '\n': o += "\\n"; break;
            case '\r': o += "\\r"; break;
            case '\t': o += "\\t"; break;
            case '\b': o += "\\b"; break;
            case '\f': o += "\\f"; break;
            default:
                if (c < 0x20) {
                    char b[8];
                    std::snprintf(b, sizeof b, "\\u%04x", c);
                    o += b;
                } else {
                    o += (char)c;
                }
        }
    }
    return o + "\"";
}

// Shortest round-trip decimal of a finite double in the layout serde_json
// writes (ryu's: plain notation with a ".0" when the decimal point falls
// inside [-5, 16] digits, else d.ddde±x); non-finite values print as null.
std::string json_f64(double v) {
    if (!std::isfinite(v)) return "null";
    if (v == 0) return std::signbit(v) ? "-0.0" : "0.0";
    char buf[40];
    int prec = 1;
    for (; prec <= 17; prec++) {
        std::snprintf(buf, sizeof buf, "%.*e", prec - 1, v);
        if (std::strtod(buf, nullptr) == v) break;
    }
    // buf = [-]d.ddde[+-]xx : digits and decimal exponent
    std::string s(buf);
    const bool neg = s[0] == '-';
    if (neg) s.erase(0, 1);
    const size_t epos = s.find('e');
    const int exp10 = std::atoi(s.c_str() + epos + 1);
    std::string digits;
    for (size_t k = 0; k < epos; k++)
        if (s[k] != '.') digits += s[k];
    while (digits.size() > 1 && digits.back() == '0') digits.pop_back();
    const int len = (int)digits.size();
    const int kk = exp10 + 1;  // position of the decimal point from the left
    std::string out;
    if (kk >= len && kk <= 16) {
        out = digits + std::string(kk - len, '0') + ".0";
    } else if (kk > 0 && kk <= 16) {
        out = digits.substr(0, kk) + "." + digits.substr(kk);
    } else if (kk > -5 && kk <= 0) {
        out = "0." + std::string(-kk, '0') + digits;
    } else {
        out = digits.substr(0, 1);
        if (len > 1) out += "." + digits.substr(1);
        out += "e" + std::to_string(kk - 1);
    }
    return neg ? "-" + out : out;
}

// ------------------------------------------------------------------ meta

struct LayerConfig {
    uint32_t min_size = 128, growth_factor = 4;
    // src/io/meta.rs:33-61
    uint32_t layer_for_size(uint32_t size) const {
        if (size < min_size) return 0;
        const bool p2 = (min_size & (min_size - 1)) == 0 && (growth_factor & (growth_factor - 1)) == 0 &&
                        growth_factor >= 2;
        if (p2) {
            const uint32_t l2s = 31 - __builtin_clz(size), l2m = 31 - __builtin_clz(min_size);
            const uint32_t l2g = __builtin_ctz(growth_factor);
            return (l2s > l2m ? l2s - l2m : 0) / l2g + 1;
        }
        uint64_t th = min_size;
        uint32_t layer = 0;
        while (th <= size) {
            th = std::min<uint64_t>(th * growth_factor, UINT64_MAX / 2);
            layer++;
        }
        return layer;
    }
    // src/io/meta.rs:66-81
    uint32_t layer_max_size(uint32_t layer) const {
        uint64_t v = min_size;
        for (uint32_t k = 0; k < layer; k++) v = std::min<uint64_t>(v * growth_factor, 0xFFFFFFFFull);
        return (uint32_t)v;
    }
    uint32_t tile_size(uint32_t layer) const { return std::max<uint32_t>(1, layer_max_size(layer) / 2); }
};

struct Meta {
    std::map<uint32_t, std::string> sid_map;
    LayerConfig lc;
    std::vector<std::string> shards;
    uint32_t num_layers = 0;

    static Meta load(const std::string& db) {
        const JVal root = JParser(read_text(join(db, "meta.json"))).value();
        Meta m;
        if (const JVal* sm = root.get("sid_map"))
            for (auto& kv : sm->obj) {
                const JVal* name = kv.second.get("name");
                m.sid_map[(uint32_t)std::strtoul(kv.first.c_str(), nullptr, 10)] = name ? name->str : "";
            }
        if (const JVal* lc = root.get("layer_config")) {
            if (const JVal* a = lc->get("min_size")) m.lc.min_size = (uint32_t)a->num;
            if (const JVal* b = lc->get("growth_factor")) m.lc.growth_factor = (uint32_t)b->num;
        }
        if (const JVal* sh = root.get("shards"))
            for (auto& v : sh->arr) m.shards.push_back(v.str);
        if (const JVal* nl = root.get("num_layers")) m.num_layers = (uint32_t)nl->num;
        return m;
    }

    // serde_json::to_string_pretty layout of src/io/meta.rs:101-110
    void save(const std::string& db) const {
        std::string o = "{\n  \"sid_map\": {";
        bool first = true;
        for (auto& kv : sid_map) {
            o += first ? "\n" : ",\n";
            first = false;
            o += "    \"" + std::to_string(kv.first) + "\": {\n      \"name\": " + json_escape(kv.second) + "\n    }";
        }
        o += sid_map.empty() ? "},\n" : "\n  },\n";
        o += "  \"layer_config\": {\n    \"min_size\": " + std::to_string(lc.min_size) +
             ",\n    \"growth_factor\": " + std::to_string(lc.growth_factor) + "\n  },\n";
        o += "  \"shards\": [";
        for (size_t k = 0; k < shards.size(); k++) o += (k ? ",\n    " : "\n    ") + json_escape(shards[k]);
        o += shards.empty() ? "],\n" : "\n  ],\n";
        o += "  \"num_layers\": " + std::to_string(num_layers) + "\n}";
        write_bytes(join(db, "meta.json"), o.data(), o.size());
    }

    uint32_t next_sid() const { return sid_map.empty() ? 0 : sid_map.rbegin()->first + 1; }
};

// ------------------------------------------------------------------ layers

struct Pos {
    uint32_t start, end;
};

std::string layer_path(const std::string& db, const std::string& shard, uint32_t k, const char* ext) {
    return join(join(join(db, "shards"), shard), "layer_" + std::to_string(k) + "." + ext);
}

// src/io/jump.rs:66-113: table[t] = #{start < t * tile}, t = 0..=max_start/tile;
// an empty layer gets [0].
std::vector<uint64_t> jump_table(const std::vector<Pos>& p, uint32_t tile) {
    if (p.empty()) return {0};
    const uint64_t max_tile = p.back().start / tile;
    std::vector<uint64_t> t(max_tile + 1, 0);
    size_t i = 0;
    for (uint64_t k = 0; k <= max_tile; k++) {
        const uint64_t lim = k * (uint64_t)tile;
        while (i < p.size() && p[i].start < lim) i++;
        t[k] = i;
    }
    return t;
}

// All intervals of a database as flat arrays (shard index, start, end, sid).
struct Flat {
    std::vector<uint32_t> shard, start, end, sid;
};

Flat load_db_intervals(const std::string& db, const Meta& m) {
    Flat f;
    for (uint32_t s = 0; s < m.shards.size(); s++)
        for (uint32_t k = 0; k < m.num_layers; k++) {
            const auto pos = read_array<Pos>(layer_path(db, m.shards[s], k, "pos"));
            const auto sid = read_array<uint32_t>(layer_path(db, m.shards[s], k, "sid"));
            if (pos.size() != sid.size())
                throw Fail("Position count " + std::to_string(pos.size()) + " does not match sid count " +
                           std::to_string(sid.size()));
            for (size_t i = 0; i < pos.size(); i++) {
                f.shard.push_back(s);
                f.start.push_back(pos[i].start);
                f.end.push_back(pos[i].end);
                f.sid.push_back(sid[i]);
            }
        }
    return f;
}

// ------------------------------------------------------------------ moments

const std::pair<const char*, uint64_t> HG38[] = {
    {"chr1", 248956422}, {"chr2", 242193529}, {"chr3", 198295559}, {"chr4", 190214555},
    {"chr5", 181538259}, {"chr6", 170805979}, {"chr7", 159345973}, {"chr8", 145138636},
    {"chr9", 138394717}, {"chr10", 133797422}, {"chr11", 135086622}, {"chr12", 133275309},
    {"chr13", 114364328}, {"chr14", 107043718}, {"chr15", 101991189}, {"chr16", 90338345},
    {"chr17", 83257441}, {"chr18", 80373285}, {"chr19", 58617616}, {"chr20", 64444167},
    {"chr21", 46709983}, {"chr22", 50818468}, {"chrX", 156040895}, {"chrY", 57227415}};

// Per-shard genome lengths: the hg38 table (src/fourier/genome.rs:7-34), 0
// for shards outside it or outside the whitelist.
std::vector<uint64_t> genome_for(const std::vector<std::string>& shards, const std::set<std::string>* wl) {
    std::vector<uint64_t> g(shards.size(), 0);
    for (size_t s = 0; s < shards.size(); s++)
        for (auto& e : HG38)
            if (shards[s] == e.first && (!wl || wl->count(shards[s]))) g[s] = e.second;
    return g;
}

// DB/momentmap.chk: "CHKMOM01", u32 nfiles, u32 nshards, u64 ndb_intervals,
// u64 total, then per shard (u32 name length, bytes), i64 slot_off[F*S],
// u32 slot_len[F*S], f64 mean[total], f64 var[total].
void save_moments(const std::string& db, const Meta& m, uint32_t F, uint64_t n, chk_index* ix) {
    uint64_t total = 0;
    chk(chk_moments_download(ix, &total, nullptr, nullptr, nullptr, nullptr));
    const size_t ns = (size_t)F * m.shards.size();
    std::vector<int64_t> so(ns);
    std::vector<uint32_t> sl(ns);
    std::vector<double> mean(std::max<uint64_t>(total, 1)), var(std::max<uint64_t>(total, 1));
    chk(chk_moments_download(ix, &total, so.data(), sl.data(), mean.data(), var.data()));
    std::string o = "CHKMOM01";
    auto put = [&](const void* p, size_t k) { o.append((const char*)p, k); };
    const uint32_t S = (uint32_t)m.shards.size();
    put(&F, 4);
    put(&S, 4);
    put(&n, 8);
    put(&total, 8);
    for (auto& s : m.shards) {
        const uint32_t L = (uint32_t)s.size();
        put(&L, 4);
        put(s.data(), L);
    }
    put(so.data(), ns * 8);
    put(sl.data(), ns * 4);
    put(mean.data(), total * 8);
    put(var.data(), total * 8);
    write_bytes(join(db, "momentmap.chk"), o.data(), o.size());
}

bool load_moments(const std::string& db, const Meta& m, uint32_t F, uint64_t n, chk_index* ix) {
    const std::string p = join(db, "momentmap.chk");
    if (!exists(p)) return false;
    const std::string s = read_text(p);
    size_t i = 8;
    auto get = [&](void* d, size_t k) {
        if (i + k > s.size()) throw Fail("truncated " + p);
        std::memcpy(d, s.data() + i, k);
        i += k;
    };
    if (s.compare(0, 8, "CHKMOM01")) return false;
    uint32_t F2, S2;
    uint64_t n2, total;
    get(&F2, 4);
    get(&S2, 4);
    get(&n2, 8);
    get(&total, 8);
    if (F2 != F || S2 != m.shards.size() || n2 != n) return false;  // stale
    for (uint32_t k = 0; k < S2; k++) {
        uint32_t L;
        get(&L, 4);
        std::string name(L, '\0');
        get(&name[0], L);
        if (name != m.shards[k]) return false;
    }
    const size_t ns = (size_t)F * S2;
    std::vector<int64_t> so(ns);
    std::vector<uint32_t> sl(ns);
    std::vector<double> mean(std::max<uint64_t>(total, 1)), var(std::max<uint64_t>(total, 1));
    get(so.data(), ns * 8);
    get(sl.data(), ns * 4);
    get(mean.data(), total * 8);
    get(var.data(), total * 8);
    chk(chk_moments_upload(ix, so.data(), sl.data(), total, mean.data(), var.data()));
    return true;
}

struct IndexHandle {
    chk_index* p = nullptr;
    ~IndexHandle() {
        if (p) chk_index_free(p);
    }
};

// ------------------------------------------------------------------ add

// src/tasks/build.rs:86-276: parse -> bucket by (shard, layer) -> sort by
// start -> write or merge the layer and its jump table -> update meta.json.
void run_add(const std::string& input, const std::string& db, size_t batch_size, bool stats,
             const std::vector<std::string>& chroms) {
    if (!exists(input)) throw Fail("Input path not found: " + input);
    std::vector<std::string> beds;
    if (is_dir(input)) {
        for (auto& p : list_dir(input))
            if (is_bed_file(p)) beds.push_back(p);
    } else if (is_bed_file(input)) {
        beds.push_back(input);
    }
    if (beds.empty()) throw Fail("No BED files found in " + input);
    mkdirs(db);
    Meta meta;
    if (exists(join(db, "meta.json"))) meta = Meta::load(db);
    const size_t bs = std::max<size_t>(1, batch_size ? batch_size : beds.size());
    for (size_t b0 = 0; b0 < beds.size(); b0 += bs) {
        const uint32_t next = meta.next_sid();
        std::map<std::pair<std::string, uint32_t>, std::vector<Iv>> buckets;
        for (size_t i = b0; i < std::min(beds.size(), b0 + bs); i++) {
            const uint32_t sid = next + (uint32_t)(i - b0);
            auto shards = parse_bed_file(beds[i], sid);
            meta.sid_map[sid] = basename_of(beds[i]);
            for (auto& kv : shards)
                for (auto& iv : kv.second)
                    buckets[{kv.first, meta.lc.layer_for_size(iv.end - iv.start)}].push_back(iv);
        }
        uint32_t max_layer = meta.num_layers ? meta.num_layers - 1 : 0;
        for (auto& kv : buckets) {
            auto& ivs = kv.second;
            if (ivs.empty()) continue;
            std::stable_sort(ivs.begin(), ivs.end(), [](const Iv& a, const Iv& c) { return a.start < c.start; });
            const std::string& shard = kv.first.first;
            const uint32_t layer = kv.first.second;
            mkdirs(join(join(db, "shards"), shard));
            const std::string pp = layer_path(db, shard, layer, "pos"), sp = layer_path(db, shard, layer, "sid");
            std::vector<Pos> pos = read_array<Pos>(pp);
            std::vector<uint32_t> sid = read_array<uint32_t>(sp);
            if (pos.size() != sid.size()) throw Fail("Position count does not match sid count in " + pp);
            // merge: existing intervals first among equal starts (src/io/layer.rs:135-199)
            std::vector<Pos> mp;
            std::vector<uint32_t> ms;
            mp.reserve(pos.size() + ivs.size());
            ms.reserve(pos.size() + ivs.size());
            size_t a = 0, c = 0;
            while (a < pos.size() || c < ivs.size()) {
                if (c == ivs.size() || (a < pos.size() && pos[a].start <= ivs[c].start)) {
                    mp.push_back(pos[a]);
                    ms.push_back(sid[a++]);
                } else {
                    mp.push_back({ivs[c].start, ivs[c].end});
                    ms.push_back(ivs[c++].sid);
                }
            }
            write_bytes(pp, mp.data(), mp.size() * sizeof(Pos));
            write_bytes(sp, ms.data(), ms.size() * 4);
            const auto jt = jump_table(mp, meta.lc.tile_size(layer));
            write_bytes(layer_path(db, shard, layer, "idx"), jt.data(), jt.size() * 8);
            if (std::find(meta.shards.begin(), meta.shards.end(), shard) == meta.shards.end())
                meta.shards.push_back(shard);
            max_layer = std::max(max_layer, layer);
        }
        meta.num_layers = max_layer + 1;
        meta.save(db);
    }
    if (stats) {
        // moment tables of every file on the device, from the written layers
        const Flat f = load_db_intervals(db, meta);
        const uint32_t F = meta.next_sid();
        IndexHandle ix;
        chk(chk_index_create(0, (uint32_t)meta.shards.size(), F, f.start.size(), f.shard.data(),
                             f.start.data(), f.end.data(), f.sid.data(), &ix.p));
        std::set<std::string> wl(chroms.begin(), chroms.end());
        const auto g = genome_for(meta.shards, chroms.empty() ? nullptr : &wl);
        chk(chk_index_set_genome(ix.p, g.data()));
        int32_t rows = 0, exact = 0;
        chk(chk_moments_build(ix.p, &rows, &exact));
        save_moments(db, meta, F, f.start.size(), ix.p);
    }
}

// ------------------------------------------------------------------ query

struct QFile {
    std::string name;
    std::vector<uint32_t> shard, start, end;  // shard = DB shard index or nshards
};

void run_query(const std::string& db, const std::string& qpath, const std::string& out, bool stats,
               bool intervals, size_t batch_size) {
    if (intervals)
        throw Fail("--intervals (pair listing) is not supported by this build; use counts or --stats");
    if (!exists(join(db, "meta.json"))) throw Fail("Database not found at " + db);
    const Meta meta = Meta::load(db);
    const uint32_t S = (uint32_t)meta.shards.size();
    const uint32_t F = meta.next_sid();
    std::unordered_map<std::string, uint32_t> shard_of;
    for (uint32_t s = 0; s < S; s++) shard_of[meta.shards[s]] = s;
    // query files sorted by path (src/tasks/query/files.rs:26-39)
    std::vector<std::string> files;
    if (is_dir(qpath)) {
        for (auto& p : list_dir(qpath))
            if (is_bed_file(p)) files.push_back(p);
        std::sort(files.begin(), files.end());
    } else {
        files.push_back(qpath);
    }
    const Flat f = load_db_intervals(db, meta);
    IndexHandle ix;
    chk(chk_index_create(0, S, std::max<uint32_t>(F, 1), f.start.size(), f.shard.data(), f.start.data(),
                         f.end.data(), f.sid.data(), &ix.p));
    if (stats) {
        const auto g = genome_for(meta.shards, nullptr);
        chk(chk_index_set_genome(ix.p, g.data()));
        if (!load_moments(db, meta, F, f.start.size(), ix.p)) {
            int32_t rows = 0, exact = 0;
            chk(chk_moments_build(ix.p, &rows, &exact));
        }
    }
    struct Rec {
        std::string q, d;
        uint32_t count;
        double obs, p, llr;
        bool has_llr;
    };
    std::vector<Rec> recs;
    size_t nquery = 0;
    const size_t bs = std::max<size_t>(1, batch_size ? batch_size : files.size());
    for (size_t b0 = 0; b0 < files.size(); b0 += bs) {
        // parse; empty files get no Q_SID (files.rs:79-101)
        std::vector<QFile> qs;
        for (size_t i = b0; i < std::min(files.size(), b0 + bs); i++) {
            auto shards = parse_bed_file(files[i], 0);
            QFile qf;
            qf.name = basename_of(files[i]);
            for (auto& kv : shards) {
                auto it = shard_of.find(kv.first);
                const uint32_t s = it == shard_of.end() ? S : it->second;
                for (auto& iv : kv.second) {
                    qf.shard.push_back(s);
                    qf.start.push_back(iv.start);
                    qf.end.push_back(iv.end);
                }
            }
            if (!qf.start.empty()) qs.push_back(std::move(qf));
        }
        if (qs.empty()) continue;
        nquery += qs.size();
        std::vector<uint32_t> sh, st, en, qf;
        for (uint32_t r = 0; r < qs.size(); r++) {
            sh.insert(sh.end(), qs[r].shard.begin(), qs[r].shard.end());
            st.insert(st.end(), qs[r].start.begin(), qs[r].start.end());
            en.insert(en.end(), qs[r].end.begin(), qs[r].end.end());
            qf.insert(qf.end(), qs[r].start.size(), r);
        }
        chk_query* q = nullptr;
        chk(chk_query_create(ix.p, (uint32_t)qs.size(), st.size(), sh.data(), st.data(), en.data(),
                             qf.data(), &q));
        const size_t cells = qs.size() * (size_t)F;
        std::vector<uint64_t> cnt(cells), bp(cells);
        std::vector<double> p(cells), llr(cells);
        std::vector<uint8_t> fl(cells);
        int rc = stats ? chk_query_stats(ix.p, q, cnt.data(), bp.data(), p.data(), llr.data(), fl.data())
                       : chk_sweep(ix.p, q, cnt.data(), bp.data());
        chk_query_free(q);
        chk(rc);
        for (uint32_t r = 0; r < qs.size(); r++)
            for (uint32_t d = 0; d < F; d++) {
                const size_t c = (size_t)r * F + d;
                const auto nm = meta.sid_map.find(d);
                const std::string dn = nm == meta.sid_map.end() ? "" : nm->second;
                if (stats) {
                    if (!(fl[c] & 1)) continue;  // no p-value record
                    recs.push_back({qs[r].name, dn, (uint32_t)cnt[c], (double)bp[c] / 100.0, p[c], llr[c],
                                    (fl[c] & 2) != 0});
                } else if (cnt[c]) {
                    recs.push_back({qs[r].name, dn, (uint32_t)cnt[c], 0, 0, 0, false});
                }
            }
    }
    // src/main.rs:194-201: by p ascending when present, else count descending
    if (stats)
        std::stable_sort(recs.begin(), recs.end(), [](const Rec& a, const Rec& b) { return a.p < b.p; });
    else
        std::stable_sort(recs.begin(), recs.end(), [](const Rec& a, const Rec& b) { return a.count > b.count; });
    std::string o = "{\n  \"results\": [";
    for (size_t k = 0; k < recs.size(); k++) {
        const Rec& r = recs[k];
        o += k ? ",\n    {\n" : "\n    {\n";
        o += "      \"query_name\": " + json_escape(r.q) + ",\n";
        o += "      \"db_name\": " + json_escape(r.d) + ",\n";
        o += "      \"overlap_count\": " + std::to_string(r.count);
        if (stats) {
            o += ",\n      \"observed_bins\": " + json_f64(r.obs);
            o += ",\n      \"p_value\": " + json_f64(r.p);
            if (r.has_llr) o += ",\n      \"llr\": " + json_f64(r.llr);
        }
        o += "\n    }";
    }
    o += recs.empty() ? "]\n}" : "\n  ]\n}";
    write_bytes(out, o.data(), o.size());
    std::printf("Query complete: %zu query files \xc3\x97 %zu database sources \xe2\x86\x92 %s\n", nquery,
                meta.sid_map.size(), out.c_str());
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
    const auto qbed = parse_bed_file(qpath, 0);
    size_t n_ivs = 0;
    for (auto& kv : qbed) n_ivs += kv.second.size();
    std::fprintf(stderr, "Query: %zu intervals across %zu chromosomes\n", n_ivs, qbed.size());
    std::map<std::string, std::vector<Iv>> wl;
    const bool have_wl = !wl_path.empty();
    if (have_wl) {
        wl = parse_bed_file(wl_path, 0);
        std::fprintf(stderr, "Whitelist: active \xe2\x80\x94 shifts clipped to annotated regions\n");
    }
    std::fprintf(stderr, "Mode: %s\n", shuffle ? "shuffle (per-interval independent placement)"
                                                : "circular (global shift)");
    std::map<std::string, uint32_t> chrom_size;
    int64_t max_shift = 0;
    for (auto& c : HG38) {
        chrom_size[c.first] = (uint32_t)c.second;
        max_shift = std::max<int64_t>(max_shift, (int64_t)c.second);
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
        const Flat f = load_db_intervals(db, meta);
        chk(chk_index_create(0, S, std::max<uint32_t>(F, 1), f.start.size(), f.shard.data(), f.start.data(),
                             f.end.data(), f.sid.data(), &ix.p));
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
                    if (w) intersect_whitelist(a, b, *w, [&](uint32_t x, uint32_t y) { out.push_back({kv.first, {x, y, 0}}); });
                    else out.push_back({kv.first, {a, b, 0}});
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
            mkdirs(bd);
            for (size_t k = 0; k < trials.size(); k++) {
                char nm[32];
                std::snprintf(nm, sizeof nm, "trial_%08zu.bed", k);
                std::string txt;
                for (auto& ci : trials[k])
                    txt += ci.first + "\t" + std::to_string(ci.second.start) + "\t" + std::to_string(ci.second.end) + "\n";
                write_bytes(join(bd, nm), txt.data(), txt.size());
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
            run_add(input, db, batch, stats, chroms);
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
