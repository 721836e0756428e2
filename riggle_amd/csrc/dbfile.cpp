// dbfile.cpp -- host side of the reference's database and BED formats (see
// dbfile.hpp): parsing over host threads, layer files, meta.json, the
// moment-table file and the result JSON.
#include "dbfile.hpp"

#include <dirent.h>
#include <sys/stat.h>
#include <zlib.h>

#include <algorithm>
#include <atomic>
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <memory>
#include <mutex>
#include <thread>
#include <unordered_map>

#include "engine.hpp"

namespace chk {
namespace db {

namespace {

[[noreturn]] void io_fail(const std::string& msg) { throw Error(CHK_ERR_IO, msg); }

// Runs f(k) for k in [0, n) over up to `threads` host threads.
template <typename F>
void parallel_for(size_t n, unsigned threads, F&& f) {
    threads = (unsigned)std::max<size_t>(1, std::min<size_t>(threads, n));
    if (threads <= 1) {
        for (size_t k = 0; k < n; k++) f(k);
        return;
    }
    std::atomic<size_t> next{0};
    host_run(threads, [&](unsigned) {
        for (size_t k; (k = next.fetch_add(1)) < n;) f(k);
    });
}

// One host_run call: task indices are claimed by the workers and the caller;
// every index is claimed and counted, and after a task throws the remaining
// ones are claimed without running.
struct HostJob {
    const std::function<void(unsigned)>* f = nullptr;
    unsigned T = 0;
    std::atomic<unsigned> next{0}, done{0};
    std::atomic<bool> failed{false};
    std::mutex mu;
    std::condition_variable cv;
    std::exception_ptr err;

    void work() {
        for (unsigned t; (t = next.fetch_add(1)) < T;) {
            if (!failed.load()) {
                try {
                    (*f)(t);
                } catch (...) {
                    std::lock_guard<std::mutex> g(mu);
                    if (!err) err = std::current_exception();
                    failed = true;
                }
            }
            if (done.fetch_add(1) + 1 == T) {
                std::lock_guard<std::mutex> g(mu);
                cv.notify_all();
            }
        }
    }
};

// Persistent workers, never destroyed: at process exit they are idle on
// the condition variable.
struct HostPool {
    std::mutex mu;
    std::condition_variable cv;
    std::deque<std::shared_ptr<HostJob>> q;
    std::vector<std::thread> workers;

    explicit HostPool(unsigned n) {
        for (unsigned i = 0; i < n; i++) workers.emplace_back([this] { loop(); });
        for (auto& w : workers) w.detach();
    }
    void loop() {
        for (;;) {
            std::shared_ptr<HostJob> j;
            {
                std::unique_lock<std::mutex> lk(mu);
                cv.wait(lk, [&] { return !q.empty(); });
                j = q.front();
                if (j->next.load() >= j->T) {  // every task taken: drop the job
                    q.pop_front();
                    continue;
                }
            }
            j->work();
        }
    }
};

HostPool& host_pool() {
    static HostPool* p = new HostPool(host_threads() > 1 ? host_threads() - 1 : 0);
    return *p;
}

}  // namespace

void host_run(unsigned T, const std::function<void(unsigned)>& f) {
    if (T == 0) return;
    if (T == 1) {
        f(0);
        return;
    }
    auto j = std::make_shared<HostJob>();
    j->f = &f;
    j->T = T;
    HostPool& P = host_pool();
    {
        std::lock_guard<std::mutex> g(P.mu);
        P.q.push_back(j);
    }
    P.cv.notify_all();
    j->work();  // the caller takes every task no worker has started
    {
        std::unique_lock<std::mutex> lk(j->mu);
        j->cv.wait(lk, [&] { return j->done.load() == j->T; });
    }
    {  // a worker normally drops the finished job; without workers (a forked child) the caller does
        std::lock_guard<std::mutex> g(P.mu);
        for (auto it = P.q.begin(); it != P.q.end(); ++it)
            if (*it == j) {
                P.q.erase(it);
                break;
            }
    }
    if (j->err) std::rethrow_exception(j->err);
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
    if (p.empty() || is_dir(p)) return;
    const size_t k = p.find_last_of('/');
    if (k != std::string::npos && k > 0) mkdirs(p.substr(0, k));
    if (::mkdir(p.c_str(), 0755) != 0 && errno != EEXIST)
        io_fail("IO error: cannot create directory " + p + ": " + std::strerror(errno));
}

std::string join(const std::string& a, const std::string& b) {
    return a.empty() || a.back() == '/' ? a + b : a + "/" + b;
}

std::string basename_of(const std::string& p) {
    const size_t k = p.find_last_of('/');
    return k == std::string::npos ? p : p.substr(k + 1);
}

// src/io/parse.rs:16-19 (the file name ends in .bed or .bed.gz)
bool is_bed_file(const std::string& p) {
    const std::string n = basename_of(p);
    auto ends = [&](const char* suf) {
        const size_t k = std::strlen(suf);
        return n.size() >= k && n.compare(n.size() - k, k, suf) == 0;
    };
    return ends(".bed") || ends(".bed.gz");
}

std::vector<std::string> list_dir(const std::string& d) {
    std::vector<std::string> out;
    DIR* dp = ::opendir(d.c_str());
    if (!dp) io_fail("IO error: cannot read directory " + d + ": " + std::strerror(errno));
    while (struct dirent* e = ::readdir(dp)) {
        if (!std::strcmp(e->d_name, ".") || !std::strcmp(e->d_name, "..")) continue;
        out.push_back(join(d, e->d_name));
    }
    ::closedir(dp);
    return out;
}

std::string read_file(const std::string& p) {
    FILE* f = std::fopen(p.c_str(), "rb");
    if (!f) io_fail("IO error: cannot open " + p + ": " + std::strerror(errno));
    std::string s;
    if (std::fseek(f, 0, SEEK_END) == 0) {
        const long sz = std::ftell(f);
        if (sz > 0) s.reserve((size_t)sz);
        std::fseek(f, 0, SEEK_SET);
    }
    char buf[1 << 16];
    size_t k;
    while ((k = std::fread(buf, 1, sizeof buf, f)) > 0) s.append(buf, k);
    std::fclose(f);
    return s;
}

void write_file(const std::string& p, const void* data, size_t n) {
    FILE* f = std::fopen(p.c_str(), "wb");
    if (!f) io_fail("IO error: cannot write " + p + ": " + std::strerror(errno));
    if (n && std::fwrite(data, 1, n, f) != n) {
        std::fclose(f);
        io_fail("IO error: short write to " + p);
    }
    std::fclose(f);
}

unsigned host_threads() {
    const unsigned h = std::thread::hardware_concurrency();
    return std::max(1u, std::min(16u, h ? h : 1u));
}

template <typename T>
static std::vector<T> read_array(const std::string& p) {
    if (!exists(p)) return {};
    const std::string s = read_file(p);
    if (s.size() % sizeof(T))
        io_fail(p + " has invalid size: " + std::to_string(s.size()) + " bytes");
    std::vector<T> v(s.size() / sizeof(T));
    if (!v.empty()) std::memcpy(v.data(), s.data(), s.size());
    return v;
}

// ------------------------------------------------------------------ BED

namespace {

// The whole file, decompressed when it is gzip (zlib reads plain files as
// they are).
std::string slurp_bed(const std::string& path) {
    gzFile gz = gzopen(path.c_str(), "rb");
    if (!gz) io_fail("IO error: cannot open " + path);
    gzbuffer(gz, 1 << 20);
    std::string s;
    struct stat st;
    if (::stat(path.c_str(), &st) == 0 && st.st_size > 0) s.reserve((size_t)st.st_size);
    std::vector<char> buf(1 << 22);
    for (;;) {
        const int k = gzread(gz, buf.data(), (unsigned)buf.size());
        if (k < 0) {
            gzclose(gz);
            io_fail("IO error: cannot read " + path);
        }
        if (k == 0) break;
        s.append(buf.data(), (size_t)k);
    }
    gzclose(gz);
    return s;
}

// A token parses as u32 when it is an optional '+' then decimal digits whose
// value fits (Rust's u32::from_str).
inline bool parse_u32(const char* b, const char* e, uint32_t* out) {
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

inline bool is_ws(char c) { return c == ' ' || c == '\t' || c == '\n' || c == '\r' || c == '\v' || c == '\f'; }

// Parse of one run of whole lines.
struct Chunk {
    BedFile out;
    std::unordered_map<std::string, uint32_t> idx;
    uint64_t lines = 0;  // lines consumed (up to and including a failing one)
    std::string err;     // message without the line number
};

void parse_lines(const char* p, const char* end, Chunk& c) {
    std::string shard;
    while (p < end) {
        const char* nl = static_cast<const char*>(std::memchr(p, '\n', (size_t)(end - p)));
        const char* le = nl ? nl : end;
        c.lines++;
        const char* a = p;
        const char* e = le;
        p = nl ? nl + 1 : end;
        while (a < e && is_ws(*a)) a++;
        while (e > a && is_ws(e[-1])) e--;
        if (a == e || *a == '#') continue;
        bool have_shard = false, have_s = false, have_e = false;
        uint32_t s = 0, en = 0;
        int cols = 0;
        const char* sh0 = nullptr;
        const char* sh1 = nullptr;
        while (a < e) {
            while (a < e && is_ws(*a)) a++;
            if (a == e) break;
            const char* t = a;
            while (a < e && !is_ws(*a)) a++;
            cols++;
            uint32_t v;
            if (parse_u32(t, a, &v)) {
                if (!have_s) {
                    s = v;
                    have_s = true;
                } else if (!have_e) {
                    en = v;
                    have_e = true;
                }
            } else if (!have_shard) {
                sh0 = t;
                sh1 = a;
                have_shard = true;
            }
            if (have_shard && have_e) break;
        }
        if (cols < 2 || !have_s || !have_e) {
            c.err = "invalid format - expected at least 2 integer columns";
            return;
        }
        if (s >= en) {
            c.err = "start (" + std::to_string(s) + ") must be less than end (" + std::to_string(en) + ")";
            return;
        }
        if (have_shard) shard.assign(sh0, sh1);
        else shard.assign("default");
        auto it = c.idx.find(shard);
        uint32_t k;
        if (it == c.idx.end()) {
            k = (uint32_t)c.out.shards.size();
            c.idx.emplace(shard, k);
            c.out.shards.push_back(shard);
            c.out.ivs.emplace_back();
        } else {
            k = it->second;
        }
        c.out.ivs[k].push_back({s, en});
        c.out.n++;
    }
}

}  // namespace

BedFile parse_bed(const std::string& path, unsigned threads) {
    const std::string text = slurp_bed(path);
    const char* b = text.data();
    const size_t n = text.size();
    // chunks of whole lines, about 8 MiB or more each
    const size_t T = std::max<size_t>(1, std::min<size_t>(threads ? threads : 1, n >> 23));
    std::vector<size_t> cut{0};
    for (size_t t = 1; t < T; t++) {
        size_t k = std::max(cut.back(), n * t / T);
        const void* nl = k < n ? std::memchr(b + k, '\n', n - k) : nullptr;
        k = nl ? (size_t)(static_cast<const char*>(nl) - b) + 1 : n;
        cut.push_back(k);
    }
    cut.push_back(n);
    std::vector<Chunk> ch(cut.size() - 1);
    parallel_for(ch.size(), (unsigned)ch.size(), [&](size_t k) { parse_lines(b + cut[k], b + cut[k + 1], ch[k]); });
    uint64_t line0 = 0;
    for (auto& c : ch) {
        if (!c.err.empty()) throw Error(CHK_ERR_INVALID, "Line " + std::to_string(line0 + c.lines) + ": " + c.err);
        line0 += c.lines;
    }
    if (ch.size() == 1) return std::move(ch[0].out);
    BedFile out;
    std::unordered_map<std::string, uint32_t> idx;
    for (auto& c : ch)
        for (size_t s = 0; s < c.out.shards.size(); s++) {
            auto it = idx.find(c.out.shards[s]);
            uint32_t k;
            if (it == idx.end()) {
                k = (uint32_t)out.shards.size();
                idx.emplace(c.out.shards[s], k);
                out.shards.push_back(c.out.shards[s]);
                out.ivs.emplace_back();
            } else {
                k = it->second;
            }
            auto& dst = out.ivs[k];
            dst.insert(dst.end(), c.out.ivs[s].begin(), c.out.ivs[s].end());
        }
    out.n = 0;
    for (auto& v : out.ivs) out.n += v.size();
    return out;
}

std::vector<BedFile> parse_beds(const std::vector<std::string>& paths, unsigned threads) {
    std::vector<BedFile> out(paths.size());
    if (paths.size() >= threads) {
        parallel_for(paths.size(), threads, [&](size_t k) { out[k] = parse_bed(paths[k], 1); });
    } else {
        // few files: each one split over the threads
        for (size_t k = 0; k < paths.size(); k++) out[k] = parse_bed(paths[k], threads);
    }
    return out;
}

// ------------------------------------------------------------------ JSON

const JVal* JVal::get(const std::string& k) const {
    for (auto& kv : obj)
        if (kv.first == k) return &kv.second;
    return nullptr;
}

namespace {

struct JParser {
    const std::string& s;
    const std::string& what;
    size_t i = 0;
    JParser(const std::string& x, const std::string& w) : s(x), what(w) {}
    void ws() {
        while (i < s.size() && std::isspace((unsigned char)s[i])) i++;
    }
    [[noreturn]] void bad() {
        throw Error(CHK_ERR_IO, "JSON error: malformed " + what + " at byte " + std::to_string(i));
    }
    char at() {
        if (i >= s.size()) bad();
        return s[i];
    }
    static void utf8(std::string& out, unsigned cp) {
        if (cp < 0x80) {
            out += (char)cp;
        } else if (cp < 0x800) {
            out += (char)(0xC0 | (cp >> 6));
            out += (char)(0x80 | (cp & 63));
        } else if (cp < 0x10000) {
            out += (char)(0xE0 | (cp >> 12));
            out += (char)(0x80 | ((cp >> 6) & 63));
            out += (char)(0x80 | (cp & 63));
        } else {
            out += (char)(0xF0 | (cp >> 18));
            out += (char)(0x80 | ((cp >> 12) & 63));
            out += (char)(0x80 | ((cp >> 6) & 63));
            out += (char)(0x80 | (cp & 63));
        }
    }
    std::string string() {
        if (at() != '"') bad();
        i++;
        std::string out;
        while (at() != '"') {
            if (s[i] != '\\') {
                out += s[i++];
                continue;
            }
            i++;
            const char c = at();
            i++;
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
                        const unsigned lo = (unsigned)std::strtoul(s.substr(i + 2, 4).c_str(), nullptr, 16);
                        i += 6;
                        cp = 0x10000 + ((cp - 0xD800) << 10) + (lo - 0xDC00);
                    }
                    utf8(out, cp);
                    break;
                }
                default: out += c;
            }
        }
        i++;
        return out;
    }
    JVal value() {
        ws();
        JVal v;
        const char c = at();
        if (c == '{') {
            v.kind = JVal::Obj;
            i++;
            ws();
            if (at() == '}') {
                i++;
                return v;
            }
            for (;;) {
                ws();
                std::string k = string();
                ws();
                if (at() != ':') bad();
                i++;
                v.obj.emplace_back(k, value());
                ws();
                if (at() == ',') {
                    i++;
                    continue;
                }
                if (at() == '}') {
                    i++;
                    break;
                }
                bad();
            }
        } else if (c == '[') {
            v.kind = JVal::Arr;
            i++;
            ws();
            if (at() == ']') {
                i++;
                return v;
            }
            for (;;) {
                v.arr.push_back(value());
                ws();
                if (at() == ',') {
                    i++;
                    continue;
                }
                if (at() == ']') {
                    i++;
                    break;
                }
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

}  // namespace

JVal parse_json(const std::string& text, const std::string& what) { return JParser(text, what).value(); }

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

uint32_t LayerConfig::layer_for_size(uint32_t size) const {
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

uint32_t LayerConfig::layer_max_size(uint32_t layer) const {
    uint64_t v = min_size;
    for (uint32_t k = 0; k < layer; k++) v = std::min<uint64_t>(v * growth_factor, 0xFFFFFFFFull);
    return (uint32_t)v;
}

uint32_t LayerConfig::tile_size(uint32_t layer) const {
    return std::max<uint32_t>(1, layer_max_size(layer) / 2);
}

Meta Meta::load(const std::string& db) {
    const std::string path = join(db, "meta.json");
    if (!exists(path)) throw Error(CHK_ERR_NOT_FOUND, "Database not found at " + db);
    const JVal root = parse_json(read_file(path), "meta.json");
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

void Meta::save(const std::string& db) const {
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
    write_file(join(db, "meta.json"), o.data(), o.size());
}

std::string layer_path(const std::string& db, const std::string& shard, uint32_t k, const char* ext) {
    return join(join(join(db, "shards"), shard), "layer_" + std::to_string(k) + "." + ext);
}

std::vector<uint64_t> jump_table(const Pos* p, size_t n, uint32_t tile) {
    if (n == 0) return {0};
    const uint64_t max_tile = p[n - 1].start / tile;
    std::vector<uint64_t> t(max_tile + 1, 0);
    size_t i = 0;
    for (uint64_t k = 0; k <= max_tile; k++) {
        const uint64_t lim = k * (uint64_t)tile;
        while (i < n && p[i].start < lim) i++;
        t[k] = i;
    }
    return t;
}

Flat load_intervals(const std::string& db, const Meta& m) {
    struct Part {
        uint32_t shard;
        std::string pos, sid;
        uint64_t n = 0, off = 0;
    };
    std::vector<Part> parts;
    for (uint32_t s = 0; s < m.shards.size(); s++)
        for (uint32_t k = 0; k < m.num_layers; k++) {
            Part p;
            p.shard = s;
            p.pos = layer_path(db, m.shards[s], k, "pos");
            p.sid = layer_path(db, m.shards[s], k, "sid");
            struct stat a, b;
            const bool ha = ::stat(p.pos.c_str(), &a) == 0, hb = ::stat(p.sid.c_str(), &b) == 0;
            if (!ha && !hb) continue;
            const uint64_t np = ha ? (uint64_t)a.st_size : 0, ns = hb ? (uint64_t)b.st_size : 0;
            if (np % 8 || ns % 4 || np / 8 != ns / 4)
                throw Error(CHK_ERR_IO, "Position count " + std::to_string(np / 8) +
                                            " does not match sid count " + std::to_string(ns / 4) +
                                            " in " + p.pos);
            p.n = np / 8;
            parts.push_back(std::move(p));
        }
    uint64_t total = 0;
    for (auto& p : parts) {
        p.off = total;
        total += p.n;
    }
    Flat f;
    f.shard.resize(total);
    f.start.resize(total);
    f.end.resize(total);
    f.sid.resize(total);
    parallel_for(parts.size(), host_threads(), [&](size_t k) {
        const Part& p = parts[k];
        const auto pos = read_array<Pos>(p.pos);
        const auto sid = read_array<uint32_t>(p.sid);
        if (pos.size() != p.n || sid.size() != p.n) throw Error(CHK_ERR_IO, p.pos + " changed while read");
        for (uint64_t i = 0; i < p.n; i++) {
            f.shard[p.off + i] = p.shard;
            f.start[p.off + i] = pos[i].start;
            f.end[p.off + i] = pos[i].end;
        }
        if (p.n) std::memcpy(f.sid.data() + p.off, sid.data(), p.n * 4);
    });
    return f;
}

// ------------------------------------------------------------------ moments

namespace {

const std::pair<const char*, uint64_t> HG38[] = {
    {"chr1", 248956422}, {"chr2", 242193529}, {"chr3", 198295559}, {"chr4", 190214555},
    {"chr5", 181538259}, {"chr6", 170805979}, {"chr7", 159345973}, {"chr8", 145138636},
    {"chr9", 138394717}, {"chr10", 133797422}, {"chr11", 135086622}, {"chr12", 133275309},
    {"chr13", 114364328}, {"chr14", 107043718}, {"chr15", 101991189}, {"chr16", 90338345},
    {"chr17", 83257441}, {"chr18", 80373285}, {"chr19", 58617616}, {"chr20", 64444167},
    {"chr21", 46709983}, {"chr22", 50818468}, {"chrX", 156040895}, {"chrY", 57227415}};

}  // namespace

std::vector<uint64_t> genome_for(const std::vector<std::string>& shards,
                                 const std::set<std::string>* wl) {
    std::vector<uint64_t> g(shards.size(), 0);
    for (size_t s = 0; s < shards.size(); s++)
        for (auto& e : HG38)
            if (shards[s] == e.first && (!wl || wl->count(shards[s]))) g[s] = e.second;
    return g;
}

// DB/momentmap.chk: "CHKMOM01", u32 nfiles, u32 nshards, u64 ndb_intervals,
// u64 total, then per shard (u32 name length, bytes), i64 slot_off[F*S],
// u32 slot_len[F*S], f64 mean[total], f64 var[total].
void save_moments(const std::string& db, Index& ix, const Meta& m) {
    const uint64_t total = ix.mom.total;
    const uint32_t F = ix.nfiles, S = ix.nshards;
    const size_t ns = (size_t)F * S;
    std::vector<int64_t> so(ns);
    std::vector<uint32_t> sl(ns);
    std::vector<double> mean(std::max<uint64_t>(total, 1)), var(std::max<uint64_t>(total, 1));
    moments_download(ix, so.data(), sl.data(), mean.data(), var.data());
    std::string o = "CHKMOM01";
    auto put = [&](const void* p, size_t k) { o.append((const char*)p, k); };
    const uint64_t n = ix.n;
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
    write_file(join(db, "momentmap.chk"), o.data(), o.size());
}

bool load_moments(const std::string& db, Index& ix, const Meta& m) {
    const std::string p = join(db, "momentmap.chk");
    if (!exists(p)) return false;
    const std::string s = read_file(p);
    size_t i = 8;
    auto get = [&](void* d, size_t k) {
        if (i + k > s.size()) io_fail("truncated " + p);
        std::memcpy(d, s.data() + i, k);
        i += k;
    };
    if (s.size() < 8 || s.compare(0, 8, "CHKMOM01")) return false;
    uint32_t F2, S2;
    uint64_t n2, total;
    get(&F2, 4);
    get(&S2, 4);
    get(&n2, 8);
    get(&total, 8);
    if (F2 != ix.nfiles || S2 != m.shards.size() || S2 != ix.nshards || n2 != ix.n) return false;
    for (uint32_t k = 0; k < S2; k++) {
        uint32_t L;
        get(&L, 4);
        std::string name(L, '\0');
        if (L) get(&name[0], L);
        if (name != m.shards[k]) return false;
    }
    const size_t ns = (size_t)F2 * S2;
    std::vector<int64_t> so(ns);
    std::vector<uint32_t> sl(ns);
    std::vector<double> mean(std::max<uint64_t>(total, 1)), var(std::max<uint64_t>(total, 1));
    get(so.data(), ns * 8);
    get(sl.data(), ns * 4);
    get(mean.data(), total * 8);
    get(var.data(), total * 8);
    moments_upload(ix, so.data(), sl.data(), total, mean.data(), var.data());
    return true;
}

// ------------------------------------------------------------------ results

void write_results(const std::string& path, const std::vector<std::string>& qnames,
                   const std::vector<std::string>& dnames, const uint64_t* counts, const uint64_t* bp,
                   const double* p, const double* llr, const uint8_t* flags, bool stats) {
    struct Rec {
        uint32_t r, d;
        double p;
        uint64_t count;
    };
    const size_t R = qnames.size(), F = dnames.size();
    std::vector<Rec> recs;
    for (size_t r = 0; r < R; r++)
        for (size_t d = 0; d < F; d++) {
            const size_t c = r * F + d;
            if (stats ? (flags[c] & 1) != 0 : counts[c] != 0)
                recs.push_back({(uint32_t)r, (uint32_t)d, stats ? p[c] : 0.0, counts[c]});
        }
    // src/main.rs:194-201: by p ascending when present, else count descending
    if (stats)
        std::stable_sort(recs.begin(), recs.end(), [](const Rec& a, const Rec& b) { return a.p < b.p; });
    else
        std::stable_sort(recs.begin(), recs.end(),
                         [](const Rec& a, const Rec& b) { return (uint32_t)a.count > (uint32_t)b.count; });
    std::string o = "{\n  \"results\": [";
    o.reserve(recs.size() * (stats ? 200 : 120) + 32);
    for (size_t k = 0; k < recs.size(); k++) {
        const Rec& x = recs[k];
        const size_t c = (size_t)x.r * F + x.d;
        o += k ? ",\n    {\n" : "\n    {\n";
        o += "      \"query_name\": " + json_escape(qnames[x.r]) + ",\n";
        o += "      \"db_name\": " + json_escape(dnames[x.d]) + ",\n";
        // the reference's counts are u32 (src/matrix/dense.rs)
        o += "      \"overlap_count\": " + std::to_string((uint32_t)counts[c]);
        if (stats) {
            o += ",\n      \"observed_bins\": " + json_f64((double)bp[c] / 100.0);
            o += ",\n      \"p_value\": " + json_f64(p[c]);
            if (flags[c] & 2) o += ",\n      \"llr\": " + json_f64(llr[c]);
        }
        o += "\n    }";
    }
    o += recs.empty() ? "]\n}" : "\n  ]\n}";
    write_file(path, o.data(), o.size());
}

}  // namespace db
}  // namespace chk
