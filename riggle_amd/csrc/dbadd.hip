// dbadd.hip -- `chuckle add` (src/tasks/build.rs:86-276) with the bucket
// sort on the device.
//
// Each batch of BED files is parsed over host threads.  Every interval gets
// the key (bucket << 32 | start), bucket = shard * layers + size layer, and
// one stable device radix sort orders the whole batch by (bucket, start):
// within a bucket, intervals of equal start keep file order (D_SID order)
// and line order.  A gather kernel writes the sorted intervals in the layer
// files' own byte layout (u32 start, end pairs and u32 D_SIDs), so each
// bucket's slice is copied out as it is written.  Buckets are then written
// (or merged into the existing layer files, existing intervals first among
// equal starts as src/io/layer.rs:135-199 does) with their jump tables over
// host threads, and meta.json is updated.
#include <algorithm>
#include <atomic>
#include <cstring>
#include <thread>
#include <unordered_map>

#include "dbfile.hpp"
#include "engine.hpp"

namespace chk {
namespace db {

namespace {

constexpr int NT = 256;

__global__ void add_keys(const uint32_t* bucket, const uint32_t* start, uint64_t n, uint64_t* keys,
                         uint32_t* vals) {
    const uint64_t i = (uint64_t)blockIdx.x * NT + threadIdx.x;
    if (i >= n) return;
    keys[i] = ((uint64_t)bucket[i] << 32) | start[i];
    vals[i] = (uint32_t)i;
}

__global__ void add_gather(const uint32_t* perm, const uint32_t* start, const uint32_t* end,
                           const uint32_t* sid, uint64_t n, uint2* pos, uint32_t* osid) {
    const uint64_t i = (uint64_t)blockIdx.x * NT + threadIdx.x;
    if (i >= n) return;
    const uint32_t j = perm[i];
    pos[i] = make_uint2(start[j], end[j]);
    osid[i] = sid[j];
}

struct Stream {
    hipStream_t s = nullptr;
    ~Stream() {
        if (s) (void)hipStreamDestroy(s);
    }
};

template <typename T>
std::vector<T> read_vec(const std::string& p) {
    const std::string s = read_file(p);
    if (s.size() % sizeof(T)) throw Error(CHK_ERR_IO, p + " has invalid size");
    std::vector<T> v(s.size() / sizeof(T));
    if (!v.empty()) std::memcpy(v.data(), s.data(), s.size());
    return v;
}

// Writes (or merges into) one layer and its jump table.
void write_bucket(const std::string& db, const std::string& shard, uint32_t layer, uint32_t tile,
                  const Pos* pos, const uint32_t* sid, size_t n) {
    mkdirs(join(join(db, "shards"), shard));
    const std::string pp = layer_path(db, shard, layer, "pos"), sp = layer_path(db, shard, layer, "sid");
    if (!exists(pp)) {
        write_file(pp, pos, n * sizeof(Pos));
        write_file(sp, sid, n * 4);
        const auto jt = jump_table(pos, n, tile);
        write_file(layer_path(db, shard, layer, "idx"), jt.data(), jt.size() * 8);
        return;
    }
    const auto op = read_vec<Pos>(pp);
    const auto os = exists(sp) ? read_vec<uint32_t>(sp) : std::vector<uint32_t>();
    if (op.size() != os.size())
        throw Error(CHK_ERR_IO, "Position count " + std::to_string(op.size()) + " does not match sid count " +
                                    std::to_string(os.size()) + " in " + pp);
    std::vector<Pos> mp(op.size() + n);
    std::vector<uint32_t> ms(op.size() + n);
    size_t a = 0, c = 0, k = 0;
    while (a < op.size() || c < n) {
        if (c == n || (a < op.size() && op[a].start <= pos[c].start)) {
            mp[k] = op[a];
            ms[k++] = os[a++];
        } else {
            mp[k] = pos[c];
            ms[k++] = sid[c++];
        }
    }
    write_file(pp, mp.data(), mp.size() * sizeof(Pos));
    write_file(sp, ms.data(), ms.size() * 4);
    const auto jt = jump_table(mp.data(), mp.size(), tile);
    write_file(layer_path(db, shard, layer, "idx"), jt.data(), jt.size() * 8);
}

}  // namespace

void add(const std::string& input, const std::string& db, uint64_t batch_size, bool stats,
         const std::set<std::string>& chroms, bool sorted_sids, int device) {
    if (!exists(input)) throw Error(CHK_ERR_NOT_FOUND, "Input path not found: " + input);
    std::vector<std::string> beds;
    if (is_dir(input)) {
        for (auto& p : list_dir(input))
            if (is_bed_file(p)) beds.push_back(p);
    } else if (is_bed_file(input)) {
        beds.push_back(input);
    }
    if (beds.empty()) throw Error(CHK_ERR_NOT_FOUND, "No BED files found in " + input);
    if (sorted_sids) std::sort(beds.begin(), beds.end());
    mkdirs(db);
    Meta meta;
    if (exists(join(db, "meta.json"))) meta = Meta::load(db);
    // the device is first touched after the first batch parsed (a parse
    // error is reported without one)
    Stream st;
    SortScratch scratch;
    DevBuf<uint32_t> d_bucket, d_start, d_end, d_sid, d_vals, d_osid;
    DevBuf<uint64_t> d_keys;
    DevBuf<uint2> d_pos;
    const unsigned threads = host_threads();
    const size_t bs = std::max<size_t>(1, batch_size ? batch_size : beds.size());
    for (size_t b0 = 0; b0 < beds.size(); b0 += bs) {
        const std::vector<std::string> batch(beds.begin() + b0, beds.begin() + std::min(beds.size(), b0 + bs));
        const std::vector<BedFile> files = parse_beds(batch, threads);
        const uint32_t next = meta.next_sid();
        // shard ids: the DB's shards, then new ones in first-seen order
        std::vector<std::string> names = meta.shards;
        std::unordered_map<std::string, uint32_t> id;
        for (uint32_t s = 0; s < names.size(); s++) id[names[s]] = s;
        uint64_t n = 0;
        uint32_t maxl = 0;
        for (auto& f : files) {
            n += f.n;
            for (size_t s = 0; s < f.shards.size(); s++) {
                if (!id.count(f.shards[s])) {
                    id[f.shards[s]] = (uint32_t)names.size();
                    names.push_back(f.shards[s]);
                }
                for (auto& iv : f.ivs[s]) maxl = std::max(maxl, meta.lc.layer_for_size(iv.end - iv.start));
            }
        }
        CHK_REQUIRE(n < (1ull << 32), CHK_ERR_INVALID, "an add batch holds at most 2^32-1 intervals");
        const uint32_t NL = maxl + 1;
        const uint64_t nbuckets = (uint64_t)names.size() * NL;
        CHK_REQUIRE(nbuckets < (1ull << 32), CHK_ERR_INVALID, "too many (shard, layer) buckets");
        std::vector<uint32_t> bucket(n), start(n), end(n), sid(n);
        std::vector<uint64_t> boff(nbuckets + 1, 0);
        {
            uint64_t k = 0;
            for (size_t i = 0; i < files.size(); i++)
                for (size_t s = 0; s < files[i].shards.size(); s++) {
                    const uint32_t sh = id[files[i].shards[s]];
                    for (auto& iv : files[i].ivs[s]) {
                        const uint32_t bk = sh * NL + meta.lc.layer_for_size(iv.end - iv.start);
                        bucket[k] = bk;
                        start[k] = iv.start;
                        end[k] = iv.end;
                        sid[k] = next + (uint32_t)i;
                        boff[bk + 1]++;
                        k++;
                    }
                }
        }
        for (uint64_t b = 0; b < nbuckets; b++) boff[b + 1] += boff[b];
        std::vector<Pos> pos(n);
        std::vector<uint32_t> osid(n);
        if (n) {
            if (!st.s) {
                CHK_HIP(hipSetDevice(device));
                CHK_HIP(hipStreamCreateWithFlags(&st.s, hipStreamNonBlocking));
            }
            d_bucket.reserve(n);
            d_start.reserve(n);
            d_end.reserve(n);
            d_sid.reserve(n);
            d_keys.reserve(n);
            d_vals.reserve(n);
            d_pos.reserve(n);
            d_osid.reserve(n);
            CHK_HIP(hipMemcpyAsync(d_bucket.p, bucket.data(), n * 4, hipMemcpyHostToDevice, st.s));
            CHK_HIP(hipMemcpyAsync(d_start.p, start.data(), n * 4, hipMemcpyHostToDevice, st.s));
            CHK_HIP(hipMemcpyAsync(d_end.p, end.data(), n * 4, hipMemcpyHostToDevice, st.s));
            CHK_HIP(hipMemcpyAsync(d_sid.p, sid.data(), n * 4, hipMemcpyHostToDevice, st.s));
            const unsigned g = (unsigned)((n + NT - 1) / NT);
            hipLaunchKernelGGL(add_keys, dim3(g), dim3(NT), 0, st.s, d_bucket.p, d_start.p, n, d_keys.p,
                               d_vals.p);
            CHK_HIP(hipGetLastError());
            radix_sort_u64(d_keys.p, d_vals.p, n, 32 + bits_for(nbuckets - 1), scratch, st.s);
            hipLaunchKernelGGL(add_gather, dim3(g), dim3(NT), 0, st.s, d_vals.p, d_start.p, d_end.p,
                               d_sid.p, n, d_pos.p, d_osid.p);
            CHK_HIP(hipGetLastError());
            CHK_HIP(hipMemcpyAsync(pos.data(), d_pos.p, n * sizeof(Pos), hipMemcpyDeviceToHost, st.s));
            CHK_HIP(hipMemcpyAsync(osid.data(), d_osid.p, n * 4, hipMemcpyDeviceToHost, st.s));
            CHK_HIP(hipStreamSynchronize(st.s));
        }
        std::vector<uint64_t> used;
        for (uint64_t b = 0; b < nbuckets; b++)
            if (boff[b + 1] > boff[b]) used.push_back(b);
        std::atomic<size_t> nextb{0};
        std::vector<std::exception_ptr> err(threads);
        std::vector<std::thread> th;
        for (unsigned t = 0; t < std::min<size_t>(threads, std::max<size_t>(1, used.size())); t++)
            th.emplace_back([&, t] {
                try {
                    for (size_t k; (k = nextb.fetch_add(1)) < used.size();) {
                        const uint64_t b = used[k];
                        const uint32_t layer = (uint32_t)(b % NL);
                        write_bucket(db, names[b / NL], layer, meta.lc.tile_size(layer), pos.data() + boff[b],
                                     osid.data() + boff[b], boff[b + 1] - boff[b]);
                    }
                } catch (...) {
                    err[t] = std::current_exception();
                    nextb = used.size();
                }
            });
        for (auto& x : th) x.join();
        for (auto& e : err)
            if (e) std::rethrow_exception(e);
        for (size_t i = 0; i < batch.size(); i++) meta.sid_map[next + (uint32_t)i] = basename_of(batch[i]);
        uint32_t max_layer = meta.num_layers ? meta.num_layers - 1 : 0;
        for (uint64_t b : used) {
            const std::string& sh = names[b / NL];
            if (std::find(meta.shards.begin(), meta.shards.end(), sh) == meta.shards.end())
                meta.shards.push_back(sh);
            max_layer = std::max(max_layer, (uint32_t)(b % NL));
        }
        meta.num_layers = max_layer + 1;
        meta.save(db);
    }
    if (stats) {
        // moment tables of every file on the device, from the written layers
        const Flat f = load_intervals(db, meta);
        Index ix;
        index_init(ix, device, (uint32_t)meta.shards.size(), std::max<uint32_t>(1, meta.next_sid()),
                   f.start.size());
        index_build(ix, f.shard.data(), f.start.data(), f.end.data(), f.sid.data());
        const auto g = genome_for(meta.shards, chroms.empty() ? nullptr : &chroms);
        set_genome(ix, g.data());
        moments_build(ix);
        save_moments(db, ix, meta);
    }
}

}  // namespace db
}  // namespace chk
