// pool_stress.cpp -- chk::db::host_run (the library's persistent host worker
// pool, dbfile.cpp) in a CPU build with ThreadSanitizer: six threads call it
// at once with 1 to 16 tasks, tasks call it again from inside, and some calls
// throw; every task must run exactly once, the nested calls must finish, and
// each exception must reach its caller.  Prints "bad=0" and exits 0 when all
// checks hold.
#include <atomic>
#include <cstdio>
#include <stdexcept>
#include <thread>
#include <vector>

#include "dbfile.hpp"

namespace chk {
// the moment-table file functions of dbfile.cpp are not exercised here
void moments_download(Index&, int64_t*, uint32_t*, double*, double*) { throw Error(CHK_ERR_INTERNAL, "unused"); }
void moments_upload(Index&, const int64_t*, const uint32_t*, uint64_t, const double*, const double*) {
    throw Error(CHK_ERR_INTERNAL, "unused");
}
void index_init(Index&, int, uint32_t, uint32_t, uint64_t) { throw Error(CHK_ERR_INTERNAL, "unused"); }
}  // namespace chk

using chk::db::host_run;

int main() {
    std::atomic<long> bad{0};
    auto caller = [&](unsigned seed) {
        for (int it = 0; it < 300; it++) {
            const unsigned T = 1 + (seed * 7 + it) % 16;
            std::vector<std::atomic<int>> runs(T);
            for (auto& r : runs) r = 0;
            host_run(T, [&](unsigned t) {
                runs[t]++;
                if (t % 5 == 0) {  // a call from inside a task
                    std::atomic<long> s{0};
                    host_run(3, [&](unsigned u) { s += u + 1; });
                    if (s != 6) bad++;
                }
            });
            for (auto& r : runs)
                if (r != 1) bad++;
            if (it % 7 == 0) {
                bool caught = false;
                try {
                    host_run(T + 1, [&](unsigned t) {
                        if (t == T / 2) throw std::runtime_error("task failed");
                    });
                } catch (const std::runtime_error&) {
                    caught = true;
                }
                if (!caught) bad++;
            }
        }
    };
    std::vector<std::thread> th;
    for (unsigned s = 0; s < 6; s++) th.emplace_back(caller, s);
    for (auto& x : th) x.join();
    std::printf("bad=%ld\n", bad.load());
    return bad ? 1 : 0;
}
