// Test infrastructure: drives the solve server's C ABI from a client thread
// the way an open-loop serving benchmark does (Poisson arrivals, a mix of
// sizes, radii and storage types), then checks that every accepted request
// was answered once with its analytic norm 2(K+1) (written by the HIP
// stand-in in mock_hip.cc) and that the pool is empty at the end.  A second
// phase keeps some outputs, fetches and releases them while the worker runs,
// under a budget small enough to make requests wait for memory.  A third
// phase fetches kept outputs on one thread while another releases them.  A
// fourth makes one window fail to instantiate: the server must report the
// error, refuse new requests, and free every request it holds when it is
// destroyed (LeakSanitizer checks that at exit).
//
// usage: server_driver <rate per s> <seconds> <budget GiB> <check every entry 0|1>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <map>
#include <random>
#include <thread>
#include <vector>

#include "prk_stencil.h"

namespace prk {
extern long long g_items, g_final;
extern int g_check_all;
extern long long g_instantiations, g_fail_instantiation, g_updates, g_error_launches;
}  // namespace prk

static prk_solve random_solve(std::mt19937_64 &rng, int nmin, int nmax, int rmax = 4)
{
    std::uniform_real_distribution<double> u(std::log((double)nmin), std::log((double)nmax));
    prk_solve s{};
    s.pattern = (int)(rng() % 2);
    s.radius = 1 + (int)(rng() % rmax);
    s.n = (int)std::lround(std::exp(u(rng)));
    s.iterations = 1 + (int)(rng() % 200);
    const int d = (int)(rng() % 10);
    s.dtype = d < 6 ? PRK_F32 : d < 9 ? PRK_F64 : PRK_BF16;
    s.init = PRK_INIT_SEEDED;
    s.seed = rng();
    return s;
}

static int check_results(prk_server *s, const std::map<uint64_t, prk_solve> &ids, const char *phase)
{
    std::vector<prk_server_result> res(ids.size() + 16);
    int cnt = 0;
    const int rc = prk_server_poll(s, res.data(), (int)res.size(), &cnt);
    int bad = 0;
    std::map<uint64_t, int> seen;
    for (int i = 0; i < cnt; ++i) {
        auto it = ids.find(res[i].id);
        if (it == ids.end() || seen[res[i].id]++ || res[i].norm != 2.0 * (it->second.iterations + 1) ||
            !(res[i].t_submit <= res[i].t_admit && res[i].t_admit <= res[i].t_done))
            ++bad;
    }
    std::printf("%s: poll rc %d, accepted %zu, answered %d, wrong %d\n", phase, rc, ids.size(), cnt, bad);
    return (rc || bad || cnt != (int)ids.size()) ? 1 : 0;
}

int main(int argc, char **argv)
{
    const double rate = argc > 1 ? std::atof(argv[1]) : 500;
    const double secs = argc > 2 ? std::atof(argv[2]) : 2;
    const double budget_gb = argc > 3 ? std::atof(argv[3]) : 200;
    prk::g_check_all = argc > 4 ? std::atoi(argv[4]) : 1;
    int fails = 0;

    // phase 1: open-loop arrivals, nothing kept
    {
        std::mt19937_64 rng(11);
        std::exponential_distribution<double> gap(rate);
        prk_server_config cfg{(long long)(budget_gb * (1ll << 30)), 256, 4096, 16384};
        prk_server *s = nullptr;
        if (prk_server_create(&cfg, &s)) {
            std::printf("create failed: %s\n", prk_last_error());
            return 1;
        }
        std::map<uint64_t, prk_solve> ids;
        int refused = 0;
        const double t0 = prk_server_clock();
        for (double t = gap(rng); t < secs; t += gap(rng)) {
            while (prk_server_clock() < t0 + t) std::this_thread::sleep_for(std::chrono::microseconds(100));
            const prk_solve sv = random_solve(rng, 128, 8192);
            uint64_t id = 0;
            if (prk_server_submit(s, &sv, 0, &id) == PRK_OK) ids[id] = sv;
            else ++refused;
        }
        const int drc = prk_server_drain(s, 600);
        fails += drc != PRK_OK;
        fails += check_results(s, ids, "open loop");
        prk_server_stats st{};
        prk_server_stats_get(s, &st);
        std::printf("open loop: refused %d, windows %lld, launches %lld, graph updates %lld, pool in use %lld\n",
                    refused, st.windows, st.launches, st.graph_updates, st.pool_used);
        fails += st.pool_used != 0 || st.served != (long long)ids.size();
        prk_server_destroy(s);
    }
    // phase 2: tight budget, kept outputs fetched and released while running
    {
        std::mt19937_64 rng(12);
        prk_server_config cfg{64ll << 20, 64, 4096, 16384};
        prk_server *s = nullptr;
        if (prk_server_create(&cfg, &s)) return 1;
        std::map<uint64_t, prk_solve> ids;
        std::vector<uint64_t> kept;
        std::vector<unsigned char> host(8u << 20);
        int fetched = 0;
        for (int i = 0; i < 400; ++i) {
            const prk_solve sv = random_solve(rng, 17, 700, 8);
            uint64_t id = 0;
            const int keep = (int)(i % 3 == 0);
            const int rc = prk_server_submit(s, &sv, keep, &id);
            if (rc == PRK_OK) {
                ids[id] = sv;
                if (keep) kept.push_back(id);
            } else if (rc != PRK_EBUSY) {
                ++fails;
            }
            // release kept outputs as they finish so the queue keeps moving
            for (size_t k = 0; k < kept.size();) {
                if (prk_server_fetch(s, kept[k], 0, host.data(), host.size()) == PRK_OK) {
                    ++fetched;
                    fails += prk_server_release(s, kept[k]) != PRK_OK;
                    kept.erase(kept.begin() + (long)k);
                } else {
                    ++k;
                }
            }
        }
        while (!kept.empty()) {
            for (size_t k = 0; k < kept.size();) {
                if (prk_server_fetch(s, kept[k], 0, host.data(), host.size()) == PRK_OK) {
                    ++fetched;
                    fails += prk_server_release(s, kept[k]) != PRK_OK;
                    kept.erase(kept.begin() + (long)k);
                } else {
                    ++k;
                }
            }
            std::this_thread::sleep_for(std::chrono::microseconds(200));
        }
        fails += prk_server_drain(s, 600) != PRK_OK;
        fails += check_results(s, ids, "tight budget");
        prk_server_stats st{};
        prk_server_stats_get(s, &st);
        std::printf("tight budget: fetched %d, pool in use %lld\n", fetched, st.pool_used);
        fails += st.pool_used != 0;
        prk_server_destroy(s);
    }
    // phase 3: one thread fetches kept outputs while another releases them
    {
        std::mt19937_64 rng(13);
        prk_server_config cfg{256ll << 20, 256, 4096, 16384};
        prk_server *s = nullptr;
        if (prk_server_create(&cfg, &s)) return 1;
        std::vector<uint64_t> ids;
        for (int i = 0; i < 60; ++i) {
            const prk_solve sv = random_solve(rng, 16, 300);
            uint64_t id = 0;
            if (prk_server_submit(s, &sv, 1, &id) == PRK_OK) ids.push_back(id);
        }
        fails += prk_server_drain(s, 600) != PRK_OK;
        int fetched = 0, gone = 0, released = 0;
        std::thread reader([&] {
            std::vector<unsigned char> host(1u << 20);
            for (int round = 0; round < 3; ++round)
                for (uint64_t id : ids) {
                    const int rc = prk_server_fetch(s, id, round & 1, host.data(), host.size());
                    if (rc == PRK_OK) ++fetched;
                    else if (rc == PRK_ENOTFOUND) ++gone;
                }
        });
        for (uint64_t id : ids) released += prk_server_release(s, id) == PRK_OK;
        reader.join();
        prk_server_stats st{};
        prk_server_stats_get(s, &st);
        std::printf("fetch/release race: fetched %d, already released %d, released %d of %zu, pool in use %lld\n",
                    fetched, gone, released, ids.size(), st.pool_used);
        fails += released != (int)ids.size() || st.pool_used != 0 || fetched + gone != 3 * (int)ids.size();
        fails += prk_server_release(s, ids.empty() ? 0 : ids[0]) != PRK_ENOTFOUND;
        prk_server_destroy(s);
    }
    // phase 4: the third window of a server fails to instantiate
    {
        std::mt19937_64 rng(17);
        prk_server_config cfg{1ll << 30, 256, 4096, 16384};
        prk_server *s = nullptr;
        if (prk_server_create(&cfg, &s)) return 1;
        prk::g_fail_instantiation = prk::g_instantiations + 3;
        int accepted = 0;
        for (int i = 0; i < 200; ++i) {
            prk_solve sv = random_solve(rng, 16, 200);
            sv.iterations = (int)(rng() % 4);  // finishes inside the window it starts in
            uint64_t id = 0;
            accepted += prk_server_submit(s, &sv, i % 3 == 0, &id) == PRK_OK;
            if (i % 20 == 0) std::this_thread::sleep_for(std::chrono::milliseconds(1));
        }
        const int drc = prk_server_drain(s, 600);
        prk_solve sv = random_solve(rng, 16, 200);
        uint64_t id = 0;
        const int src = prk_server_submit(s, &sv, 0, &id);
        std::printf("failed window: drain rc %d, submit after failure rc %d, accepted %d\n", drc, src, accepted);
        fails += drc != PRK_EHIP || src != PRK_EHIP || prk::g_instantiations < 3;
        prk_server_destroy(s);
        prk::g_fail_instantiation = 0;
    }
    // phase 5: every second FP32/BF16 request of each class also runs in FP64
    // (an internal twin): every request is answered once, the shadowed ones
    // with an error and none of the FP64 ones, the twins never, and the pool
    // is empty once the kept outputs are released
    {
        std::mt19937_64 rng(19);
        prk_server_config cfg{2ll << 30, 256, 4096, 16384, 2};
        prk_server *s = nullptr;
        if (prk_server_create(&cfg, &s)) return 1;
        std::map<uint64_t, prk_solve> ids;
        for (int i = 0; i < 300; ++i) {
            const prk_solve sv = random_solve(rng, 16, 400);
            uint64_t id = 0;
            const int rc = prk_server_submit(s, &sv, i % 5 == 0, &id);
            if (rc == PRK_OK) ids[id] = sv;
            else fails += rc != PRK_EBUSY;
            if (i % 25 == 0) std::this_thread::sleep_for(std::chrono::milliseconds(1));
        }
        fails += prk_server_drain(s, 600) != PRK_OK;
        std::vector<prk_server_result> res(ids.size() + 16);
        int cnt = 0, with_err = 0, bad = 0;
        fails += prk_server_poll(s, res.data(), (int)res.size(), &cnt) != PRK_OK;
        std::map<uint64_t, int> seen;
        for (int i = 0; i < cnt; ++i) {
            auto it = ids.find(res[i].id);
            if (it == ids.end() || seen[res[i].id]++ || res[i].norm != 2.0 * (it->second.iterations + 1)) {
                ++bad;
                continue;
            }
            if (res[i].err_valid) {
                ++with_err;
                bad += it->second.dtype == PRK_F64 || res[i].t_done < res[i].t_admit;
            }
        }
        prk_server_stats st{};
        prk_server_stats_get(s, &st);
        for (auto &kv : ids) (void)prk_server_release(s, kv.first);  // the kept ones
        prk_server_stats end{};
        prk_server_stats_get(s, &end);
        std::printf("shadow solves: accepted %zu, answered %d, wrong %d, with error %d, shadowed %lld, "
                    "pool in use %lld\n", ids.size(), cnt, bad, with_err, st.shadow_solves, end.pool_used);
        fails += bad || cnt != (int)ids.size() || with_err != st.shadow_solves || st.shadow_solves == 0 ||
                 end.pool_used != 0;
        prk_server_destroy(s);
    }
    std::printf("entries %lld, final norms %lld, graph updates %lld, shadow comparisons %lld, failures %d\n",
                prk::g_items, prk::g_final, prk::g_updates, prk::g_error_launches, fails);
    return fails ? 1 : 0;
}
