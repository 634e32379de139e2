// Host-code sanitizer driver (test infrastructure, built by tests/test_host_sanitizers.py with
// AddressSanitizer and UndefinedBehaviorSanitizer on the host objects). For every input file it runs
// the host-only entry points: circuit parse + compile statistics, the reference sample, the m2d
// converter's column build, and the sampler compile (which stops at the device check on a machine
// without a GPU); DEM files go through the DEM parser. Prefixes of the first 40 files exercise the
// parsers' error paths.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>
#include "stim_amd.h"

static void run_circuit(const std::string &t, bool full) {
    uint64_t st[10] = {0};
    int rc = mfs_parse_stats(t.data(), t.size(), st);
    if (rc == 0) {
        std::vector<uint8_t> ref(st[0] + 8);
        (void)mfs_reference_sample(t.data(), t.size(), ref.data(), ref.size());
        if (full) {
            mfs_m2d *c = nullptr;
            if (mfs_m2d_compile(t.data(), t.size(), 0, &c) == 0) mfs_m2d_destroy(c);
            mfs_sampler *s = nullptr;
            if (mfs_compile(t.data(), t.size(), 1, 1, &s) == 0) mfs_destroy(s);
        }
    }
}

int main(int argc, char **argv) {
    int files = 0, prefixes = 0;
    for (int i = 1; i < argc; i++) {
        std::ifstream f(argv[i]);
        std::stringstream ss;
        ss << f.rdbuf();
        const std::string t = ss.str();
        const bool dem = strstr(argv[i], ".dem") != nullptr;
        if (dem) {
            uint64_t st[3];
            (void)mfs_dem_parse_stats(t.data(), t.size(), st);
            for (size_t k = 0; k < t.size() && k < 4000; k += 7) (void)mfs_dem_parse_stats(t.data(), k, st);
        } else {
            run_circuit(t, true);
            if (i <= 40) {
                for (size_t k = 0; k < t.size() && k < 3000; k += 5) { run_circuit(t.substr(0, k), false); prefixes++; }
            }
        }
        files++;
    }
    printf("files %d prefixes %d\n", files, prefixes);
    return 0;
}
