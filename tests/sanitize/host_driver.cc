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
    // st[0] = qubits, st[1] = measurements. The reference sample's tableau grows with the square of
    // the qubit count; mutants with huge qubit indices only go through parsing and compiling.
    if (rc == 0 && st[0] <= 4096) {
        std::vector<uint8_t> ref((st[1] + 7) / 8 + 1);
        (void)mfs_reference_sample(t.data(), t.size(), ref.data(), ref.size());
        mfs_m2d *m2d = nullptr;
        if (mfs_m2d_compile(t.data(), t.size(), 0, &m2d) == 0) mfs_m2d_destroy(m2d);
        // The measurement sampler's expansion plan, and the host expansion of a few shots where it applies.
        uint64_t xs[8] = {0};
        if (mfs_expand_plan_stats(t.data(), t.size(), xs) == 0 && xs[0] && st[1] < (1u << 16)) {
            const uint64_t shots = 3, Y = st[2] + st[3];
            std::vector<uint8_t> y(shots * (Y + 1), 1), fr(shots * (xs[3] + 1), 1), m(shots * (st[1] + 1));
            (void)mfs_expand_shots_host(t.data(), t.size(), shots, y.data(), Y + 1, fr.data(), xs[3] + 1, m.data(),
                                        st[1] + 1);
        }
        if (full) {
            mfs_sampler *s = nullptr;
            if (mfs_compile(t.data(), t.size(), 1, 1, &s) == 0) mfs_destroy(s);
        }
    }
}

// Seeded mutations of a circuit text: replaced characters drawn from the syntax's alphabet, inserted
// digits (huge counts, indices, probabilities) and deleted spans.
static std::string mutate(const std::string &t, uint64_t &state) {
    static const char alphabet[] = "0123456789 \n\t()[]{}!*^-.,:#rec_XYZMRDETCOBSAPIHNLQ";
    auto next = [&state]() {
        state = state * 6364136223846793005ull + 1442695040888963407ull;
        return (uint32_t)(state >> 33);
    };
    std::string s = t;
    const int edits = 1 + (int)(next() % 4);
    for (int e = 0; e < edits && !s.empty(); e++) {
        const size_t at = next() % s.size();
        switch (next() % 3) {
            case 0: s[at] = alphabet[next() % (sizeof(alphabet) - 1)]; break;
            case 1: s.insert(at, std::string(1 + next() % 20, (char)('0' + next() % 10))); break;
            default: s.erase(at, 1 + next() % 8); break;
        }
    }
    return s;
}

int main(int argc, char **argv) {
    int files = 0, prefixes = 0, mutants = 0;
    uint64_t state = 12345;
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
            if (t.size() < 200000) {
                for (int k = 0; k < 40; k++) {
                    const std::string u = mutate(t, state);
                    (void)mfs_dem_parse_stats(u.data(), u.size(), st);
                    mutants++;
                }
            }
        } else {
            run_circuit(t, true);
            if (i <= 40) {
                for (size_t k = 0; k < t.size() && k < 3000; k += 5) { run_circuit(t.substr(0, k), false); prefixes++; }
            }
            if (t.size() < 20000) {
                for (int k = 0; k < 20; k++) { run_circuit(mutate(t, state), false); mutants++; }
            }
        }
        files++;
    }
    printf("files %d prefixes %d mutants %d\n", files, prefixes, mutants);
    return 0;
}
