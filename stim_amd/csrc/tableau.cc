// Stabilizer simulation for reference samples (see tableau.h).
//
// Form: the inverse tableau of the circuit's Clifford U (the rows U^dagger X_q U and U^dagger Z_q U,
// each an n-qubit Pauli string with a sign). A gate rewrites only its own qubits' rows (products of
// old rows), a deterministic Z measurement reads one row's sign, and a random one is resolved by a
// few input-side column operations (details at class Tableau). The reference keeps the same kind
// of state (TableauSimulator::inv_state, tableau_simulator.h:41). Gates act through conjugation
// tables (U P U^dagger for every local Pauli P, with sign), built from the images of X and Z that
// define each gate (doc/gates.md, "Stabilizer Generators"). A random result is always 0 (the
// reference's sign bias +1). The random results an instruction collapses are taken in ascending
// qubit order and read afterwards in target order, as the reference does
// (tableau_simulator.inl:330-504, 1267-1290): the order decides the sample when random results are
// correlated. Pauli-product measurements and phases rotate every factor to Z, fold the product onto
// its lowest qubit with CX, act there and undo the rotation (src/stim/circuit/gate_decomposition.cc:
// 88-170, 163-217, 245-274).
#include "tableau.h"

#include <algorithm>
#include <array>
#include <stdexcept>
#include <string>

namespace sa {
namespace {

// Single-qubit Pauli codes: bit 0 = X component, bit 1 = Z component (0 = I, 1 = X, 2 = Z, 3 = Y).
// Exponent k in P * Q = i^k (P xor Q): XY = iZ, YZ = iX, ZX = iY, and the reversed products give -i.
int pauli_phase(int p, int q) {
    if (p == 0 || q == 0 || p == q) return 0;
    if ((p == 1 && q == 3) || (p == 3 && q == 2) || (p == 2 && q == 1)) return 1;
    return 3;
}

int code_of(char c) {
    switch (c) {
        case '_': return 0;
        case 'X': return 1;
        case 'Z': return 2;
        case 'Y': return 3;
    }
    throw std::logic_error("bad Pauli letter in a gate table");
}

struct Map1 {
    uint8_t out[4];
    uint8_t neg[4];
};

struct Map2 {
    uint8_t oa[16], ob[16], neg[16];   // index: code at a | code at b << 2
};

// "X -> x_image", "Z -> z_image" with images like "-Y".
Map1 make1(const char *xi, const char *zi) {
    const bool nx = xi[0] == '-', nz = zi[0] == '-';
    const int px = code_of(xi[nx]), pz = code_of(zi[nz]);
    Map1 m{};
    m.out[1] = (uint8_t)px;
    m.neg[1] = nx;
    m.out[2] = (uint8_t)pz;
    m.neg[2] = nz;
    // Y = i X Z, so its image is i (X image)(Z image).
    const int e = (1 + pauli_phase(px, pz) + 2 * (nx + nz)) & 3;
    if ((e & 1) || px == pz || !px || !pz) throw std::logic_error("single-qubit gate table is not a Clifford");
    m.out[3] = (uint8_t)(px ^ pz);
    m.neg[3] = e == 2;
    return m;
}

struct P2 {
    int a, b;
    bool neg;
};

P2 parse2(const char *s) {
    const bool neg = s[0] == '-';
    return {code_of(s[neg]), code_of(s[neg + 1]), neg};
}

bool anticommute(const P2 &p, const P2 &q) {
    const int k = (pauli_phase(p.a, q.a) & 1) + (pauli_phase(p.b, q.b) & 1);
    return k & 1;
}

// Images of X_, Z_, _X, _Z (qubit a first).
Map2 make2(const char *xa, const char *za, const char *xb, const char *zb) {
    const P2 gen[4] = {parse2(xa), parse2(za), parse2(xb), parse2(zb)};
    // The images must keep the Pauli commutation relations.
    for (int i = 0; i < 4; i++) {
        for (int j = i + 1; j < 4; j++) {
            const bool want = (i == 0 && j == 1) || (i == 2 && j == 3);
            if (anticommute(gen[i], gen[j]) != want) throw std::logic_error("two-qubit gate table is not a Clifford");
        }
    }
    Map2 m{};
    for (int c = 0; c < 16; c++) {
        const int ca = c & 3, cb = c >> 2;
        // P = i^{[ca == Y] + [cb == Y]} X_a^x Z_a^z X_b^x Z_b^z, mapped factor by factor.
        int e = (ca == 3) + (cb == 3);
        int ra = 0, rb = 0;
        const int use[4] = {ca & 1, ca >> 1, cb & 1, cb >> 1};
        for (int g = 0; g < 4; g++) {
            if (!use[g]) continue;
            e += pauli_phase(ra, gen[g].a) + pauli_phase(rb, gen[g].b) + 2 * gen[g].neg;
            ra ^= gen[g].a;
            rb ^= gen[g].b;
        }
        e &= 3;
        if (e & 1) throw std::logic_error("two-qubit gate table is not a Clifford");
        m.oa[c] = (uint8_t)ra;
        m.ob[c] = (uint8_t)rb;
        m.neg[c] = e == 2;
    }
    return m;
}

struct Tables {
    std::array<Map1, (size_t)Gate::NUM_GATES> one{};
    std::array<Map2, (size_t)Gate::NUM_GATES> two{};
    std::array<bool, (size_t)Gate::NUM_GATES> has1{}, has2{};

    void def1(Gate g, const char *x, const char *z) {
        one[(size_t)g] = make1(x, z);
        has1[(size_t)g] = true;
    }
    void def2(Gate g, const char *xa, const char *za, const char *xb, const char *zb) {
        two[(size_t)g] = make2(xa, za, xb, zb);
        has2[(size_t)g] = true;
    }

    // Gate definitions: images of the single-qubit (X, Z) and two-qubit (X_, Z_, _X, _Z) generators
    // under conjugation, as listed in doc/gates.md.
    Tables() {
        def1(Gate::I, "X", "Z");
        def1(Gate::X, "X", "-Z");
        def1(Gate::Y, "-X", "-Z");
        def1(Gate::Z, "-X", "Z");
        def1(Gate::H, "Z", "X");
        def1(Gate::H_XY, "Y", "-Z");
        def1(Gate::H_YZ, "-X", "Y");
        def1(Gate::H_NXY, "-Y", "-Z");
        def1(Gate::H_NXZ, "-Z", "-X");
        def1(Gate::H_NYZ, "-X", "-Y");
        def1(Gate::C_XYZ, "Y", "X");
        def1(Gate::C_NXYZ, "-Y", "-X");
        def1(Gate::C_XNYZ, "-Y", "X");
        def1(Gate::C_XYNZ, "Y", "-X");
        def1(Gate::C_ZYX, "Z", "Y");
        def1(Gate::C_ZYNX, "-Z", "Y");
        def1(Gate::C_ZNYX, "Z", "-Y");
        def1(Gate::C_NZYX, "-Z", "-Y");
        def1(Gate::SQRT_X, "X", "-Y");
        def1(Gate::SQRT_X_DAG, "X", "Y");
        def1(Gate::SQRT_Y, "-Z", "X");
        def1(Gate::SQRT_Y_DAG, "Z", "-X");
        def1(Gate::S, "Y", "Z");
        def1(Gate::S_DAG, "-Y", "Z");

        def2(Gate::CX, "XX", "Z_", "_X", "ZZ");
        def2(Gate::CY, "XY", "Z_", "ZX", "ZZ");
        def2(Gate::CZ, "XZ", "Z_", "ZX", "_Z");
        def2(Gate::XCX, "X_", "ZX", "_X", "XZ");
        def2(Gate::XCY, "X_", "ZY", "XX", "XZ");
        def2(Gate::XCZ, "X_", "ZZ", "XX", "_Z");
        def2(Gate::YCX, "XX", "ZX", "_X", "YZ");
        def2(Gate::YCY, "XY", "ZY", "YX", "YZ");
        def2(Gate::YCZ, "XZ", "ZZ", "YX", "_Z");
        def2(Gate::SWAP, "_X", "_Z", "X_", "Z_");
        def2(Gate::ISWAP, "ZY", "_Z", "YZ", "Z_");
        def2(Gate::ISWAP_DAG, "-ZY", "_Z", "-YZ", "Z_");
        def2(Gate::CXSWAP, "XX", "_Z", "X_", "ZZ");
        def2(Gate::SWAPCX, "_X", "ZZ", "XX", "Z_");
        def2(Gate::CZSWAP, "ZX", "_Z", "XZ", "Z_");
        def2(Gate::SQRT_XX, "X_", "-YX", "_X", "-XY");
        def2(Gate::SQRT_XX_DAG, "X_", "YX", "_X", "XY");
        def2(Gate::SQRT_YY, "-ZY", "XY", "-YZ", "YX");
        def2(Gate::SQRT_YY_DAG, "ZY", "-XY", "YZ", "-YX");
        def2(Gate::SQRT_ZZ, "YZ", "Z_", "ZY", "_Z");
        def2(Gate::SQRT_ZZ_DAG, "-YZ", "Z_", "-ZY", "_Z");
        def2(Gate::II, "X_", "Z_", "_X", "_Z");
    }
};

const Tables &tables() {
    static const Tables t;
    return t;
}

// Inverse of a single-qubit gate's conjugation map: U P U^dagger = (-1)^neg Q gives
// U^dagger Q U = (-1)^neg P.
Map1 inverse1(const Map1 &m) {
    Map1 r{};
    for (int c = 1; c < 4; c++) {
        r.out[m.out[c]] = (uint8_t)c;
        r.neg[m.out[c]] = m.neg[c];
    }
    return r;
}

Map2 inverse2(const Map2 &m) {
    Map2 r{};
    for (int c = 1; c < 16; c++) {
        const int d = m.oa[c] | (m.ob[c] << 2);
        r.oa[d] = (uint8_t)(c & 3);
        r.ob[d] = (uint8_t)(c >> 2);
        r.neg[d] = m.neg[c];
    }
    return r;
}

// The state U|0...0> held as the inverse tableau of U: row X'_q = U^dagger X_q U and row Z'_q =
// U^dagger Z_q U, each an n-qubit Pauli string (x bits, z bits, 64 qubits per word) with a sign.
// Applying a gate G (U -> G U) rewrites only its qubits' rows: X'_q becomes the image of
// G^dagger X_q G, a product of the old rows of G's qubits. Z_q is deterministic in U|0> exactly when
// Z'_q has no X part (it is then a signed product of Z's, which fix |0...0>), and its value is the
// sign. A random Z_q measurement picks a pivot input qubit p where Z'_q has X, clears Z'_q's other X
// components with CX gates from p on the input side (U -> U C with C fixing |0...0>, so the state is
// unchanged), turns p's X (or Y) into Z with H (or H_YZ) on the input side, which moves the state onto
// the measured eigenspace, and picks result 0 with an X on input p when the sign says 1. Input-side
// gates conjugate every row on the input qubits' columns.
class Tableau {
public:
    explicit Tableau(uint32_t n)
        : n_(n), nw_((n + 63) / 64), x_(2 * (size_t)n * nw_, 0), z_(x_.size(), 0), r_(2 * (size_t)n, 0),
          sx_(4 * nw_), sz_(4 * nw_) {
        for (uint32_t q = 0; q < n; q++) {
            set(x_, 2 * (size_t)q, q, true);
            set(z_, 2 * (size_t)q + 1, q, true);
        }
    }

    // Output-side gate: U -> G U. m is G's conjugation map (G P G^dagger).
    void apply1(const Map1 &m, uint32_t q) {
        const Map1 inv = inverse1(m);
        const size_t rx = 2 * (size_t)q, rz = rx + 1;
        const uint8_t sx = image1(inv.out[1], rx, rz, &sx_[0], &sz_[0]) ^ inv.neg[1];
        const uint8_t sz = image1(inv.out[2], rx, rz, &sx_[nw_], &sz_[nw_]) ^ inv.neg[2];
        store(rx, &sx_[0], &sz_[0], sx);
        store(rz, &sx_[nw_], &sz_[nw_], sz);
    }

    void apply2(const Map2 &m, uint32_t a, uint32_t b) {
        const Map2 inv = inverse2(m);
        const size_t rows[4] = {2 * (size_t)a, 2 * (size_t)a + 1, 2 * (size_t)b, 2 * (size_t)b + 1};
        const int codes[4] = {1, 2, 1 << 2, 2 << 2};   // X_a, Z_a, X_b, Z_b
        uint8_t sg[4];
        bool same[4];
        for (int k = 0; k < 4; k++) {
            const int c = codes[k];
            // A generator the gate leaves alone keeps its row (CX keeps Z_a and X_b).
            same[k] = (inv.oa[c] | (inv.ob[c] << 2)) == c && !inv.neg[c];
            if (!same[k]) sg[k] = image2(inv.oa[c], inv.ob[c], rows, &sx_[k * nw_], &sz_[k * nw_]) ^ inv.neg[c];
        }
        for (int k = 0; k < 4; k++)
            if (!same[k]) store(rows[k], &sx_[k * nw_], &sz_[k * nw_], sg[k]);
    }

    // Measures Z on qubit a; a random result is taken as 0 and the state collapses onto it.
    bool measure_z(uint32_t a) {
        const size_t row = 2 * (size_t)a + 1;
        const uint64_t *zx = &x_[row * nw_];
        size_t w = 0;
        while (w < nw_ && !zx[w]) w++;
        if (w == nw_) return r_[row];
        random_++;
        const uint32_t p = (uint32_t)(64 * w + __builtin_ctzll(zx[w]));
        const Tables &T = tables();
        std::copy_n(zx, nw_, sx_.data());
        sx_[p >> 6] &= ~(1ull << (p & 63));
        cx_fan_out(p, sx_.data());
        input1(T.one[(size_t)(get(z_, row, p) ? Gate::H_YZ : Gate::H)], p);
        if (r_[row]) input1(T.one[(size_t)Gate::X], p);
        return false;
    }

    // Measurements so far whose result was random (collapsed onto 0).
    uint64_t random_collapses() const { return random_; }

private:
    uint32_t n_;
    size_t nw_;
    uint64_t random_ = 0;
    std::vector<uint64_t> x_, z_;
    std::vector<uint8_t> r_;
    std::vector<uint64_t> sx_, sz_;   // new rows of the gate being applied

    bool get(const std::vector<uint64_t> &v, size_t row, uint32_t q) const {
        return (v[row * nw_ + (q >> 6)] >> (q & 63)) & 1;
    }
    void set(std::vector<uint64_t> &v, size_t row, uint32_t q, bool b) {
        uint64_t &w = v[row * nw_ + (q >> 6)];
        const uint64_t bit = 1ull << (q & 63);
        w = b ? (w | bit) : (w & ~bit);
    }
    void store(size_t row, const uint64_t *x, const uint64_t *z, uint8_t sign) {
        std::copy_n(x, nw_, &x_[row * nw_]);
        std::copy_n(z, nw_, &z_[row * nw_]);
        r_[row] = sign;
    }

    // acc (x, z, exponent e of i^e) <- acc * row; returns the new exponent.
    int mul_row(uint64_t *ax, uint64_t *az, int e, size_t row) const {
        const uint64_t *rx = &x_[row * nw_], *rz = &z_[row * nw_];
        e += 2 * r_[row];
        for (size_t w = 0; w < nw_; w++) {
            const uint64_t x1 = ax[w], z1 = az[w], x2 = rx[w], z2 = rz[w];
            const uint64_t X1 = x1 & ~z1, Y1 = x1 & z1, Z1 = ~x1 & z1;
            const uint64_t X2 = x2 & ~z2, Y2 = x2 & z2, Z2 = ~x2 & z2;
            e += __builtin_popcountll((X1 & Y2) | (Y1 & Z2) | (Z1 & X2));
            e -= __builtin_popcountll((X1 & Z2) | (Y1 & X2) | (Z1 & Y2));
            ax[w] = x1 ^ x2;
            az[w] = z1 ^ z2;
        }
        return e & 3;
    }

    // Sign of the real product i^e (e even for a Hermitian image).
    static uint8_t sign_of(int e) {
        if (e & 1) throw std::logic_error("non-Hermitian Pauli product in the tableau");
        return (uint8_t)((e & 3) == 2);
    }

    // Image under the inverse tableau of a single-qubit Pauli code c on the qubit whose rows are
    // (rx, rz): X -> row X', Z -> row Z', Y = i X Z -> i X' Z'.
    uint8_t image1(int c, size_t rx, size_t rz, uint64_t *ox, uint64_t *oz) const {
        std::fill_n(ox, nw_, 0);
        std::fill_n(oz, nw_, 0);
        int e = c == 3 ? 1 : 0;
        if (c & 1) e = mul_row(ox, oz, e, rx);
        if (c & 2) e = mul_row(ox, oz, e, rz);
        return sign_of(e);
    }

    // Image of the two-qubit Pauli (code ca on a, cb on b) = i^{#Y} X_a^. Z_a^. X_b^. Z_b^.
    uint8_t image2(int ca, int cb, const size_t rows[4], uint64_t *ox, uint64_t *oz) const {
        std::fill_n(ox, nw_, 0);
        std::fill_n(oz, nw_, 0);
        int e = (ca == 3) + (cb == 3);
        if (ca & 1) e = mul_row(ox, oz, e, rows[0]);
        if (ca & 2) e = mul_row(ox, oz, e, rows[1]);
        if (cb & 1) e = mul_row(ox, oz, e, rows[2]);
        if (cb & 2) e = mul_row(ox, oz, e, rows[3]);
        return sign_of(e);
    }

    // Input-side gate (U -> U C): every row is conjugated by C on the input columns. m is C's map;
    // C is self-inverse for every use here (H, H_YZ, X).
    void input1(const Map1 &m, uint32_t q) {
        const size_t w = q >> 6;
        const uint64_t bit = 1ull << (q & 63);
        for (size_t i = 0; i < 2 * (size_t)n_; i++) {
            uint64_t &xw = x_[i * nw_ + w], &zw = z_[i * nw_ + w];
            const int c = ((xw & bit) ? 1 : 0) | ((zw & bit) ? 2 : 0);
            if (!c) continue;
            const int o = m.out[c];
            xw = (o & 1) ? (xw | bit) : (xw & ~bit);
            zw = (o & 2) ? (zw | bit) : (zw & ~bit);
            r_[i] ^= m.neg[c];
        }
    }

    // Input-side CX from control p to every target in `mask` (p not in it), in one pass over the rows.
    // Per row: x_k ^= x_p for k in mask, z_p ^= parity(z & mask), and the sign flips as the CX gates
    // applied one after another would flip it (each CX(p, k) flips it when x_p z_k (x_k + z_p + 1) is
    // odd, with z_p already updated by the earlier targets). With a = z & mask, b = x & mask,
    // m = |a|, that is x_p (m + |a & b| + z_p m + m (m - 1) / 2) mod 2.
    void cx_fan_out(uint32_t p, const uint64_t *mask) {
        const size_t wp = p >> 6;
        const uint64_t bp = 1ull << (p & 63);
        std::vector<uint32_t> words;   // the mask's nonzero words (usually a few)
        for (size_t w = 0; w < nw_; w++)
            if (mask[w]) words.push_back((uint32_t)w);
        if (words.empty()) return;
        for (size_t i = 0; i < 2 * (size_t)n_; i++) {
            uint64_t *xr = &x_[i * nw_], *zr = &z_[i * nw_];
            const bool xp = (xr[wp] & bp) != 0, zp = (zr[wp] & bp) != 0;
            uint64_t m = 0, ab = 0;
            for (uint32_t w : words) {
                const uint64_t a = zr[w] & mask[w];
                m += __builtin_popcountll(a);
                ab += __builtin_popcountll(a & xr[w]);
                if (xp) xr[w] ^= mask[w];
            }
            if (m & 1) zr[wp] ^= bp;
            if (xp) r_[i] ^= (uint8_t)((m + ab + (zp ? m : 0) + m * (m - 1) / 2) & 1);
        }
    }
};

// A '*'-joined Pauli product: combined factors in order of first appearance and the sign (inverted
// factors and phases of repeated qubits).
struct Product {
    std::vector<std::pair<uint32_t, int>> terms;
    bool negate = false;
};

struct Runner {
    Tableau tab;
    std::vector<uint8_t> rec;
    uint64_t random_recorded = 0;   // random collapses whose result went into the record
    const Tables &T = tables();

    explicit Runner(uint32_t n) : tab(n) {}

    void g1(Gate g, uint32_t q) { tab.apply1(T.one[(size_t)g], q); }
    void g2(Gate g, uint32_t a, uint32_t b) { tab.apply2(T.two[(size_t)g], a, b); }

    uint8_t bit_of(const Target &t) const {
        if (t.kind == TK::REC) {
            if (t.value == 0 || t.value > rec.size())
                throw std::out_of_range("Referred to a measurement record before the beginning of time.");
            return rec[rec.size() - t.value];
        }
        return 0;   // sweep bits are 0 in the reference sample
    }

    // basis: 1 = X, 2 = Z, 3 = Y. Rotates the basis onto Z (H for X, H_YZ for Y; both self-inverse).
    void rotate(uint32_t q, int basis) {
        if (basis == 1) g1(Gate::H, q);
        if (basis == 3) g1(Gate::H_YZ, q);
    }

    // Collapses the given qubits in the basis (a random result becomes 0); results read afterwards are
    // deterministic. The order matters when random results are correlated: Z-basis collapses go in
    // target order, X- and Y-basis collapses in ascending qubit order (tableau_simulator.inl:1267-1338).
    void collapse(std::vector<uint32_t> qs, int basis) {
        if (basis != 2) {
            std::sort(qs.begin(), qs.end());
            qs.erase(std::unique(qs.begin(), qs.end()), qs.end());
        }
        for (uint32_t q : qs) {
            rotate(q, basis);
            (void)tab.measure_z(q);
            rotate(q, basis);
        }
    }

    // Reads a (collapsed) qubit in the basis; with reset, forces it to the basis' +1 eigenstate.
    bool read(uint32_t q, int basis, bool reset) {
        rotate(q, basis);
        const bool b = tab.measure_z(q);
        if (reset && b) g1(Gate::X, q);
        rotate(q, basis);
        return b;
    }

    // M / MR / R in one basis: collapse every target, then record (and reset) in target order.
    void measure_targets(const std::vector<Target> &ts, int basis, bool record, bool reset) {
        std::vector<uint32_t> qs;
        for (const Target &t : ts) qs.push_back(t.value);
        const uint64_t r0 = tab.random_collapses();
        collapse(qs, basis);
        if (record) random_recorded += tab.random_collapses() - r0;
        for (const Target &t : ts) {
            const bool b = read(t.value, basis, reset);
            if (record) rec.push_back(b ^ t.inverted);
        }
    }

    // Measures a group of qubit-disjoint products: each is rotated to Z and folded onto its first
    // factor's qubit with CX; the pivots are collapsed together in the collapse basis' order and read
    // in product order.
    void measure_group(const std::vector<Product> &group, int collapse_basis = 2) {
        const uint64_t r0 = tab.random_collapses();
        std::vector<uint32_t> pivots;
        for (const Product &p : group) {
            for (const auto &t : p.terms) rotate(t.first, t.second);
            for (size_t k = 1; k < p.terms.size(); k++) g2(Gate::CX, p.terms[k].first, p.terms[0].first);
            pivots.push_back(p.terms[0].first);
        }
        if (collapse_basis == 2) {
            collapse(pivots, 2);
        } else {
            // Rotated pivots are collapsed in ascending order, like X- and Y-basis measurements.
            std::vector<uint32_t> order(pivots);
            std::sort(order.begin(), order.end());
            for (uint32_t q : order) (void)tab.measure_z(q);
        }
        for (size_t i = 0; i < group.size(); i++) rec.push_back(tab.measure_z(pivots[i]) ^ group[i].negate);
        random_recorded += tab.random_collapses() - r0;
        for (const Product &p : group) {
            for (size_t k = 1; k < p.terms.size(); k++) g2(Gate::CX, p.terms[k].first, p.terms[0].first);
            for (const auto &t : p.terms) rotate(t.first, t.second);
        }
    }

    void phase_product(const Product &p, bool dag) {
        if (p.terms.empty()) return;
        const bool neg = p.negate ^ dag;
        for (const auto &t : p.terms) rotate(t.first, t.second);
        for (size_t k = 1; k < p.terms.size(); k++) g2(Gate::CX, p.terms[k].first, p.terms[0].first);
        g1(neg ? Gate::S_DAG : Gate::S, p.terms[0].first);
        for (size_t k = 1; k < p.terms.size(); k++) g2(Gate::CX, p.terms[k].first, p.terms[0].first);
        for (const auto &t : p.terms) rotate(t.first, t.second);
    }

    // MPP: products join a group until one overlaps the group's qubits; an identity product flushes
    // the group and records its sign (src/stim/circuit/gate_decomposition.cc:88-161).
    void do_mpp(const Instr &in) {
        std::vector<Product> group;
        std::vector<uint32_t> used;
        auto flush = [&]() {
            if (!group.empty()) measure_group(group);
            group.clear();
            used.clear();
        };
        for (Product &p : products(in)) {
            if (p.terms.empty()) {
                flush();
                rec.push_back(p.negate);
                continue;
            }
            bool overlap = false;
            for (const auto &t : p.terms) overlap |= std::find(used.begin(), used.end(), t.first) != used.end();
            if (overlap) flush();
            for (const auto &t : p.terms) used.push_back(t.first);
            group.push_back(std::move(p));
        }
        flush();
    }

    // MXX / MYY / MZZ: pairs join a segment until one reuses a qubit of the segment
    // (gate_decomposition.cc:245-274); each segment is measured as a group of two-factor products
    // folded onto each pair's first qubit, collapsed in the pair basis' order
    // (tableau_simulator.inl:232-288).
    void do_pairs(const Instr &in, int c) {
        const auto &ts = in.targets;
        std::vector<Product> group;
        std::vector<uint32_t> used;
        auto flush = [&]() {
            if (!group.empty()) measure_group(group, c);
            group.clear();
            used.clear();
        };
        for (size_t i = 0; i + 1 < ts.size(); i += 2) {
            const uint32_t a = ts[i].value, b = ts[i + 1].value;
            if (std::find(used.begin(), used.end(), a) != used.end() || std::find(used.begin(), used.end(), b) != used.end())
                flush();
            used.push_back(a);
            used.push_back(b);
            if (a == b) throw std::invalid_argument(std::string(gate_info(in.gate).name) + " targets a qubit twice in one pair");
            Product p;
            p.terms = {{a, c}, {b, c}};
            p.negate = ts[i].inverted ^ ts[i + 1].inverted;
            group.push_back(std::move(p));
        }
        flush();
    }

    // Splits an MPP/SPP target list into products.
    std::vector<Product> products(const Instr &in) {
        std::vector<Product> out;
        const auto &ts = in.targets;
        size_t i = 0;
        while (i < ts.size()) {
            Product p;
            std::vector<std::pair<uint32_t, int>> acc;
            int e = 0;
            bool more = true;
            while (more) {
                const Target &t = ts[i];
                if (t.kind == TK::PAULI) {
                    const int c = t.pauli & 3;
                    bool found = false;
                    for (auto &a : acc) {
                        if (a.first == t.value) {
                            e += pauli_phase(a.second, c);
                            a.second ^= c;
                            found = true;
                            break;
                        }
                    }
                    if (!found) acc.push_back({t.value, c});
                    p.negate ^= t.inverted;
                } else {
                    throw std::invalid_argument("Found an unsupported target in a Pauli product.");
                }
                i++;
                more = i + 1 < ts.size() && ts[i].kind == TK::COMBINER;
                if (more) i++;
            }
            e &= 3;
            if (e & 1) throw std::invalid_argument("Acted on an anti-Hermitian operator (e.g. X0*Z0 instead of Y0).");
            p.negate ^= e == 2;
            for (auto &a : acc) {
                if (a.second) p.terms.push_back(a);
            }
            std::sort(p.terms.begin(), p.terms.end());   // the pivot is the lowest qubit
            out.push_back(std::move(p));
        }
        return out;
    }

    // Classically controlled Pauli of a CX/CY/CZ/XCZ/YCZ pair whose control is a measurement or sweep bit.
    void feedback(Gate g, const Target &a, const Target &b) {
        Target c = a, t = b;
        if (g == Gate::XCZ || g == Gate::YCZ) std::swap(c, t);
        const Gate eff = g == Gate::XCZ ? Gate::CX : g == Gate::YCZ ? Gate::CY : g;
        const bool cbit = c.kind != TK::QUBIT, tbit = t.kind != TK::QUBIT;
        if (cbit && tbit) return;
        if (eff == Gate::CZ) {
            const Target &bit = cbit ? c : t;
            const Target &q = cbit ? t : c;
            if (bit_of(bit)) g1(Gate::Z, q.value);
            return;
        }
        if (tbit) throw std::invalid_argument("Controlled gate had a bit as its target, instead of its control.");
        if (bit_of(c)) g1(eff == Gate::CX ? Gate::X : Gate::Y, t.value);
    }

    void step(const Instr &in) {
        const auto &ts = in.targets;
        const Gate g = in.gate;
        switch (g) {
            case Gate::DETECTOR: case Gate::OBSERVABLE_INCLUDE: case Gate::TICK: case Gate::QUBIT_COORDS:
            case Gate::SHIFT_COORDS: case Gate::I: case Gate::II:
            case Gate::DEPOLARIZE1: case Gate::DEPOLARIZE2: case Gate::X_ERROR: case Gate::Y_ERROR:
            case Gate::Z_ERROR: case Gate::I_ERROR: case Gate::II_ERROR: case Gate::PAULI_CHANNEL_1:
            case Gate::PAULI_CHANNEL_2: case Gate::E: case Gate::ELSE_CORRELATED_ERROR:
                return;
            case Gate::MPAD:
                for (const Target &t : ts) rec.push_back(t.value != 0);
                return;
            case Gate::HERALDED_ERASE: case Gate::HERALDED_PAULI_CHANNEL_1:
                for (size_t i = 0; i < ts.size(); i++) rec.push_back(0);
                return;
            case Gate::M: case Gate::MX: case Gate::MY: case Gate::MR: case Gate::MRX: case Gate::MRY: {
                const int basis = (g == Gate::MX || g == Gate::MRX) ? 1 : (g == Gate::MY || g == Gate::MRY) ? 3 : 2;
                measure_targets(ts, basis, true, g == Gate::MR || g == Gate::MRX || g == Gate::MRY);
                return;
            }
            case Gate::R: case Gate::RX: case Gate::RY:
                measure_targets(ts, g == Gate::RX ? 1 : g == Gate::RY ? 3 : 2, false, true);
                return;
            case Gate::MXX: case Gate::MYY: case Gate::MZZ:
                do_pairs(in, g == Gate::MXX ? 1 : g == Gate::MYY ? 3 : 2);
                return;
            case Gate::MPP:
                do_mpp(in);
                return;
            case Gate::SPP: case Gate::SPP_DAG:
                for (const Product &p : products(in)) phase_product(p, g == Gate::SPP_DAG);
                return;
            default:
                break;
        }
        if (T.has1[(size_t)g]) {
            for (const Target &t : ts) g1(g, t.value);
            return;
        }
        if (T.has2[(size_t)g]) {
            for (size_t i = 0; i + 1 < ts.size(); i += 2) {
                const Target &a = ts[i], &b = ts[i + 1];
                if (a.kind == TK::QUBIT && b.kind == TK::QUBIT) {
                    g2(g, a.value, b.value);
                } else if (g == Gate::CX || g == Gate::CY || g == Gate::CZ || g == Gate::XCZ || g == Gate::YCZ) {
                    feedback(g, a, b);
                } else {
                    throw std::invalid_argument(std::string("Gate ") + gate_info(g).name + " can't target measurement bits.");
                }
            }
            return;
        }
        throw std::invalid_argument(std::string("Not supported by the reference sample: ") + gate_info(g).name);
    }
};

}  // namespace

std::vector<uint8_t> reference_sample(const std::vector<const Instr *> &flat, RefStats *stats) {
    const Stats st = compute_stats(flat);
    if (stats) *stats = RefStats{};
    if (st.num_measurements == 0) return {};
    // The tableau holds 2n rows of n X bits and n Z bits (about n^2 / 2 bytes), as the reference's
    // TableauSimulator does; circuits whose tableau would exceed the limit are refused up front rather
    // than failing inside the allocator.
    constexpr double MAX_TABLEAU_BYTES = 16.0 * (1ull << 30);
    const uint64_t n = std::max<uint64_t>(st.num_qubits, 1);
    const double bytes = 2.0 * (2.0 * (double)n) * (double)((n + 63) / 64) * 8.0;
    if (n >= (1ull << 31) || bytes > MAX_TABLEAU_BYTES) {
        throw std::runtime_error("the reference sample of a circuit on " + std::to_string(n) +
                                 " qubits needs a stabilizer tableau of " + std::to_string((uint64_t)(bytes / (1 << 20))) +
                                 " MiB (at most 16 GiB); skip the reference sample or pass one");
    }
    Runner run((uint32_t)n);
    run.rec.reserve(st.num_measurements);
    for (const Instr *in : flat) run.step(*in);
    if (stats) {
        stats->random_recorded = run.random_recorded;
        stats->random_total = run.tab.random_collapses();
    }
    return run.rec;
}

}  // namespace sa
