// .stim text parser for the stim_amd sampler host. See circuit.h for the reference citations.
#include "circuit.h"

#include <algorithm>
#include <cctype>
#include <cstdlib>
#include <cstring>
#include <unordered_map>

namespace sa {

namespace {

// Canonical gate table: name, shape, argument count range, flags.
// Names, target shapes and argument counts follow src/stim/gates/gate_data_*.cc.
const GateInfo GATES[] = {
    {"DETECTOR", Gate::DETECTOR, Shape::RECORDS, -1, -1, false, false},
    {"OBSERVABLE_INCLUDE", Gate::OBSERVABLE_INCLUDE, Shape::ANNOTATION_ANY, 1, 1, false, false},
    {"TICK", Gate::TICK, Shape::NONE, 0, 0, false, false},
    {"QUBIT_COORDS", Gate::QUBIT_COORDS, Shape::QUBITS, -1, -1, false, false},
    {"SHIFT_COORDS", Gate::SHIFT_COORDS, Shape::NONE, -1, -1, false, false},
    {"MPAD", Gate::MPAD, Shape::QUBITS, 0, 1, false, true},
    {"MX", Gate::MX, Shape::QUBITS, 0, 1, false, true},
    {"MY", Gate::MY, Shape::QUBITS, 0, 1, false, true},
    {"M", Gate::M, Shape::QUBITS, 0, 1, false, true},
    {"MRX", Gate::MRX, Shape::QUBITS, 0, 1, false, true},
    {"MRY", Gate::MRY, Shape::QUBITS, 0, 1, false, true},
    {"MR", Gate::MR, Shape::QUBITS, 0, 1, false, true},
    {"RX", Gate::RX, Shape::QUBITS, 0, 0, false, false},
    {"RY", Gate::RY, Shape::QUBITS, 0, 0, false, false},
    {"R", Gate::R, Shape::QUBITS, 0, 0, false, false},
    {"XCX", Gate::XCX, Shape::PAIRS, 0, 0, false, false},
    {"XCY", Gate::XCY, Shape::PAIRS, 0, 0, false, false},
    {"XCZ", Gate::XCZ, Shape::PAIRS, 0, 0, true, false},
    {"YCX", Gate::YCX, Shape::PAIRS, 0, 0, false, false},
    {"YCY", Gate::YCY, Shape::PAIRS, 0, 0, false, false},
    {"YCZ", Gate::YCZ, Shape::PAIRS, 0, 0, true, false},
    {"CX", Gate::CX, Shape::PAIRS, 0, 0, true, false},
    {"CY", Gate::CY, Shape::PAIRS, 0, 0, true, false},
    {"CZ", Gate::CZ, Shape::PAIRS, 0, 0, true, false},
    {"H", Gate::H, Shape::QUBITS, 0, 0, false, false},
    {"H_XY", Gate::H_XY, Shape::QUBITS, 0, 0, false, false},
    {"H_YZ", Gate::H_YZ, Shape::QUBITS, 0, 0, false, false},
    {"H_NXY", Gate::H_NXY, Shape::QUBITS, 0, 0, false, false},
    {"H_NXZ", Gate::H_NXZ, Shape::QUBITS, 0, 0, false, false},
    {"H_NYZ", Gate::H_NYZ, Shape::QUBITS, 0, 0, false, false},
    {"DEPOLARIZE1", Gate::DEPOLARIZE1, Shape::QUBITS, 1, 1, false, false},
    {"DEPOLARIZE2", Gate::DEPOLARIZE2, Shape::PAIRS, 1, 1, false, false},
    {"X_ERROR", Gate::X_ERROR, Shape::QUBITS, 1, 1, false, false},
    {"Y_ERROR", Gate::Y_ERROR, Shape::QUBITS, 1, 1, false, false},
    {"Z_ERROR", Gate::Z_ERROR, Shape::QUBITS, 1, 1, false, false},
    {"I_ERROR", Gate::I_ERROR, Shape::QUBITS, -1, -1, false, false},
    {"II_ERROR", Gate::II_ERROR, Shape::PAIRS, -1, -1, false, false},
    {"PAULI_CHANNEL_1", Gate::PAULI_CHANNEL_1, Shape::QUBITS, 3, 3, false, false},
    {"PAULI_CHANNEL_2", Gate::PAULI_CHANNEL_2, Shape::PAIRS, 15, 15, false, false},
    {"E", Gate::E, Shape::PAULI_PRODUCTS, 1, 1, false, false},
    {"ELSE_CORRELATED_ERROR", Gate::ELSE_CORRELATED_ERROR, Shape::PAULI_PRODUCTS, 1, 1, false, false},
    {"HERALDED_ERASE", Gate::HERALDED_ERASE, Shape::QUBITS, 1, 1, false, true},
    {"HERALDED_PAULI_CHANNEL_1", Gate::HERALDED_PAULI_CHANNEL_1, Shape::QUBITS, 4, 4, false, true},
    {"I", Gate::I, Shape::QUBITS, 0, 0, false, false},
    {"X", Gate::X, Shape::QUBITS, 0, 0, false, false},
    {"Y", Gate::Y, Shape::QUBITS, 0, 0, false, false},
    {"Z", Gate::Z, Shape::QUBITS, 0, 0, false, false},
    {"C_XYZ", Gate::C_XYZ, Shape::QUBITS, 0, 0, false, false},
    {"C_NXYZ", Gate::C_NXYZ, Shape::QUBITS, 0, 0, false, false},
    {"C_XNYZ", Gate::C_XNYZ, Shape::QUBITS, 0, 0, false, false},
    {"C_XYNZ", Gate::C_XYNZ, Shape::QUBITS, 0, 0, false, false},
    {"C_ZYX", Gate::C_ZYX, Shape::QUBITS, 0, 0, false, false},
    {"C_ZYNX", Gate::C_ZYNX, Shape::QUBITS, 0, 0, false, false},
    {"C_ZNYX", Gate::C_ZNYX, Shape::QUBITS, 0, 0, false, false},
    {"C_NZYX", Gate::C_NZYX, Shape::QUBITS, 0, 0, false, false},
    {"SQRT_X", Gate::SQRT_X, Shape::QUBITS, 0, 0, false, false},
    {"SQRT_X_DAG", Gate::SQRT_X_DAG, Shape::QUBITS, 0, 0, false, false},
    {"SQRT_Y", Gate::SQRT_Y, Shape::QUBITS, 0, 0, false, false},
    {"SQRT_Y_DAG", Gate::SQRT_Y_DAG, Shape::QUBITS, 0, 0, false, false},
    {"S", Gate::S, Shape::QUBITS, 0, 0, false, false},
    {"S_DAG", Gate::S_DAG, Shape::QUBITS, 0, 0, false, false},
    {"II", Gate::II, Shape::PAIRS, 0, 0, false, false},
    {"SQRT_XX", Gate::SQRT_XX, Shape::PAIRS, 0, 0, false, false},
    {"SQRT_XX_DAG", Gate::SQRT_XX_DAG, Shape::PAIRS, 0, 0, false, false},
    {"SQRT_YY", Gate::SQRT_YY, Shape::PAIRS, 0, 0, false, false},
    {"SQRT_YY_DAG", Gate::SQRT_YY_DAG, Shape::PAIRS, 0, 0, false, false},
    {"SQRT_ZZ", Gate::SQRT_ZZ, Shape::PAIRS, 0, 0, false, false},
    {"SQRT_ZZ_DAG", Gate::SQRT_ZZ_DAG, Shape::PAIRS, 0, 0, false, false},
    {"SWAP", Gate::SWAP, Shape::PAIRS, 0, 0, false, false},
    {"ISWAP", Gate::ISWAP, Shape::PAIRS, 0, 0, false, false},
    {"ISWAP_DAG", Gate::ISWAP_DAG, Shape::PAIRS, 0, 0, false, false},
    {"CXSWAP", Gate::CXSWAP, Shape::PAIRS, 0, 0, false, false},
    {"SWAPCX", Gate::SWAPCX, Shape::PAIRS, 0, 0, false, false},
    {"CZSWAP", Gate::CZSWAP, Shape::PAIRS, 0, 0, false, false},
    {"MXX", Gate::MXX, Shape::PAIRS, 0, 1, false, true},
    {"MYY", Gate::MYY, Shape::PAIRS, 0, 1, false, true},
    {"MZZ", Gate::MZZ, Shape::PAIRS, 0, 1, false, true},
    {"MPP", Gate::MPP, Shape::PAULI_PRODUCTS, 0, 1, false, true},
    {"SPP", Gate::SPP, Shape::PAULI_PRODUCTS, 0, 0, false, false},
    {"SPP_DAG", Gate::SPP_DAG, Shape::PAULI_PRODUCTS, 0, 0, false, false},
};

const std::unordered_map<std::string, Gate> &name_map() {
    static const std::unordered_map<std::string, Gate> m = [] {
        std::unordered_map<std::string, Gate> r;
        for (const auto &g : GATES) {
            r[g.name] = g.gate;
        }
        // Aliases (src/stim/gates/gate_data_*.cc add_gate_alias calls).
        r["MZ"] = Gate::M;
        r["MRZ"] = Gate::MR;
        r["RZ"] = Gate::R;
        r["ZCX"] = Gate::CX;
        r["CNOT"] = Gate::CX;
        r["ZCY"] = Gate::CY;
        r["ZCZ"] = Gate::CZ;
        r["H_XZ"] = Gate::H;
        r["CORRELATED_ERROR"] = Gate::E;
        r["SQRT_Z"] = Gate::S;
        r["SQRT_Z_DAG"] = Gate::S_DAG;
        r["SWAPCZ"] = Gate::CZSWAP;
        return r;
    }();
    return m;
}

struct Cursor {
    const char *p;
    const char *end;
    size_t line = 1;

    [[noreturn]] void fail(const std::string &msg) const {
        throw ParseError("Parse error on line " + std::to_string(line) + ": " + msg);
    }
};

bool is_space(char c) {
    return c == ' ' || c == '\t' || c == '\r';
}

uint64_t parse_uint(const std::string &s, const Cursor &cur, const char *what) {
    if (s.empty() || s.size() > 19) {
        cur.fail(std::string("bad ") + what + " '" + s + "'");
    }
    uint64_t v = 0;
    for (char c : s) {
        if (c < '0' || c > '9') {
            cur.fail(std::string("bad ") + what + " '" + s + "'");
        }
        v = v * 10 + (uint64_t)(c - '0');
    }
    return v;
}

Target parse_target(const std::string &tok, const Cursor &cur) {
    Target t{0, TK::QUBIT, false, 0};
    if (tok == "*") {
        t.kind = TK::COMBINER;
        return t;
    }
    if (tok.rfind("rec[-", 0) == 0 && tok.back() == ']') {
        t.kind = TK::REC;
        uint64_t k = parse_uint(tok.substr(5, tok.size() - 6), cur, "record target");
        if (k == 0 || k >= (1u << 24)) {
            cur.fail("record lookback out of range in '" + tok + "'");
        }
        t.value = (uint32_t)k;
        return t;
    }
    if (tok.rfind("sweep[", 0) == 0 && tok.back() == ']') {
        t.kind = TK::SWEEP;
        t.value = (uint32_t)parse_uint(tok.substr(6, tok.size() - 7), cur, "sweep target");
        return t;
    }
    size_t i = 0;
    if (tok[i] == '!') {
        t.inverted = true;
        i++;
    }
    if (i < tok.size()) {
        char c = (char)std::toupper((unsigned char)tok[i]);
        if (c == 'X' || c == 'Y' || c == 'Z') {
            t.kind = TK::PAULI;
            t.pauli = c == 'X' ? 1 : c == 'Z' ? 2 : 3;
            i++;
        }
    }
    uint64_t q = parse_uint(tok.substr(i), cur, "target");
    if (q >= (1u << 24)) {
        cur.fail("qubit index too large in '" + tok + "'");
    }
    t.value = (uint32_t)q;
    return t;
}

void validate(const Instr &in, const Cursor &cur) {
    const GateInfo &gi = gate_info(in.gate);
    int na = (int)in.args.size();
    if ((gi.min_args >= 0 && na < gi.min_args) || (gi.max_args >= 0 && na > gi.max_args)) {
        cur.fail(std::string("gate ") + gi.name + " got " + std::to_string(na) + " parens arguments");
    }
    for (double a : in.args) {
        bool prob = in.gate != Gate::DETECTOR && in.gate != Gate::QUBIT_COORDS && in.gate != Gate::SHIFT_COORDS &&
                    in.gate != Gate::OBSERVABLE_INCLUDE;
        if (prob && !(a >= 0 && a <= 1)) {
            cur.fail(std::string("probability argument out of range for ") + gi.name);
        }
    }
    if (in.gate == Gate::OBSERVABLE_INCLUDE) {
        double a = in.args[0];
        if (!(a >= 0) || a != (double)(uint64_t)a || a > 1e8) {
            cur.fail("OBSERVABLE_INCLUDE index must be a non-negative integer");
        }
    }
    double sum = 0;
    for (double a : in.args) sum += a;
    if ((in.gate == Gate::PAULI_CHANNEL_1 || in.gate == Gate::PAULI_CHANNEL_2 ||
         in.gate == Gate::HERALDED_PAULI_CHANNEL_1) && sum > 1 + 1e-12) {
        cur.fail(std::string("disjoint probabilities sum above 1 for ") + gi.name);
    }
    for (const Target &t : in.targets) {
        switch (gi.shape) {
            case Shape::NONE:
                cur.fail(std::string(gi.name) + " takes no targets");
            case Shape::QUBITS:
                if (t.kind != TK::QUBIT) cur.fail(std::string(gi.name) + " only takes qubit targets");
                if (in.gate == Gate::MPAD && t.value > 1) cur.fail("MPAD targets must be 0 or 1");
                break;
            case Shape::PAIRS:
                if (t.kind == TK::PAULI || t.kind == TK::COMBINER) cur.fail(std::string(gi.name) + " takes qubit targets");
                if ((t.kind == TK::REC || t.kind == TK::SWEEP) && !gi.can_target_bits)
                    cur.fail(std::string(gi.name) + " cannot target measurement records or sweep bits");
                break;
            case Shape::PAULI_PRODUCTS:
                if (t.kind == TK::QUBIT) cur.fail(std::string(gi.name) + " takes Pauli targets like X0 or Z1*Z2");
                if ((in.gate == Gate::E || in.gate == Gate::ELSE_CORRELATED_ERROR) && t.kind != TK::PAULI)
                    cur.fail(std::string(gi.name) + " only takes Pauli targets");
                if (in.gate == Gate::MPP && (t.kind == TK::REC || t.kind == TK::SWEEP))
                    cur.fail("MPP cannot target measurement records");
                if ((in.gate == Gate::SPP || in.gate == Gate::SPP_DAG) && (t.kind == TK::REC || t.kind == TK::SWEEP))
                    cur.fail(std::string(gi.name) + " only takes Pauli targets ('X2', 'Y3', 'Z5', etc).");
                break;
            case Shape::RECORDS:
                if (t.kind != TK::REC) cur.fail("DETECTOR only takes measurement record targets like rec[-1]");
                break;
            case Shape::ANNOTATION_ANY:
                if (t.kind != TK::REC && t.kind != TK::PAULI)
                    cur.fail("OBSERVABLE_INCLUDE takes rec[-k] or Pauli targets");
                break;
        }
    }
    if (gi.shape == Shape::PAIRS) {
        if (in.targets.size() % 2) cur.fail(std::string(gi.name) + " needs an even number of targets");
        for (size_t k = 0; k < in.targets.size(); k += 2) {
            const Target &a = in.targets[k], &b = in.targets[k + 1];
            if (a.kind == TK::QUBIT && b.kind == TK::QUBIT && a.value == b.value) {
                cur.fail(std::string(gi.name) + " can't target the same qubit twice in a pair");
            }
            if (b.kind != TK::QUBIT && (in.gate == Gate::CX || in.gate == Gate::CY)) {
                cur.fail(std::string("controlled gate ") + gi.name + " had a bit as its target instead of its control");
            }
            if (a.kind != TK::QUBIT && (in.gate == Gate::XCZ || in.gate == Gate::YCZ)) {
                cur.fail(std::string("controlled gate ") + gi.name + " had a bit as its target instead of its control");
            }
        }
    }
    if (gi.shape == Shape::PAULI_PRODUCTS) {
        // Combiners must join terms.
        bool expect_term = true;
        for (const Target &t : in.targets) {
            if (t.kind == TK::COMBINER) {
                if (expect_term) cur.fail("misplaced combiner '*'");
                expect_term = true;
            } else {
                expect_term = false;
            }
        }
        if (!in.targets.empty() && expect_term) cur.fail("dangling combiner '*'");
        if ((in.gate == Gate::E || in.gate == Gate::ELSE_CORRELATED_ERROR)) {
            for (const Target &t : in.targets) {
                if (t.kind == TK::COMBINER) cur.fail(std::string(gi.name) + " does not take combiners");
            }
        }
    }
}

}  // namespace

const GateInfo &gate_info(Gate g) {
    static const std::vector<const GateInfo *> by_id = [] {
        std::vector<const GateInfo *> v((size_t)Gate::NUM_GATES, nullptr);
        for (const auto &gi : GATES) v[(size_t)gi.gate] = &gi;
        return v;
    }();
    return *by_id[(size_t)g];
}

Block parse_circuit(const char *text, size_t len) {
    Cursor cur{text, text + len};
    Block root;
    std::vector<Block *> stack{&root};
    std::vector<std::shared_ptr<Block>> owners;
    const auto &names = name_map();
    const char *p = text;
    const char *end = text + len;
    while (p < end) {
        const char *eol = (const char *)memchr(p, '\n', (size_t)(end - p));
        if (!eol) eol = end;
        std::string line(p, (size_t)(eol - p));
        p = eol < end ? eol + 1 : end;
        // Strip comments (but not inside tags).
        size_t cpos = std::string::npos;
        bool in_tag = false;
        for (size_t i = 0; i < line.size(); i++) {
            if (line[i] == '[' && !in_tag && i > 0 && (std::isalnum((unsigned char)line[i - 1]) || line[i - 1] == '_')) {
                // A tag only directly follows the gate name; rec[...] / sweep[...] are handled by tokenizing.
                in_tag = true;
            } else if (line[i] == ']' && in_tag) {
                in_tag = false;
            } else if (line[i] == '#' && !in_tag) {
                cpos = i;
                break;
            }
        }
        if (cpos != std::string::npos) line.resize(cpos);
        size_t a = 0, b = line.size();
        while (a < b && is_space(line[a])) a++;
        while (b > a && is_space(line[b - 1])) b--;
        line = line.substr(a, b - a);
        if (line.empty()) {
            cur.line++;
            continue;
        }
        if (line == "}") {
            if (stack.size() == 1) cur.fail("unmatched '}'");
            stack.pop_back();
            cur.line++;
            continue;
        }
        bool opens = false;
        if (line.back() == '{') {
            opens = true;
            line.pop_back();
            while (!line.empty() && is_space(line.back())) line.pop_back();
        }
        size_t i = 0;
        if (!std::isalpha((unsigned char)line[0])) cur.fail("expected an instruction name");
        while (i < line.size() && (std::isalnum((unsigned char)line[i]) || line[i] == '_')) i++;
        std::string name = line.substr(0, i);
        for (char &c : name) c = (char)std::toupper((unsigned char)c);
        // Optional tag.
        if (i < line.size() && line[i] == '[') {
            size_t close = line.find(']', i);
            if (close == std::string::npos) cur.fail("unterminated tag");
            i = close + 1;
        }
        std::vector<double> args;
        size_t j = i;
        while (j < line.size() && is_space(line[j])) j++;
        if (j < line.size() && line[j] == '(') {
            size_t close = line.find(')', j);
            if (close == std::string::npos) cur.fail("unterminated parens arguments");
            std::string inner = line.substr(j + 1, close - j - 1);
            size_t s = 0;
            bool any = false;
            for (char c : inner) any |= !is_space(c);
            if (any) {
                while (true) {
                    size_t comma = inner.find(',', s);
                    std::string tok = inner.substr(s, comma == std::string::npos ? std::string::npos : comma - s);
                    size_t ta = 0, tb = tok.size();
                    while (ta < tb && is_space(tok[ta])) ta++;
                    while (tb > ta && is_space(tok[tb - 1])) tb--;
                    tok = tok.substr(ta, tb - ta);
                    char *endp = nullptr;
                    double v = std::strtod(tok.c_str(), &endp);
                    if (tok.empty() || *endp != '\0') cur.fail("bad parens argument '" + tok + "'");
                    args.push_back(v);
                    if (comma == std::string::npos) break;
                    s = comma + 1;
                }
            }
            i = close + 1;
        }
        // Tokenize targets; '*' is its own token even without surrounding spaces.
        std::vector<std::string> toks;
        std::string curtok;
        for (size_t k = i; k < line.size(); k++) {
            char c = line[k];
            if (is_space(c)) {
                if (!curtok.empty()) toks.push_back(curtok), curtok.clear();
            } else if (c == '*') {
                if (!curtok.empty()) toks.push_back(curtok), curtok.clear();
                toks.push_back("*");
            } else {
                curtok += c;
            }
        }
        if (!curtok.empty()) toks.push_back(curtok);
        if (i < line.size() && !is_space(line[i]) && line[i] != '*' && !toks.empty() && i == 0) {
            cur.fail("malformed instruction");
        }
        if (name == "REPEAT") {
            if (!opens || toks.size() != 1 || !args.empty()) cur.fail("malformed REPEAT block");
            uint64_t reps = parse_uint(toks[0], cur, "repeat count");
            if (reps == 0) cur.fail("REPEAT 0 blocks are not allowed");
            Item it;
            it.is_repeat = true;
            it.reps = reps;
            it.body = std::make_shared<Block>();
            Block *body = it.body.get();
            stack.back()->items.push_back(std::move(it));
            stack.push_back(body);
            cur.line++;
            continue;
        }
        if (opens) cur.fail("only REPEAT can start a block");
        auto f = names.find(name);
        if (f == names.end()) cur.fail("unknown gate '" + name + "'");
        Instr in;
        in.gate = f->second;
        in.args = std::move(args);
        for (const auto &tk : toks) in.targets.push_back(parse_target(tk, cur));
        validate(in, cur);
        Item it;
        it.instr = std::move(in);
        stack.back()->items.push_back(std::move(it));
        cur.line++;
    }
    if (stack.size() != 1) cur.fail("unterminated REPEAT block");
    return root;
}

namespace {
constexpr size_t MAX_UNROLLED_INSTRUCTIONS = (size_t)1 << 27;

void flatten_into(const Block &b, std::vector<const Instr *> &out) {
    for (const Item &it : b.items) {
        if (it.is_repeat) {
            // The body is flattened once and copied: an empty body costs nothing whatever the count,
            // and the size is checked before copying.
            std::vector<const Instr *> body;
            flatten_into(*it.body, body);
            if (body.empty()) continue;
            if (out.size() > MAX_UNROLLED_INSTRUCTIONS ||
                it.reps > (MAX_UNROLLED_INSTRUCTIONS - out.size()) / body.size()) {
                throw std::invalid_argument("circuit too long to unroll (more than 2^27 instructions)");
            }
            for (uint64_t r = 0; r < it.reps; r++) out.insert(out.end(), body.begin(), body.end());
        } else {
            out.push_back(&it.instr);
        }
    }
}
}  // namespace

std::vector<const Instr *> flatten(const Block &b) {
    std::vector<const Instr *> out;
    flatten_into(b, out);
    return out;
}

std::vector<std::pair<size_t, size_t>> product_groups(const Instr &in) {
    std::vector<std::pair<size_t, size_t>> groups;
    size_t k = 0;
    const auto &T = in.targets;
    while (k < T.size()) {
        size_t start = k;
        k++;
        while (k < T.size() && T[k].kind == TK::COMBINER) k += 2;
        groups.push_back({start, std::min(k, T.size())});
    }
    return groups;
}

uint64_t num_results(const Instr &in) {
    switch (in.gate) {
        case Gate::M: case Gate::MX: case Gate::MY: case Gate::MR: case Gate::MRX: case Gate::MRY:
        case Gate::MPAD: case Gate::HERALDED_ERASE: case Gate::HERALDED_PAULI_CHANNEL_1:
            return in.targets.size();
        case Gate::MXX: case Gate::MYY: case Gate::MZZ:
            return in.targets.size() / 2;
        case Gate::MPP:
            return product_groups(in).size();
        default:
            return 0;
    }
}

Stats compute_stats(const std::vector<const Instr *> &flat) {
    Stats s;
    for (const Instr *in : flat) {
        for (const Target &t : in->targets) {
            if ((t.kind == TK::QUBIT && in->gate != Gate::MPAD) || t.kind == TK::PAULI) {
                s.num_qubits = std::max<uint64_t>(s.num_qubits, (uint64_t)t.value + 1);
            } else if (t.kind == TK::REC) {
                s.max_lookback = std::max<uint64_t>(s.max_lookback, t.value);
            }
        }
        s.num_measurements += num_results(*in);
        if (in->gate == Gate::DETECTOR) {
            s.num_detectors++;
        } else if (in->gate == Gate::OBSERVABLE_INCLUDE) {
            s.num_observables = std::max<uint64_t>(s.num_observables, (uint64_t)in->args[0] + 1);
        }
    }
    return s;
}

}  // namespace sa
