// Detector error model parser (see dem.h).
#include "dem.h"

#include <algorithm>
#include <cctype>
#include <cerrno>
#include <cstdlib>
#include <stdexcept>

namespace sa {

namespace {

struct Line {
    std::string name;              // lower case
    std::vector<double> args;
    struct Target {
        char kind;                 // 'D', 'L', '#' (number), '^'
        uint64_t value;
    };
    std::vector<Target> targets;
    bool opens_block = false;
    bool closes_block = false;
    size_t line_no = 0;
};

[[noreturn]] void fail(size_t line_no, const std::string &msg) {
    throw std::invalid_argument("detector error model line " + std::to_string(line_no) + ": " + msg);
}

uint64_t parse_uint(const std::string &s, size_t line_no) {
    if (s.empty() || !std::all_of(s.begin(), s.end(), [](char c) { return c >= '0' && c <= '9'; })) {
        fail(line_no, "bad integer '" + s + "'");
    }
    errno = 0;
    const unsigned long long v = std::strtoull(s.c_str(), nullptr, 10);
    if (errno) fail(line_no, "integer out of range '" + s + "'");
    return v;
}

Line parse_line(const std::string &raw, size_t line_no) {
    Line ln;
    ln.line_no = line_no;
    std::string t = raw.substr(0, raw.find('#'));
    auto trim = [](std::string &x) {
        size_t a = 0, b = x.size();
        while (a < b && std::isspace((unsigned char)x[a])) a++;
        while (b > a && std::isspace((unsigned char)x[b - 1])) b--;
        x = x.substr(a, b - a);
    };
    trim(t);
    if (t.empty()) return ln;
    if (t == "}") {
        ln.closes_block = true;
        return ln;
    }
    if (t.back() == '{') {
        ln.opens_block = true;
        t.pop_back();
        trim(t);
    }
    size_t i = 0;
    if (i >= t.size() || !std::isalpha((unsigned char)t[i])) fail(line_no, "expected an instruction name");
    while (i < t.size() && (std::isalnum((unsigned char)t[i]) || t[i] == '_')) {
        ln.name.push_back((char)std::tolower((unsigned char)t[i]));
        i++;
    }
    if (i < t.size() && t[i] == '[') {   // tag: skipped
        const size_t e = t.find(']', i);
        if (e == std::string::npos) fail(line_no, "unterminated tag");
        i = e + 1;
    }
    if (i < t.size() && t[i] == '(') {
        const size_t e = t.find(')', i);
        if (e == std::string::npos) fail(line_no, "unterminated argument list");
        std::string inner = t.substr(i + 1, e - i - 1);
        size_t a = 0;
        while (a <= inner.size()) {
            size_t c = inner.find(',', a);
            if (c == std::string::npos) c = inner.size();
            std::string arg = inner.substr(a, c - a);
            trim(arg);
            if (arg.empty()) fail(line_no, "empty argument");
            char *end = nullptr;
            const double v = std::strtod(arg.c_str(), &end);
            if (!end || *end) fail(line_no, "bad argument '" + arg + "'");
            ln.args.push_back(v);
            a = c + 1;
        }
        i = e + 1;
    }
    while (i < t.size()) {
        if (!std::isspace((unsigned char)t[i])) fail(line_no, "expected whitespace before a target");
        while (i < t.size() && std::isspace((unsigned char)t[i])) i++;
        if (i >= t.size()) break;
        size_t e = i;
        while (e < t.size() && !std::isspace((unsigned char)t[e])) e++;
        const std::string tok = t.substr(i, e - i);
        if (tok == "^") {
            ln.targets.push_back({'^', 0});
        } else if (tok[0] == 'D' || tok[0] == 'd') {
            ln.targets.push_back({'D', parse_uint(tok.substr(1), line_no)});
        } else if (tok[0] == 'L' || tok[0] == 'l') {
            ln.targets.push_back({'L', parse_uint(tok.substr(1), line_no)});
        } else {
            ln.targets.push_back({'#', parse_uint(tok, line_no)});
        }
        i = e;
    }
    return ln;
}

struct Block {
    std::vector<Line> lines;             // instructions of this block (REPEAT lines carry a child index)
    std::vector<size_t> child;           // per line: index of the repeated block, or SIZE_MAX
};

void cancel_pairs(std::vector<uint32_t> &v) {
    std::sort(v.begin(), v.end());
    std::vector<uint32_t> out;
    for (size_t k = 0; k < v.size();) {
        size_t e = k;
        while (e < v.size() && v[e] == v[k]) e++;
        if ((e - k) & 1) out.push_back(v[k]);
        k = e;
    }
    v.swap(out);
}

uint64_t sat_add(uint64_t a, uint64_t b) {
    return a + b < a ? UINT64_MAX : a + b;
}

uint64_t sat_mul(uint64_t a, uint64_t b) {
    return (a && b > UINT64_MAX / a) ? UINT64_MAX : a * b;
}

// Unrolled size of a block: error mechanisms and the detector shift of one pass (both saturating).
struct BlockSize {
    uint64_t errors = 0, shift = 0;
};

BlockSize block_size(const std::vector<Block> &blocks, size_t b) {
    BlockSize s;
    const Block &blk = blocks[b];
    for (size_t k = 0; k < blk.lines.size(); k++) {
        const Line &ln = blk.lines[k];
        if (blk.child[k] != SIZE_MAX) {
            const BlockSize c = block_size(blocks, blk.child[k]);
            s.errors = sat_add(s.errors, sat_mul(c.errors, ln.targets[0].value));
            s.shift = sat_add(s.shift, sat_mul(c.shift, ln.targets[0].value));
        } else if (ln.name == "error") {
            s.errors = sat_add(s.errors, 1);
        } else if (ln.name == "shift_detectors") {
            for (const auto &tg : ln.targets) s.shift = sat_add(s.shift, tg.value);
        }
    }
    return s;
}

// Error mechanisms the sampler unrolls at most (the model is materialized for the GPU).
constexpr uint64_t MAX_DEM_MECHANISMS = 1ull << 28;

struct Flattener {
    const std::vector<Block> &blocks;
    DemModel &model;
    uint64_t det_offset = 0;

    void run(size_t b) {
        const Block &blk = blocks[b];
        for (size_t k = 0; k < blk.lines.size(); k++) {
            const Line &ln = blk.lines[k];
            if (blk.child[k] != SIZE_MAX) {
                const uint64_t reps = ln.targets[0].value;
                if (reps == 0) continue;
                const BlockSize c = block_size(blocks, blk.child[k]);
                if (c.errors == 0) {
                    // No mechanisms: only detector declarations and shifts, which never shrink, so
                    // the last pass (at its final offset) declares the largest detectors.
                    const uint64_t before = sat_mul(c.shift, reps - 1);
                    if (sat_add(det_offset, before) >= (1ull << 62)) fail(ln.line_no, "detector shift too large");
                    det_offset += before;
                    run(blk.child[k]);
                    continue;
                }
                if (sat_add(model.errors.size(), sat_mul(c.errors, reps)) > MAX_DEM_MECHANISMS) {
                    fail(ln.line_no, "detector error model too large to unroll (more than 2^28 error mechanisms)");
                }
                for (uint64_t r = 0; r < reps; r++) run(blk.child[k]);
                continue;
            }
            if (ln.name == "error") {
                DemError e;
                e.p = ln.args[0];
                for (const auto &tg : ln.targets) {
                    if (tg.kind == 'D') {
                        const uint64_t d = det_offset + tg.value;
                        if (d >= (1ull << 31)) fail(ln.line_no, "detector index too large");
                        e.dets.push_back((uint32_t)d);
                        model.num_detectors = std::max(model.num_detectors, d + 1);
                    } else if (tg.kind == 'L') {
                        if (tg.value >= (1ull << 31)) fail(ln.line_no, "observable index too large");
                        e.obs.push_back((uint32_t)tg.value);
                        model.num_observables = std::max(model.num_observables, tg.value + 1);
                    }
                }
                cancel_pairs(e.dets);
                cancel_pairs(e.obs);
                model.errors.push_back(std::move(e));
            } else if (ln.name == "detector") {
                for (const auto &tg : ln.targets)
                    model.num_detectors = std::max(model.num_detectors, det_offset + tg.value + 1);
            } else if (ln.name == "logical_observable") {
                for (const auto &tg : ln.targets)
                    model.num_observables = std::max(model.num_observables, tg.value + 1);
            } else if (ln.name == "shift_detectors") {
                for (const auto &tg : ln.targets) det_offset += tg.value;
            }
        }
    }
};

}  // namespace

DemModel parse_dem(const char *text, size_t len) {
    std::vector<Block> blocks(1);
    std::vector<size_t> stack{0};
    size_t line_no = 0, pos = 0;
    const std::string src(text ? text : "", text ? len : 0);
    while (pos <= src.size()) {
        size_t e = src.find('\n', pos);
        if (e == std::string::npos) e = src.size();
        line_no++;
        Line ln = parse_line(src.substr(pos, e - pos), line_no);
        pos = e + 1;
        if (ln.closes_block) {
            if (stack.size() < 2) fail(line_no, "'}' without a matching block start");
            stack.pop_back();
            continue;
        }
        if (ln.name.empty()) continue;
        const std::string &n = ln.name;
        if (n == "repeat") {
            if (!ln.opens_block) fail(line_no, "repeat needs a '{' block");
            if (ln.targets.size() != 1 || ln.targets[0].kind != '#' || !ln.args.empty())
                fail(line_no, "repeat takes one repetition count");
            blocks.push_back({});
            const size_t child = blocks.size() - 1;
            Block &parent = blocks[stack.back()];
            parent.lines.push_back(std::move(ln));
            parent.child.push_back(child);
            stack.push_back(child);
            continue;
        }
        if (ln.opens_block) fail(line_no, "only repeat opens a block");
        if (n == "error") {
            if (ln.args.size() != 1) fail(line_no, "error takes exactly one probability argument");
            if (!(ln.args[0] >= 0 && ln.args[0] <= 1)) fail(line_no, "error probability not in [0, 1]");
            for (const auto &tg : ln.targets)
                if (tg.kind == '#') fail(line_no, "error targets must be D#, L# or ^");
        } else if (n == "detector") {
            for (const auto &tg : ln.targets)
                if (tg.kind != 'D') fail(line_no, "detector targets must be D#");
        } else if (n == "logical_observable") {
            for (const auto &tg : ln.targets)
                if (tg.kind != 'L') fail(line_no, "logical_observable targets must be L#");
        } else if (n == "shift_detectors") {
            for (const auto &tg : ln.targets)
                if (tg.kind != '#') fail(line_no, "shift_detectors takes a number");
        } else {
            fail(line_no, "unknown instruction '" + n + "'");
        }
        blocks[stack.back()].lines.push_back(std::move(ln));
        blocks[stack.back()].child.push_back(SIZE_MAX);
    }
    if (stack.size() != 1) fail(line_no, "unterminated repeat block");
    DemModel model;
    Flattener f{blocks, model};
    f.run(0);
    return model;
}

}  // namespace sa
