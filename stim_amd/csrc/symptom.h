// Sparse-symptom sampling engine.
//
// The frame rules are linear over GF(2) (src/stim/simulators/frame_simulator.inl:319-630 only XOR
// rows; measurements XOR frame rows into the record; detectors and observables XOR record and frame
// rows), so a shot's detection events are the XOR of the symptoms of the noise outcomes that fired
// in it. This engine:
//   1. runs the frame kernel in fixed-fault mode once per *basis fault* of every noise site (one
//      shot column per basis fault: X or Z on each qubit of a channel, a record flip, a herald, a
//      correlated-chain link) to get each basis fault's detector/observable symptom;
//   2. keeps the symptoms as sparse column lists in HBM;
//   3. samples shots by drawing the channels exactly as the reference does (Bernoulli(p) per site via
//      geometric gaps, RareErrorIterator src/stim/util_bot/probability_util.h:33-71; then the
//      outcome from the channel's disjoint distribution) and XORs the chosen outcome's basis
//      symptoms into an LDS image of the shot-major output rows, which is streamed to HBM.
// It is used for circuits whose detectors and observables are deterministic (frame randomization
// cannot change any output); the frame engine samples the others.
#pragma once

#include <cstdint>
#include <vector>

#include <hip/hip_runtime.h>

#include "compiler.h"

namespace sa {

struct SymClass {
    float p;         // firing probability of one site
    float inv_log;   // 1 / log2(1 - p)
    uint32_t nsites;
    uint32_t basis;  // basis columns per site
    uint32_t col0;   // first column of the class
    uint32_t woff;   // outcome table offset ((threshold, code-bitmask) pairs)
    uint32_t nout;   // outcomes | SYM_UNIFORM when all outcomes are equally likely
    uint32_t site0;  // cumulative sites of earlier classes
    uint32_t rec0;     // first slot of the class in the compact record table (a multiple of rstride <= 4)
    uint32_t rstride;  // compact slots per site: basis rounded up to 1, 2 or 4 (basis itself above 4)
    uint32_t pad0, pad1;
};
constexpr uint32_t SYM_UNIFORM = 1u << 31;
constexpr uint32_t SYM_SEQ = 1u << 30;   // uniform over the codes 1..nout (no table lookup needed)
static_assert(sizeof(SymClass) == 48, "SymClass layout");

// A group of classes sampled as one Bernoulli process: classes with equal probability, adjacent in
// site order, whose compact records are 4 slots per site, so a hit's record slot is
// rec0 + (site - site0) * rstride whatever its class; member k starts at group-relative site bound[k].
struct SymGroup {
    float p, inv_log;
    uint32_t site0, nsites;
    uint32_t rec0, rstride;
    uint32_t nmem, c0;
    uint32_t bound[4];   // unused members: 0xFFFFFFFF
    uint32_t nout[4];
    uint32_t woff[4];
    uint32_t flags;      // SYM_THIN, SYM_DEM_SLOTS (see below)
    uint32_t pad[3];
};
static_assert(sizeof(SymGroup) == 96, "SymGroup layout");
// Thinned group (detector error models, whose mechanisms have many distinct probabilities): the group
// is drawn at its largest probability p and a hit at a site with probability p_i is kept when the
// hit's 32-bit random word is at most the site's threshold round(p_i / p * 2^32) - 1. The threshold
// sits in slot 1 of the site's compact records (thr[] for 8-byte records).
constexpr uint32_t SYM_THIN = 1u;
// Detector-error-model sites with compact records: a mechanism keeps up to six detectors inline, two
// per record in slots 0, 2 and 3 (slot 0 also carries the observable flags; unused records have count
// 0), and every hit applies those three records.
constexpr uint32_t SYM_DEM_SLOTS = 2u;
constexpr uint32_t SYM_DEM_SLOT_CODE = 0xDu;   // slots 0, 2 and 3

// Basis fault of one column, as a fixed-fault entry for the frame kernel.
struct SymColumn {
    uint32_t op, site, code;
};

struct SymPlan {
    std::vector<SymClass> classes;
    std::vector<uint32_t> otab;        // outcome tables (threshold, basis-bitmask code)
    std::vector<SymColumn> columns;    // column -> basis fault
    std::vector<SymGroup> groups;      // sampling groups over the classes
    uint64_t total_sites = 0;
    uint64_t total_slots = 0;          // compact record slots (columns plus alignment padding)
    // op index -> (class, first class-site of the op's sites) for fixed-fault XOR; chain links map
    // to (class, link index) via chain_link.
    std::vector<uint32_t> op_class, op_site0, op_link;
    bool supported = true;
    std::string why_unsupported;
};

SymPlan plan_symptoms(const Program &prog);

// Column record (8 bytes): x bits 28..31 = entry count c; c <= 2: entries inline (x bits 0..27,
// then y); c == 3: x bits 0..27 = offset into the overflow entries, y = entry count.
constexpr uint32_t SYM_INLINE_MAX = 2;
// Compact column record (4 bytes), used when there are at most 2^14 detectors and 2 observables:
// bits 0..13 and 14..27 = up to two detector indices, bits 28..29 = their count c (0..2), bit 30 / 31 =
// the fault flips observable 0 / 1. c == 3 marks a symptom with more than two detectors: bits 0..27
// are then the offset of its entry count in the compact overflow list, followed by its entries
// (detector m, observable k as D + k); the observable bits stay zero. Classes with at most 4 basis
// columns take 4 slots per site, so one 16-byte load reads every basis record of a site.
constexpr uint32_t SYM_COMPACT_DETECTORS = 1u << 14;
constexpr uint32_t SYM_COMPACT_OBSERVABLES = 2;
constexpr uint32_t SYM_OTAB_LDS_WORDS = 1024;  // outcome tables up to this size are read from LDS
// LDS hit queue of the sampling kernel: 256 lanes x 8 hits x 8 bytes, plus a counter.
constexpr size_t SYM_QUEUE_BYTES = 256 * 8 * 8 + 16;

struct SymDevParams {
    const SymClass *classes;
    const SymGroup *groups;
    uint32_t ngroups;
    const uint32_t *otab;
    const uint2 *recs;             // one record per column
    const uint32_t *ovf;           // overflow entries
    const uint32_t *recs32;        // compact records by slot (null when not available)
    const uint32_t *ovf32;         // compact overflow list
    uint32_t ovf32_any;            // some compact record overflows (count 3)
    const uint32_t *thr;           // per-site thresholds of thinned groups (8-byte-record path)
    uint32_t nclasses;
    uint32_t otab_words;           // outcome-table words (copied to LDS when small)
    uint32_t wave_split;           // scan kernel: each wave draws its share of every group (DEM buckets)
    uint32_t D, O;
    uint32_t shots_per_tile;       // power of two
    uint32_t tile_log2;
    uint64_t total_sites;
    uint64_t shots;                // valid shots in this launch
    uint64_t tile0;
    uint64_t seed;
    DevOut dets;                   // packed rows (col0 in bits)
    DevOut obs;                    // packed rows; enabled=0 when observables share the dets rows
    uint32_t det_row_bytes;        // image pitch of the dets rows
    uint32_t obs_row_bytes;        // image pitch of separate observable rows (0 when unused)
    uint32_t obs_mode;             // 0: observables dropped, 1: separate image, 2: dets image at obs_col0
    uint32_t obs_col0;
    // Fired-mechanism rows of a detector-error-model launch (return_errors), or null: shot s of the
    // launch sets bit e of row s (err_stride_words 32-bit words apart, zeroed beforehand) for every
    // mechanism e that fires. site_mech maps a sampled site to its mechanism's index in the model.
    uint32_t *err_words;
    uint64_t err_stride_words;
    const uint32_t *site_mech;
    // Outcome records of the specialised scan kernel (null when it does not apply). A site's outcomes
    // (code 1..nout) have symptoms: the XOR of the basis columns each code selects, with entries that
    // cancel removed. Written relative to the site's lowest detector b, these symptoms repeat from
    // site to site (every round of a memory circuit repeats the one before, shifted), so each distinct
    // relative block of nout records is stored once. orec_sites holds, per site, (index of its block's
    // first record in orecs | nout << 28, b); group g's site s has its entry at orec_base[4 g] + s (the
    // group's members are consecutive in the table), and the hit's detectors are b + the record's entries.
    const uint2 *orecs;
    const uint32_t *orec_base;
    const uint2 *orec_sites;
    const uint32_t *orec_ovf;      // overflow lists: count, then entries (detector b + m, observable D + k)
};
// Outcome record (8 bytes): x = e0 | e1 << 14 | valid << 28, y = e2 | e3 << 14 | obs1 << 28 | obs0 << 29;
// entries are relative to the site's base detector, valid has bit e set for each of the n = 0..4
// entries in use (bits 0 .. n - 1), unused entries are 0. A symptom with more than four detectors has
// x = 0, y = SYM_OREC_OVERFLOW | offset of its list in orec_ovf.
constexpr uint32_t SYM_OREC_OVERFLOW = 1u << 31;

}  // namespace sa
