// Measurement records from detection events: the measurement sampler's fast path.
//
// The reference samples measurements with its frame simulator (sample_batch_measurements,
// src/stim/simulators/frame_simulator_util.inl:290-372): results = reference sample XOR the frame's
// flips, where the flips are the noise's effect x plus a uniformly random element g of the gauge space
// G that frame randomization spans (the random measurement results). Detectors and observables are
// parities A of the record that noiseless runs fix (A g = 0), so A x = (detection events, observable
// flips) = y. When every parity the noiseless circuit fixes is a combination of the detectors and
// observables, ker A = G, and a record with the right distribution is any solution of A m' = y with
// the solution's free part uniformly random: m' = x + (element of ker A), uniform over x + G.
//
// The plan solves A m' = y by forward substitution in record order. Gaussian elimination over GF(2)
// gives each independent row a pivot (its largest measurement index after reduction); a pivot's result
// is the XOR of the y values its reduced row combines and of the reduced row's other, earlier
// measurements. Measurements that no row pivots on are free: uniformly random words.
//
// Substitution schedule: pivots with earlier terms (items) are owned by threads of the expansion
// workgroup and run in phases separated by barriers. An item joins the thread that owns its latest
// dependency; its phase is the largest phase of its dependencies, plus one for each dependency owned by
// another thread. A thread runs its items of a phase in record order, so a chain of dependencies on
// one thread needs no barrier (a memory circuit's checks repeat round after round: one chain per check).
// A run of items on one thread that each use the previous one's result becomes a chain record; the
// items' other terms, where they are final when the phase starts, are XORed in beforehand by
// independent records in a phase of their own (at d=25: 5 phases instead of 48 substitution levels,
// no thread with more than 2 records in a phase). ker A = G is
// checked by dimension. G lies in ker A (the sampler checks that detectors and observables are
// deterministic), so dim G <= nullity(A) = the free count. The reference run's random results are
// random whatever the earlier results were (which results are random does not depend on outcomes in a
// stabilizer circuit), so the record projected onto them takes every value: dim G >= their number R.
// A free count equal to R therefore gives dim G = nullity(A), ker A = G. (A random reset that leaks
// into later results raises dim G above R, and then the free count cannot equal R.)
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "circuit.h"
#include "m2d.h"

namespace sa {

struct ExpandPlan {
    bool complete = false;
    std::string why;                  // reason when not complete
    uint32_t M = 0;                   // measurements
    uint32_t Y = 0;                   // y columns: detectors, then observables
    uint32_t pivots = 0;
    // y column j's value is XORed into result prim[j] & 0xFFFF (none when prim[j] is 0xFFFFFFFF) and,
    // when bit 31 is set, into results extra[extra_off[j] .. extra_off[j+1]) as well.
    std::vector<uint32_t> prim, extra_off, extra;
    std::vector<uint32_t> free_slots;  // results drawn uniformly at random
    // Substitution records, 8 words each (EXPAND_RECORD_WORDS). With n = word 0 >> 16 terms and
    // k = word 0 & 0xFFFF:
    // n = 1: a chain of (word 2) results k, k + w, k + 2w, ... (w = word 3): the first is XORed with
    // result word 1 (M: a word that is always zero), each later one with the result before it in the
    // chain (a check repeated round after round);
    // 2 <= n <= 7: result k ^= the XOR of results words 1 .. n (the terms inline);
    // n >= 8: result k ^= the XOR of results terms[word 1 .. word 1 + n).
    // Unused words are 0. Thread t runs records [sched_off[p * threads + t], sched_off[p * threads + t +
    // 1]) in phase p.
    uint32_t threads = 0, phases = 0;
    std::vector<uint32_t> sched_off, items, terms;
    uint32_t levels = 0;               // substitution depth (longest dependency chain), for reports
};

// Plans the expansion for a circuit with M measurements, D detectors and O observables, given the
// reference run's random-collapse counts (random_total is reported, not needed). Observables with
// Pauli-frame terms are left out of A.
// A plan whose elimination fills in beyond a fixed budget is reported incomplete.
ExpandPlan plan_expansion(const std::vector<const Instr *> &flat, uint32_t M, uint32_t D, uint32_t O,
                          uint64_t random_recorded, uint64_t random_total, uint32_t threads = 1024);

// The expansion of one shot on the host, as the kernel computes it (one byte per bit): y = the Y
// column values, free_bits = one value per free result in plan order; writes the M results (without
// any reference inversion).
void expand_shot_host(const ExpandPlan &E, const uint8_t *y, const uint8_t *free_bits, uint8_t *m);

}  // namespace sa
