"""Writes clifford_faults.npz: fixed-fault measurement flips of the 10,000-qubit, 1,000-layer random
Clifford stress circuit (stim_amd.gen.random_clifford, seed 2026, the benchmark's circuit), computed
by the oracle (oracle/frame_oracle.py) with every measurement result's flip row kept.

Shot s carries three faults drawn from the circuit's noise sites (PAULI_CHANNEL_1 and DEPOLARIZE2
outcomes); the expected row is the bit-packed (little-endian) flip of every measurement result. The
GPU test replays the same faults through the measurement sampler (skip_reference_sample=True, fixed
faults disable frame randomization) and compares bit for bit. Takes a few minutes of CPU.

    python tests/golden/make_clifford_faults.py
"""

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import frame_oracle as fo  # noqa: E402
from oracle import stim_text as st  # noqa: E402
from stim_amd import gen  # noqa: E402

SHOTS = 8
FAULTS_PER_SHOT = 3


def main():
    text = gen.random_clifford(n_qubits=10000, layers=1000, seed=2026)
    flat = st.parse_flat(text)
    chans = fo.noise_sites(flat)
    rng = np.random.default_rng(2026)
    faults = []
    for _ in range(SHOTS):
        picks = rng.choice(len(chans), size=FAULTS_PER_SHOT, replace=False)
        shot = []
        for i in picks:
            k, j, outs = chans[i]
            code = outs[rng.integers(len(outs))][0]
            shot.append((int(k), int(j), int(code)))
        faults.append(shot)
    orc = fo.FrameOracle(flat)
    orc.run(SHOTS, faults=faults)
    rec = np.array(orc.last_record, dtype=bool)          # (results, shots)
    rows = np.packbits(rec.T, axis=1, bitorder='little')  # (shots, ceil(results / 8))
    np.savez_compressed(os.path.join(HERE, 'clifford_faults.npz'), faults=np.array(faults, dtype=np.uint32),
                        rows=rows, num_results=np.uint64(rec.shape[0]))
    print('results per shot', rec.shape[0], 'flips per shot', rec.sum(axis=0).tolist())


if __name__ == '__main__':
    main()
