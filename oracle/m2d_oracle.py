"""Test infrastructure only: a CPU restatement of the reference's measurement-to-detection-event
conversion, used by tests/ as the checker for the GPU converter; never by the product path.

Follows measurements_to_detection_events_helper (src/stim/simulators/measurements_to_detection_events.inl:29-131)
without sweep data: detector d = XOR of the measurement results it names, inverted when the XOR of
their reference-sample bits is 1; observable k likewise over the measurement records that
OBSERVABLE_INCLUDE(k) names (Pauli targets add nothing without sweep data).

Pinned by: the reference CLI's `stim m2d` outputs on seeded random measurement data
(tests/golden/m2d_vectors.json), bit-exact, with and without the reference sample.
"""

import numpy as np

from oracle import stim_text as st


def columns(text):
    """(num_measurements, detector measurement lists, observable measurement lists)."""
    flat = st.parse_flat(text)
    m = 0
    dets = []
    obs = {}
    for ins in flat:
        if ins.name == 'DETECTOR':
            dets.append([m - t.value for t in ins.targets if t.kind == st.REC])
        elif ins.name == 'OBSERVABLE_INCLUDE':
            obs.setdefault(int(ins.args[0]), []).extend(m - t.value for t in ins.targets if t.kind == st.REC)
        m += st.num_results(ins)
    no = max(obs) + 1 if obs else 0
    return m, dets, [obs.get(k, []) for k in range(no)]


def convert(text, measurements, reference_sample):
    """measurements: bool (shots, M); reference_sample: bool (M,). Returns bool (shots, D + O)."""
    m, dets, obs = columns(text)
    meas = np.asarray(measurements, dtype=bool)
    ref = np.asarray(reference_sample, dtype=bool)
    out = np.zeros((meas.shape[0], len(dets) + len(obs)), dtype=bool)
    for j, idx in enumerate(dets + obs):
        col = np.zeros(meas.shape[0], dtype=bool)
        parity = False
        for i in idx:
            col ^= meas[:, i]
            parity ^= bool(ref[i])
        out[:, j] = col ^ parity
    return out
