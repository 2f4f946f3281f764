"""Write tests/golden/type21_isis_sample.npz: consecutive records of two SPK
Type 21 segments of asteroid (42) Isis, taken from the DAF file kete ships
as docs/data/20000042.bsp (used by docs/tutorials/propagation_plots.py).

The sample is data only: segment summaries, records (kete's layout,
spk_segments.rs:1010-1018) and final epochs.  Run in a checkout that holds
the kete sources:  python tests/golden/make_type21_sample.py [path/to/20000042.bsp]
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
from kete_amd.spk import read_spk  # noqa: E402

SRC = sys.argv[1] if len(sys.argv) > 1 else "/root/reference/docs/data/20000042.bsp"
# (segment index in file order, first record, record count)
PICKS = [(0, 200, 16),   # backward-integrated segment: reference epoch at record start
         (6, 0, 14)]     # forward segment from its first step: tiny steps, orders 2..15


def main():
    segs = read_spk(SRC)
    out = {"n_segments_in_file": np.int64(len(segs)),
           "file_summary": np.array([[s.obj_id, s.center_id, s.frame, s.seg_type, s.n_rec,
                                      s.n_coef] for s in segs], dtype=np.int64),
           "file_span_s": np.array([[s.start_s, s.end_s] for s in segs])}
    for k, (si, r0, nr) in enumerate(PICKS):
        s = segs[si]
        ep = s.epochs()
        out[f"s{k}_ids"] = np.array([s.obj_id, s.center_id, s.frame, s.seg_type, s.n_coef],
                                    dtype=np.int64)
        start = s.start_s if r0 == 0 else ep[r0 - 1]
        out[f"s{k}_span_s"] = np.array([start, ep[r0 + nr - 1]])
        out[f"s{k}_records"] = s.records()[r0: r0 + nr].copy()
        out[f"s{k}_epochs"] = ep[r0: r0 + nr].copy()
    np.savez_compressed(os.path.join(HERE, "type21_isis_sample.npz"), **out)


if __name__ == "__main__":
    main()
