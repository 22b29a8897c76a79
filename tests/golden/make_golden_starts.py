"""Writes tests/golden/complex_starts.npz: the reference's start search (blml_trivial.complex_starts, default
settings: 30 random runs of 400 HIO + 100 ER steps, the symmetry-breaking runs, the best 3 and 2) on the inputs of
tests/golden/cases_starts.py.

Run in a container that has the reference checkout (read from REFERENCE_DIR, default /root/reference):
    python tests/golden/make_golden_starts.py
Each case stores its inputs and, per start, the family, the score and e0."""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
from cases_starts import starts_cases  # noqa: E402

REF = os.environ.get('REFERENCE_DIR', '/root/reference')


def main():
    sys.path.insert(0, REF)
    from blml_trivial import complex_starts
    out, meta = {}, []
    for i, (name, Z, w, h, fc, kinds) in enumerate(starts_cases()):
        starts = complex_starts(Z, w, h, fc, kinds, seed=0)
        p = f'c{i}_'
        out[p + 'score'] = np.array([s[1] for s in starts])
        out[p + 'e0'] = np.stack([s[2] for s in starts])
        meta.append(dict(name=name, family=[s[0] for s in starts]))
        print(f'{name:24s} ' + ' '.join(f'{s[0]}:{s[1]:.6f}' for s in starts), flush=True)
    out['meta'] = np.array(json.dumps(meta))
    np.savez_compressed(os.path.join(HERE, 'complex_starts.npz'), **out)


if __name__ == '__main__':
    main()
