"""Writes tests/golden/binning.npz: the reference's binning of points for a counts fit (blml_gpu.py:128-176 as the
reference's own functions compute it: parse_modes, lattice_target, lattice_fold_specs, lattice_kernel(...).h and
.quantize, the unique centres by their bytes) on the inputs of tests/golden/cases_binning.py.

Run in a container that has the reference checkout (read from REFERENCE_DIR, default /root/reference):
    python tests/golden/make_golden_binning.py"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
from cases_binning import binning_cases  # noqa: E402

REF = os.environ.get('REFERENCE_DIR', '/root/reference')


def main():
    sys.path.insert(0, REF)
    import blml_trivial as bt
    out, meta = {}, []
    for i, (name, X, modes, bin_scale) in enumerate(binning_cases(bt.make_modes)):
        d = X.shape[1]
        specs, _ = bt.parse_modes(modes, d)
        kinds, periods = specs[0]['kinds'], specs[0]['periods']
        shared = bt._shared_kernel(specs)
        fc_lat = (np.asarray(specs[0]['fc'], float) if shared
                  else np.max([np.asarray(sp['fc'], float) for sp in specs], 0))
        n = float(len(X))
        h_t = bt.lattice_target(bin_scale * n ** -0.25 / fc_lat, [sp['kinds'] for sp in specs])
        kern = bt.lattice_kernel(fc_lat, h_t, kinds, periods, d, bt.lattice_fold_specs(specs))
        Xq = np.ascontiguousarray(kern.quantize(X))
        vdt = np.dtype((np.void, Xq.dtype.itemsize * d))
        _, idx, cnt = np.unique(Xq.view(vdt), return_index=True, return_counts=True)
        p = f'c{i}_'
        out[p + 'h'] = np.asarray(kern.h, float)
        out[p + 'h_target'] = np.asarray(h_t, float)
        out[p + 'Z'] = Xq[idx]
        out[p + 'w'] = cnt.astype(float)
        meta.append(dict(name=name))
        print(f'{name:28s} h {np.round(kern.h, 5)} bins {len(idx)}', flush=True)
    out['meta'] = np.array(json.dumps(meta))
    np.savez_compressed(os.path.join(HERE, 'binning.npz'), **out)


if __name__ == '__main__':
    main()
