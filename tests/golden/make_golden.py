"""Writes tests/golden/binned_fits.npz: input/output vectors of the reference estimator BLMLBinned (engine='cpu',
float64) for the fits no closed form pins -- complex and rank-K fits, which are non-convex, so their local optimum
depends on the exact start, trajectory and stop rules (blml_binned.py:223-445).  Real rank-1 fits, the kernels, the
mat-vec and the certificate are pinned without the reference (closed forms, dense sums, eigvalsh) and are not here.

Run in a container that has the reference checkout (read from REFERENCE_DIR, default /root/reference):
    python tests/golden/make_golden.py
Each case stores its inputs (counts, make_modes arguments, fit keywords) and the reference's outputs (loglik_,
density_, e_hat_, Se_, n_sweeps_, stop_reason_).  The cases are defined in tests/golden/cases.py."""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
from cases import golden_cases  # noqa: E402

REF = os.environ.get('REFERENCE_DIR', '/root/reference')


def main():
    sys.path.insert(0, REF)
    import blml_trivial as bt
    from blml_binned import BLMLBinned
    out = {}
    meta = []
    for i, (name, counts, mm, est_kw, fit_kw) in enumerate(golden_cases()):
        modes = [md for kw in mm for md in bt.make_modes(**kw)]
        m = BLMLBinned(engine='cpu', dtype=np.float64, **est_kw).fit(counts, modes, **fit_kw)
        p = f'c{i}_'
        out[p + 'counts'] = counts
        out[p + 'loglik'] = np.array(m.loglik_)
        out[p + 'density'] = np.asarray(m.density_)
        out[p + 'e_hat'] = np.asarray(m.e_hat_)
        out[p + 'Se'] = np.asarray(m.Se_)
        out[p + 'n_sweeps'] = np.array(m.n_sweeps_)
        meta.append(dict(name=name, stop_reason=m.stop_reason_, pad=list(m.pad_), shared=bool(m.shared_)))
        print(f'{name:40s} loglik {m.loglik_:.12f} sweeps {m.n_sweeps_} stop {m.stop_reason_}', flush=True)
    out['meta'] = np.array(json.dumps(meta))
    np.savez_compressed(os.path.join(HERE, 'binned_fits.npz'), **out)


if __name__ == '__main__':
    main()
