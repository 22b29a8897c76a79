"""Writes tests/golden/empirical_fits.npz: input/output vectors of the reference estimator BLMLEmpirical
(engine='cpu', float64) for the retrieve fits no closed form pins (tests/golden/cases_empirical.py).

Run in a container that has the reference checkout (read from REFERENCE_DIR, default /root/reference):
    python tests/golden/make_golden_empirical.py
Each case stores its inputs and the reference's outputs (loglik_, density_, e_hat_, Se_, J_history_, n_outer_,
n_iter_, n_anderson_, window_norm_, plane_norm_)."""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
from cases_empirical import empirical_cases  # noqa: E402

REF = os.environ.get('REFERENCE_DIR', '/root/reference')


def main():
    sys.path.insert(0, REF)
    import blml_trivial as bt
    from blml_binned import BLMLEmpirical
    out, meta = {}, []
    for i, (name, counts, mm, est_kw, fit_kw) in enumerate(empirical_cases()):
        modes = [md for kw in mm for md in bt.make_modes(**kw)]
        m = BLMLEmpirical(engine='cpu', dtype=np.float64, **est_kw).fit(counts, modes, **fit_kw)
        p = f'c{i}_'
        out[p + 'counts'] = counts
        for a in ('loglik_', 'n_outer_', 'n_iter_', 'n_anderson_', 'window_norm_', 'plane_norm_', 'residual_'):
            out[p + a.rstrip('_')] = np.array(getattr(m, a))
        for a in ('density_', 'e_hat_', 'Se_', 'J_history_'):
            out[p + a.rstrip('_')] = np.asarray(getattr(m, a))
        meta.append(dict(name=name))
        print(f'{name:36s} loglik {m.loglik_:.12f} outer {m.n_outer_} cg {m.n_iter_} anderson {m.n_anderson_}',
              flush=True)
    out['meta'] = np.array(json.dumps(meta))
    np.savez_compressed(os.path.join(HERE, 'empirical_fits.npz'), **out)


if __name__ == '__main__':
    main()
