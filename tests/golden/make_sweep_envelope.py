"""How far rounding alone moves the golden fits' sweep counts: each golden case (cases.py) refitted by the oracle
(which reproduces the reference's counts and likelihoods exactly, test_oracle_fits.py) with every Gram mat-vec
multiplied by (1 + 2e-16 z), z standard normal per value -- a perturbation at the last bit of FP64, below any FFT's
rounding -- for several seeds.  Writes sweep_envelope.json: per case the reference's count and stop reason and the
perturbed counts, their min and max.  A different FFT (the engine's) changes the rounding the same way, so its
counts are held to this envelope, not to the reference's exact count.
    python tests/golden/make_sweep_envelope.py"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
sys.path.insert(0, HERE)

from oracle import blml_oracle as O  # noqa: E402
from cases import golden_cases  # noqa: E402

SEEDS = 8
REL = 2e-16


def main():
    z = np.load(os.path.join(HERE, 'binned_fits.npz'), allow_pickle=False)
    meta = json.loads(str(z['meta']))
    out = []
    orig = O.BLMLBinnedOracle._S
    for i, (name, counts, mm, est_kw, fit_kw) in enumerate(golden_cases()):
        modes = [md for kw in mm for md in O.make_modes(**kw)]
        runs = []
        for seed in range(SEEDS):
            rng = np.random.default_rng(1000 + seed)

            def S(self, k, e, _r=rng):
                u = orig(self, k, e)
                return u * (1 + REL * _r.standard_normal(u.shape))
            O.BLMLBinnedOracle._S = S
            try:
                m = O.BLMLBinnedOracle(**est_kw).fit(counts, modes, **fit_kw)
            finally:
                O.BLMLBinnedOracle._S = orig
            runs.append(dict(n_sweeps=int(m.n_sweeps_), stop_reason=m.stop_reason_, loglik=float(m.loglik_)))
        ns = [r['n_sweeps'] for r in runs]
        out.append(dict(name=name, ref_n_sweeps=int(z[f'c{i}_n_sweeps']), ref_stop_reason=meta[i]['stop_reason'],
                        perturbed=runs, min=min(ns), max=max(ns)))
        print(name, int(z[f'c{i}_n_sweeps']), ns, flush=True)
    with open(os.path.join(HERE, 'sweep_envelope.json'), 'w') as f:
        json.dump(dict(rel=REL, seeds=SEEDS, cases=out), f, indent=1)


if __name__ == '__main__':
    main()
