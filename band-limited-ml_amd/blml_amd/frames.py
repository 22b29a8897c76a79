"""Counts frames for the binned fit: the loader and the train / validation / test thinning of the reference's frame
jobs (binned_gpu_job.py:100-110, frame_fc_gpu_job.py:22-24), and their score.

    N = load_frame(path, crop=257)              # int64 photon counts (a 2-D .npy, e.g. psf_plane5.npy)
    tr, va, te = split_frame(N, seed=0)         # binomial 70 / 10 / 20 thinning, numpy default_rng(seed)
    score(P, te)                                # mean log probability per count

The thinning keeps the reference's draws (numpy default_rng(seed): binomial(N, 0.7), then binomial of the rest
with 1/3), so a frame splits into the same three count sets as in the reference's jobs; it is host work before the
fit, a few tens of milliseconds for a 2048 x 2048 frame.  The loader never unpickles (allow_pickle=False)."""
import numpy as np

__all__ = ['load_frame', 'split_frame', 'score', 'FRAME_CENTRE']

FRAME_CENTRE = (661, 1068)     # the crop centre of the reference's bead frame (binned_gpu_job.py:105)


def load_frame(path, crop=0, centre=FRAME_CENTRE):
    """The frame's counts as int64 (a 2-D numeric .npy); crop > 0: the crop x crop window (odd sizes centred)
    around centre, as the reference's BLML_FRAME_CROP."""
    N = np.load(path, allow_pickle=False)
    if N.ndim != 2 or not np.issubdtype(N.dtype, np.number):
        raise ValueError(f"{path}: a 2-D numeric array of counts, got {N.dtype} {N.shape}")
    if np.any(N < 0) or not np.all(np.isfinite(N)):
        raise ValueError(f"{path}: counts must be finite and non-negative")
    N = N.astype(np.int64)
    if crop:
        y0, x0 = centre
        H = int(crop) // 2
        if y0 - H < 0 or x0 - H < 0 or y0 + H + 1 > N.shape[0] or x0 + H + 1 > N.shape[1]:
            raise ValueError(f"crop {crop} around {centre} leaves the frame {N.shape}")
        N = N[y0 - H:y0 + H + 1, x0 - H:x0 + H + 1]
    return N


def split_frame(N, seed=0, train=0.7, val_of_rest=1.0 / 3.0):
    """(train, validation, test) counts: train ~ binomial(N, train), validation ~ binomial(N - train, val_of_rest),
    test the rest (the 70 / 10 / 20 split of the reference's jobs at the defaults)."""
    N = np.asarray(N)
    if not np.issubdtype(N.dtype, np.integer):
        if not np.allclose(N, np.round(N)):
            raise ValueError("split_frame: integer counts")
        N = np.round(N).astype(np.int64)
    r = np.random.default_rng(seed)
    tr = r.binomial(N, train)
    rest = N - tr
    va = r.binomial(rest, val_of_rest)
    return tr, va, rest - va


def score(P, N):
    """Mean log probability per count: sum N log max(P, 1e-300) / sum N (P normalised over the frame)."""
    P = np.asarray(P, float)
    N = np.asarray(N, float)
    return float((N * np.log(np.maximum(P, 1e-300))).sum() / N.sum())
