"""Reference cropping done by the realign caller before each pair-HMM call.

getAlignedPairsWithIndelsCroppingReference (polisher.c:605-654): the
reference is cropped to [a0.x - a0.y, last.x + (readLen - last.y) + 1)
(clamped to the reference), anchors are shifted into the crop, and outputs
are shifted back by the crop start afterwards."""
from __future__ import annotations

import numpy as np


def crop_bounds(anchors, ref_len, read_len):
    a = np.asarray(anchors, dtype=np.int64).reshape(-1, 3)
    if a.shape[0] == 0:
        return 0, ref_len
    first = max(int(a[0, 0] - a[0, 1]), 0)
    end = min(int(1 + a[-1, 0] + (read_len - a[-1, 1])), ref_len)
    return first, end


def crop_pair(ref_sym, read_sym, anchors):
    """Returns (sX, sY, shifted anchors, ref offset)."""
    first, end = crop_bounds(anchors, len(ref_sym), len(read_sym))
    a = np.array(anchors, dtype=np.int64).reshape(-1, 3)
    a[:, 0] -= first
    return np.ascontiguousarray(ref_sym[first:end]), np.ascontiguousarray(read_sym), a, first
