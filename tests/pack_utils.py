"""Test helper: the activeOffsets packing, restated for checking exports."""
import numpy as np


def pack_active_offsets(active, offs, nfwd, dim):
    """activeOffsets from activeNeighs bits and per-point stencil offsets (N, NC, dim): the
    active differences in slot order, a symmetric one with the sign of its active side (bit
    nfwd + 2 (k - nfwd) is the + side), each offset packed big-endian into 8-bit fields
    (HFMInterface.hxx:554-577, EulerianStencil.hxx:176-200)."""
    active = np.asarray(active).astype(np.uint64).ravel()
    N, NC = offs.shape[:2]
    out = np.zeros((N, NC))
    j = np.zeros(N, np.int64)
    shifts = np.array([8 * (dim - 1 - q) for q in range(dim)], np.uint64)

    def packed(o):
        return ((o.astype(np.int64) & 0xFF).astype(np.uint64) << shifts).sum(-1).astype(np.float64)

    for k in range(NC):
        sides = [(k, 1)] if k < nfwd else [(nfwd + 2 * (k - nfwd), 1), (nfwd + 2 * (k - nfwd) + 1, -1)]
        for bit, sg in sides:
            on = ((active >> np.uint64(bit)) & np.uint64(1)).astype(bool)
            out[np.nonzero(on)[0], j[on]] = packed(sg * offs[on, k].astype(np.int64))
            j[on] += 1
    return out
