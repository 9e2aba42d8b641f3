"""Test helpers: the activeOffsets packing, restated for checking exports; the array orderings
of the IO boundary and the problems their fixtures are written on."""
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


ORDERINGS = ("RowMajor", "YXZ_RowMajor", "ColumnMajor", "YXZ_ColumnMajor")


def ordering_axes(ordering, d):
    """Axis permutation P such that a field given under `ordering` is the RowMajor field (grid
    axes first, components last) transposed by P on its d grid axes, components untouched
    (IO_::GetArray, JMM_CPPLibs/Output/IO.hxx:68-110; components stay last, BaseIO.hxx:130-145):
    YXZ_RowMajor swaps the first two axes, ColumnMajor reverses them, YXZ_ColumnMajor reads
    index (i, j, k) at (k, i, j) -- the identity in 2-D."""
    ax = list(range(d))
    if ordering == "RowMajor":
        return ax
    if ordering == "YXZ_RowMajor":
        return [1, 0] + ax[2:]
    if ordering == "ColumnMajor":
        return ax[::-1]
    if ordering == "YXZ_ColumnMajor":
        return ([1, 0] + ax[2:])[::-1]
    raise ValueError(ordering)


def to_ordering(a, ordering, d):
    a = np.asarray(a)
    return np.ascontiguousarray(np.transpose(a, ordering_axes(ordering, d) + list(range(d, a.ndim))))


def from_ordering(a, ordering, d):
    a = np.asarray(a)
    inv = list(np.argsort(ordering_axes(ordering, d)))
    return np.ascontiguousarray(np.transpose(a, inv + list(range(d, a.ndim))))


def ordering_params():
    """Non-square problems for the array orderings: Riemann2 40 x 56, Riemann3 12 x 16 x 20 and
    ReedsSheppForward2 20 x 24 x 16 (fields cropped from the square generators), RowMajor, with
    every export, two geodesic tips and (ReedsSheppForward2) one sensitivity target."""
    from hamiltonfastmarching_amd import synthetic as syn
    r2 = syn.riemann2(56, seed=41)
    r3 = syn.riemann3(20, seed=42)
    rs = syn.rsf2(24, 16, seed=43)
    out = {
        "riemann2": dict(model="Riemann2", dims=np.array([40.0, 56.0]), gridScale=r2["gridScale"],
                         metric=np.ascontiguousarray(r2["metric"][:40]), seeds=np.array([[0.35, 0.5]]),
                         tips=np.array([[0.05, 0.9], [0.65, 0.1]])),
        "riemann3": dict(model="Riemann3", dims=np.array([12.0, 16.0, 20.0]), gridScale=r3["gridScale"],
                         metric=np.ascontiguousarray(r3["metric"][:12, :16]), seeds=np.array([[0.3, 0.4, 0.5]]),
                         tips=np.array([[0.05, 0.7, 0.9], [0.55, 0.1, 0.2]])),
        "rsf2": dict(model="ReedsSheppForward2", dims=np.array([20.0, 24.0, 16.0]), gridScale=rs["gridScale"],
                     origin=rs["origin"], xi=rs["xi"], eps=rs["eps"], cost=np.ascontiguousarray(rs["cost"][:20]),
                     seeds=np.array([[0.4, 0.5, 1.0]]), tips=np.array([[0.7, 0.2, 2.0], [0.1, 0.8, 4.0]]),
                     inspectSensitivity=np.array([[0.75, 0.3, 3.0]]), inspectSensitivityWeights=np.array([1.0]),
                     inspectSensitivityLengths=np.array([1.0])),
    }
    for p in out.values():
        p.update(exportValues=1.0, exportActiveNeighs=1.0, exportGeodesicFlow=1.0, exportActiveOffsets=1.0)
    return out
