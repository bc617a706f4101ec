"""Error measures of a GPU state against the FP64 oracle's state (shared by
the accuracy tests; not a test module).

errors(): max and mean |error| of h, qx, qy over interior cells, relative
mass error, wet/dry mask mismatches (h > hextra), and the max |dh| split
into "inner" (cells wet in both runs and more than 3 cells from a wet/dry
edge of either run) and "shoreline" (that band)."""
import numpy as np
from scipy.ndimage import binary_dilation


def errors(case, ref, got, hextra):
    out = {}
    for k in ("h", "qx", "qy"):
        d = np.abs(got[k][1:-1, 1:-1] - ref[k][1:-1, 1:-1])
        out[f"max_d{k}"] = float(d.max())
        out[f"mean_d{k}"] = float(d.mean())
    m_ref = float(ref["h"][1:-1, 1:-1].sum())
    m_got = float(got["h"][1:-1, 1:-1].sum())
    out["rel_mass"] = abs(m_got - m_ref) / m_ref if m_ref > 0 else 0.0
    wr, wg = ref["h"][1:-1, 1:-1] > hextra, got["h"][1:-1, 1:-1] > hextra
    out["wet_mismatch"] = int((wr != wg).sum())
    out["cells"] = int(wr.size)
    edges = (wr ^ binary_dilation(wr)) | (wg ^ binary_dilation(wg)) | (wr != wg)
    band = binary_dilation(edges, iterations=3)
    inner = wr & wg & ~band
    dh = np.abs(got["h"][1:-1, 1:-1] - ref["h"][1:-1, 1:-1])
    out["max_dh_inner"] = float(dh[inner].max()) if inner.any() else 0.0
    out["max_dh_shoreline"] = float(dh[band].max()) if band.any() else 0.0
    return out


def mean_dh_vs_analytic(h_got, h_exact, hextra):
    """Mean |h - exact| over cells wet in either (interior arrays)."""
    wet = (h_exact > 0) | (h_got > hextra)
    return float(np.abs(h_got - h_exact)[wet].mean())
