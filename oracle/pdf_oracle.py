"""CPU oracle of the PDF statistics (test infrastructure: only tests load it, as the checker of
libwovengpu.so's wt_pdf). Restates cal_pdf (subroutine.f90:1-40) in the reference's FP32
arithmetic, with exact counts:

  bin_size = (max - min) / real(bin_num)                          real(4)      (:16)
  val(l)   = min + 0.5d0*bin_size + real(l-1)*bin_size            double sum of a real(4)
                                                                  product, stored real(4) (:18-20)
  a value goes to the first l with val(l) - 0.5*bin_size <= v < val(l) + 0.5*bin_size, real(4)  (:23-33)
  pdf(l)   = count(l) / (bin_size*nx*ny*nzp*nproc)                real(4), left to right (:37-39)

`cal_pdf` finds the first matching bin through the sorted upper edges (the smallest l with
hi(l) > v matches if lo(l) <= v, since the edges are non-decreasing); `cal_pdf_loop` is the
reference's literal linear search, for small inputs, and pins it.
"""
from __future__ import annotations

import numpy as np

f32 = np.float32


def bin_edges(values, nbins: int):
    v = np.asarray(values, dtype=np.float32).ravel()
    mn, mx = v.min(), v.max()
    bs = f32(f32(mx - mn) / f32(nbins))
    l = np.arange(nbins, dtype=np.float32)
    prod = (l * bs).astype(np.float32)
    val = ((np.float64(mn) + 0.5 * np.float64(bs)) + prod.astype(np.float64)).astype(np.float32)
    lo = (val - f32(0.5) * bs).astype(np.float32)
    hi = (val + f32(0.5) * bs).astype(np.float32)
    return mn, bs, val, lo, hi


def _pdf_from_counts(counts, bs, n_xy_zp):
    nx, ny, nzp = n_xy_zp
    denom = f32(f32(f32(bs * f32(nx)) * f32(ny)) * f32(nzp)) * f32(1.0)
    return (counts.astype(np.float32) / denom).astype(np.float32)


def cal_pdf(values, nbins: int, shape):
    """(val, pdf, counts); shape = (nx, ny, nzp) of the array the reference passes (3 components
    stacked along z: nzp = 3N)."""
    v = np.asarray(values, dtype=np.float32).ravel()
    mn, bs, val, lo, hi = bin_edges(v, nbins)
    la = np.searchsorted(hi, v, side="right")
    ok = la < nbins
    ok[ok] &= lo[la[ok]] <= v[ok]
    counts = np.bincount(la[ok], minlength=nbins).astype(np.int64)
    return val, _pdf_from_counts(counts, bs, shape), counts


def cal_pdf_loop(values, nbins: int, shape):
    """The reference's loop over points and bins, literally (small inputs only)."""
    v = np.asarray(values, dtype=np.float32).ravel()
    mn, bs, val, lo, hi = bin_edges(v, nbins)
    counts = np.zeros(nbins, dtype=np.int64)
    for x in v:
        for l in range(nbins):
            if x >= lo[l] and x < hi[l]:
                counts[l] += 1
                break
    return val, _pdf_from_counts(counts, bs, shape), counts
