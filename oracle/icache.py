"""CPU restatement of the irradiance cache of ICIntegrator (test
infrastructure only: used by tests/ as the checker of ilm_ic_*).

Follows Integrator/ICIntegrator.cpp:
  - root cell: scene centre +- longest edge, each corner scaled by
    1 + FLT_EPSILON (ICIntegrator::Initialise, :306-323);
  - children in SetBounds order (:54-86), SphereBoxOverlap (:88-101),
    Insert (:107-140): a record stays at a node when the depth is spent or
    its RiClamp exceeds the node's half-size (AxisAlignedBoundingBox::
    GetExtent, Geometry/BoundingBox.cpp:421-429), else goes to every child its
    sphere overlaps;
  - FindRecords (:142-183) with W_Ward (:185-196): records with w > 0 on the
    root-to-leaf path of the point, averaged by weight (GetIrradiance,
    :478-486);
  - MLICIntegrator as the reference compiles it (Tabelion weights,
    MLICIntegrator.cpp:28-34): w = 1 - max(d / RiClamp, 9.0124 (1 - cos)) /
    threshold (W_Tabelion, :285-318), insertion by the half-size along x
    (Insert, :172-176), Ri = the nearest gather hit (ComputeRecord, :717-725);
  - ComputeRecord's formulas (:498-543): stratified cosine directions
    (Montecarlo::CosineSampleHemisphere with strata, Maths/Montecarlo.h:20-29),
    E = mean, Ri = mn / sum(1 / distance), RiClamp = clamp(Ri, RMin, RMax).
Arithmetic is float32, one rounding per operation, as the reference's float
code without contraction.
"""
from __future__ import annotations

import numpy as np

F = np.float32
FLT_EPSILON = F(np.finfo(np.float32).eps)


def root_cell(pos):
    """Root cell and (RMin, RMax) factors' base length from vertex positions."""
    pos = np.asarray(pos, np.float32).reshape(-1, 3)
    lo, hi = pos.min(axis=0), pos.max(axis=0)
    ext = (hi - lo) * F(0.5)
    c = lo + ext
    L = F(np.max(np.abs(hi - lo)))
    a = (c - L) * (F(1) + FLT_EPSILON)
    b = (c + L) * (F(1) + FLT_EPSILON)
    return np.minimum(a, b).astype(np.float32), np.maximum(a, b).astype(np.float32), L


def child_bounds(lo, hi, i):
    c = lo + (hi - lo) * F(0.5)
    bits = [i & 1, (i >> 2) & 1, (i >> 1) & 1]   # x, y, z
    clo = np.where(bits, c, lo).astype(np.float32)
    chi = np.where(bits, hi, c).astype(np.float32)
    return clo, chi


def sphere_box(lo, hi, c, r):
    d = F(0)
    for i in range(3):
        if c[i] < lo[i]:
            d = F(d + F(c[i] - lo[i]) * F(c[i] - lo[i]))
        elif c[i] > hi[i]:
            d = F(d + F(c[i] - hi[i]) * F(c[i] - hi[i]))
    return d <= F(r) * F(r)


class Octree:
    def __init__(self, lo, hi, depth, variant="ward"):
        self.variant = variant
        self.lo, self.hi, self.child, self.recs = [np.asarray(lo, np.float32)], [np.asarray(hi, np.float32)], [-1], [[]]
        self.depth = depth
        self.P, self.N, self.E, self.rc = [], [], [], []

    def half(self, n):
        if self.variant == "tabelion":
            return F((self.hi[n][0] - self.lo[n][0]) * F(0.5))
        return F(np.max(np.abs((self.hi[n] - self.lo[n]) * F(0.5))))

    def insert(self, P, N, E, rc):
        self.P.append(np.asarray(P, np.float32))
        self.N.append(np.asarray(N, np.float32))
        self.E.append(np.asarray(E, np.float32))
        self.rc.append(F(rc))
        self._insert(0, len(self.P) - 1, self.depth)

    def _insert(self, n, r, depth):
        if depth <= 0 or self.rc[r] > self.half(n):
            self.recs[n].append(r)
            return
        if self.child[n] < 0:
            first = len(self.lo)
            self.child[n] = first
            for i in range(8):
                clo, chi = child_bounds(self.lo[n], self.hi[n], i)
                self.lo.append(clo)
                self.hi.append(chi)
                self.child.append(-1)
                self.recs.append([])
        first = self.child[n]
        for i in range(8):
            if sphere_box(self.lo[first + i], self.hi[first + i], self.P[r], self.rc[r]):
                self._insert(first + i, r, depth - 1)

    @staticmethod
    def _contains(lo, hi, p):
        return not (np.any(lo > p) or np.any(hi < p))

    def weight(self, p, n, r, inv_threshold):
        """W_Ward: 1 / (d / RiClamp + 1 - cos) - 1 / threshold; -1 when cos <= 0."""
        d = self.P[r] - p
        dist = F(np.sqrt(F(F(d[0] * d[0]) + F(d[1] * d[1]) + F(d[2] * d[2])))) / self.rc[r]
        cos = F(F(n[0] * self.N[r][0]) + F(n[1] * self.N[r][1]) + F(n[2] * self.N[r][2]))
        if self.variant == "tabelion":
            return F(1) - F(inv_threshold) * max(dist, F(F(F(1) - cos) * F(9.0124)))
        if cos <= 0:
            return F(-1)
        return F(F(1) / F(F(dist + F(1)) - cos)) - F(inv_threshold)

    def lookup(self, p, n, threshold):
        """(weighted mean irradiance, weight sum, accepted record ids)."""
        p, n = np.asarray(p, np.float32), np.asarray(n, np.float32)
        inv = F(1) / F(threshold)
        num, den, ids = np.zeros(3, np.float64), 0.0, []
        if not self._contains(self.lo[0], self.hi[0], p):
            return num, 0.0, ids
        node = 0
        while True:
            for r in self.recs[node]:
                w = self.weight(p, n, r, inv)
                if w > 0:
                    num += float(w) * self.E[r].astype(np.float64)
                    den += float(w)
                    ids.append(r)
            if self.child[node] < 0:
                break
            c = self.child[node]
            nxt = next((c + i for i in range(8) if self._contains(self.lo[c + i], self.hi[c + i], p)), None)
            if nxt is None:
                break
            node = nxt
        return (num / den if den > 0 else num), den, ids


def cosine_stratified(u, v, alt, az, alt_n, az_n):
    """Montecarlo::CosineSampleHemisphere(u, v, altitude, azimuth, m, n)."""
    a = (u + alt) / alt_n
    b = (v + az) / az_n
    c = 2 * np.pi * b
    d = np.sqrt(a)
    return np.stack([d * np.cos(c), d * np.sin(c), np.sqrt(np.maximum(0.0, 1 - a))], axis=-1)


def record(radiance, first_t, rmin, rmax, variant="ward"):
    """ComputeRecord: E = mean radiance, Ri = mn / sum(1 / t) (Ward) or the
    nearest hit (Tabelion), clamped radius."""
    radiance = np.asarray(radiance, np.float64).reshape(-1, 3)
    mn = len(radiance)
    t = np.asarray(first_t, np.float64)
    ri = float(t.min()) if variant == "tabelion" else mn / np.sum(1.0 / t)
    return radiance.mean(axis=0), ri, min(max(ri, rmin), rmax)
