"""Oracle for the irradiance adjoint: FP64 restatement (numpy, vectorised over
items) of the chain solve and the irradiance contribution, for central finite
differences of the image loss with respect to heightfield heights.
TEST INFRASTRUCTURE ONLY.

The chain constraint is solveConstraints' residual (pk.h:505-548), solved by
Gauss-Newton in FP64 to round-off; the contribution is the Fresnel throughput
(pk.h:809-813), dA/dw (pk.h:844-930), the receiver cosine and the clamp
(main_gpu.cu:458-471), with the tuple set, visibility and validity of a
forward solve held fixed.  Written independently of the kernels' C++ code:
forward-mode numbers are numpy arrays with a tangent axis.
"""
from __future__ import annotations

import numpy as np


class D:
    """Vectorised forward-mode number: value v (n,), tangents g (n, T)."""
    __slots__ = ("v", "g")

    def __init__(self, v, g):
        self.v = v
        self.g = g

    @staticmethod
    def const(v, T):
        v = np.asarray(v, dtype=np.float64)
        return D(v, np.zeros(v.shape + (T,)))

    def __add__(self, o):
        return D(self.v + o.v, self.g + o.g)

    def __sub__(self, o):
        return D(self.v - o.v, self.g - o.g)

    def __neg__(self):
        return D(-self.v, -self.g)

    def __mul__(self, o):
        return D(self.v * o.v, self.g * o.v[:, None] + o.g * self.v[:, None])

    def __truediv__(self, o):
        r = self.v / o.v
        return D(r, (self.g - o.g * r[:, None]) / o.v[:, None])

    def sqrt(self):
        r = np.sqrt(self.v)
        return D(r, self.g * (0.5 / r)[:, None])

    def where(self, cond, other):
        return D(np.where(cond, self.v, other.v), np.where(cond[:, None], self.g, other.g))

    def scale(self, s):
        return D(self.v * s, self.g * s)


class V:
    def __init__(self, x, y, z):
        self.x, self.y, self.z = x, y, z

    @staticmethod
    def const(a, T):  # a (n, 3)
        return V(D.const(a[:, 0], T), D.const(a[:, 1], T), D.const(a[:, 2], T))

    def __add__(self, o):
        return V(self.x + o.x, self.y + o.y, self.z + o.z)

    def __sub__(self, o):
        return V(self.x - o.x, self.y - o.y, self.z - o.z)

    def __neg__(self):
        return V(-self.x, -self.y, -self.z)

    def mul(self, s):
        return V(self.x * s, self.y * s, self.z * s)

    def div(self, s):
        return V(self.x / s, self.y / s, self.z / s)

    def dot(self, o):
        return self.x * o.x + self.y * o.y + self.z * o.z

    def cross(self, o):
        return V(self.y * o.z - self.z * o.y, self.z * o.x - self.x * o.z, self.x * o.y - self.y * o.x)

    def where(self, cond, o):
        return V(self.x.where(cond, o.x), self.y.where(cond, o.y), self.z.where(cond, o.z))


def _one(n, T):
    return D.const(np.ones(n), T)


def reflect(wi, n):
    T = wi.x.g.shape[1]
    two = D.const(np.full(wi.x.v.shape, 2.0), T)
    return n.mul(wi.dot(n) * two / n.dot(n)) - wi


def refract(wi, n, eta):
    """saka refraction_norm_free (saka.h:197-208): zero where k < 0."""
    non, win, ww = n.dot(n), wi.dot(n), wi.dot(wi)
    k = non * ww * (eta * eta - _one(len(eta.v), eta.g.shape[1])) + win * win
    tir = k.v < 0.0
    k = D(np.where(tir, 1.0, k.v), np.where(tir[:, None], 0.0, k.g))
    r = (-wi).mul(non) + n.mul(win - k.sqrt())
    zero = D.const(np.zeros(len(eta.v)), eta.g.shape[1])
    return r.where(~tir, V(zero, zero, zero))


def _at(p, u, v):  # p: list of 3 V; u, v: D
    return p[0] + (p[1] - p[0]).mul(u) + (p[2] - p[0]).mul(v)


def residual(x, verts, nrms, pb, pe, eta, event, K):
    """x: list of 2K D; verts/nrms: [K][3] V; returns 3K D."""
    vert = [pb] + [_at(verts[k], x[2 * k], x[2 * k + 1]) for k in range(K)] + [pe]
    F = []
    for k in range(K):
        wi = vert[k] - vert[k + 1]
        wo = vert[k + 2] - vert[k + 1]
        n = _at(nrms[k], x[2 * k], x[2 * k + 1])
        swap = wi.dot(n).v < 0.0
        wi, wo = wo.where(swap, wi), wi.where(swap, wo)
        is_t = ((event >> k) & 1).astype(bool)
        wopt = refract(wi, n, eta).where(is_t, reflect(wi, n))
        c = wo.cross(wopt)
        F += [c.x, c.y, c.z]
    return F


def solve(x0, verts, nrms, pb, pe, eta, event, K, iters=8):
    """Gauss-Newton in FP64 from x0 (n, 2K) to round-off: at most `iters`
    steps; an item leaves the iteration once its step is below 1e-15."""
    x = np.array(x0, dtype=np.float64)
    n = x.shape[0]
    P = 2 * K
    active = np.arange(n)
    for _ in range(iters):
        if active.size == 0:
            break
        m = active.size
        xs = [D(x[active, i], np.eye(P)[i][None, :].repeat(m, 0)) for i in range(P)]
        vv = [[V.const(verts[k][active, c], P) for c in range(3)] for k in range(K)]
        nn = [[V.const(nrms[k][active, c], P) for c in range(3)] for k in range(K)]
        F = residual(xs, vv, nn, V.const(pb[active], P), V.const(pe[active], P), D.const(eta[active], P),
                     event[active], K)
        A = np.stack([f.g for f in F], axis=1)          # (m, 3K, P)
        b = np.stack([f.v for f in F], axis=1)          # (m, 3K)
        # an item at a square root of zero (grazing refraction) has no finite
        # Jacobian: it keeps its x, and its contribution comes out non-finite
        ok = np.all(np.isfinite(A), axis=(1, 2)) & np.all(np.isfinite(b), axis=1)
        dx = np.zeros((m, P))
        dx[ok] = np.einsum("nij,nj->ni", np.linalg.pinv(A[ok]), b[ok])
        x[active] = x[active] - dx
        moving = ok & (np.max(np.abs(dx), axis=1) > 1e-15)
        active = active[moving]
    return x


def fresnel(wi, n, eta):
    nn, ww = n.dot(n), wi.dot(wi)
    c = n.dot(wi)
    flip = c.v < 0.0
    one = _one(len(c.v), c.g.shape[1])
    c = (-c).where(flip, c)
    n = (-n).where(flip, n)
    eta = (one / eta).where(flip, eta)
    k = nn * ww * (eta * eta - one) + c * c
    tir = k.v < 0.0
    k = D(np.where(tir, 1.0, k.v), np.where(tir[:, None], 0.0, k.g))
    g = k.sqrt()
    gmc, gpc = g - c, g + c
    a = gmc / gpc
    q = (c * gpc - nn * ww) / (c * gmc + nn * ww)
    half = D.const(np.full(len(c.v), 0.5), c.g.shape[1])
    r = half * (a * a) * (one + q * q)
    return r.where(~tir, one)


def footprint(ro, rd, pe, verts, nrms, eta, event, K):
    """dAdw (pk.h:844-930) on FP64 arrays: ro, rd, pe (n, 3); verts/nrms [K] (n, 3, 3)."""
    n = ro.shape[0]
    rd = rd / np.linalg.norm(rd, axis=1, keepdims=True)
    sign = np.where(rd[:, 2] < 0.0, -1.0, 1.0)
    a = -1.0 / (sign + rd[:, 2])
    b = rd[:, 0] * rd[:, 1] * a
    T0 = np.stack([1.0 + sign * rd[:, 0] ** 2 * a, sign * b, -sign * rd[:, 0]], 1)
    T1 = np.stack([b, sign + rd[:, 1] ** 2 * a, -rd[:, 1]], 1)
    ro_j = V.const(ro, 2)
    rd_j = V(D(rd[:, 0], np.stack([T0[:, 0], T1[:, 0]], 1)), D(rd[:, 1], np.stack([T0[:, 1], T1[:, 1]], 1)),
             D(rd[:, 2], np.stack([T0[:, 2], T1[:, 2]], 1)))
    eta_d = D.const(eta, 2)
    one = _one(n, 2)
    dead = np.zeros(n, dtype=bool)
    for j in range(K):
        v0, v1, v2 = (V.const(verts[j][:, c], 2) for c in range(3))
        ng = (v1 - v0).cross(v2 - v0)
        t = (v0 - ro_j).dot(ng) / ng.dot(rd_j)
        p = ro_j + rd_j.mul(t)
        v = (p - v1).dot(ng.cross(v1 - v0))
        w = (p - v2).dot(ng.cross(v2 - v1))
        u = (p - v0).dot(ng.cross(v0 - v2))
        area = u + v + w
        u, v, w = u / area, v / area, w / area
        s0, s1, s2 = (V.const(nrms[j][:, c], 2) for c in range(3))
        nrm = s0.mul(w) + s1.mul(u) + s2.mul(v)
        wi = -rd_j
        is_t = ((event >> j) & 1).astype(bool)
        inside = wi.dot(nrm).v < 0.0
        wo_in = refract(wi, -nrm, one / eta_d)
        wo_out = refract(wi, nrm, eta_d)
        wo_t = wo_in.where(inside, wo_out)
        dead |= is_t & inside & (wo_in.x.v == 0) & (wo_in.y.v == 0) & (wo_in.z.v == 0)
        wo = wo_t.where(is_t, reflect(wi, nrm))
        ro_j, rd_j = p, wo
    p_last = np.stack([ro_j.x.v, ro_j.y.v, ro_j.z.v], 1)
    nrm = V.const(p_last - pe, 2)
    t = (V.const(pe, 2) - ro_j).dot(nrm) / nrm.dot(rd_j)
    pf = ro_j + rd_j.mul(t)
    a0 = np.stack([pf.x.g[:, 0], pf.y.g[:, 0], pf.z.g[:, 0]], 1)
    a1 = np.stack([pf.x.g[:, 1], pf.y.g[:, 1], pf.z.g[:, 1]], 1)
    l0 = np.linalg.norm(a0, axis=1)
    l1 = np.linalg.norm(a1, axis=1)
    s = np.linalg.norm(np.cross(a0 / l0[:, None], a1 / l1[:, None]), axis=1)
    out = l0 * l1 * np.maximum(s, 0.01)
    return np.where(dead, 3.402823466e+38, out)


def contribution(x, verts, nrms, pb, pe, ng_end, eta, event, dielectric, K, intensity, clamp):
    """Irradiance of each item at the FP64 solution x (n, 2K)."""
    n = x.shape[0]
    xs = [D.const(x[:, i], 1) for i in range(2 * K)]
    vt = [[V.const(verts[k][:, c], 1) for c in range(3)] for k in range(K)]
    nt = [[V.const(nrms[k][:, c], 1) for c in range(3)] for k in range(K)]
    vert = [V.const(pb, 1)] + [_at(vt[k], xs[2 * k], xs[2 * k + 1]) for k in range(K)] + [V.const(pe, 1)]
    thr = np.ones(n)
    for k in range(K):
        s = _at(nt[k], xs[2 * k], xs[2 * k + 1])
        sv = np.stack([s.x.v, s.y.v, s.z.v], 1)
        sv = sv / np.linalg.norm(sv, axis=1, keepdims=True)
        wi = vert[k] - vert[k + 1]
        r = fresnel(wi, V.const(sv, 1), D.const(eta, 1)).v
        f = np.where(((event >> k) & 1).astype(bool), 1.0 - r, r)
        thr = np.where(dielectric[:, k], thr * f, thr)
    v1 = np.stack([vert[1].x.v, vert[1].y.v, vert[1].z.v], 1)
    vk = np.stack([vert[K].x.v, vert[K].y.v, vert[K].z.v], 1)
    area = footprint(pb, v1 - pb, pe, verts, nrms, eta, event, K)
    d = vk - pe
    cosr = np.einsum("ni,ni->n", d / np.linalg.norm(d, axis=1, keepdims=True), ng_end)
    c = thr * intensity / area * cosr
    c = np.where(cosr > 0.0, c, 0.0)
    return np.minimum(c, clamp)


def scene_geometry(heights, cells, extent=(-1.0, 1.0, -1.0, 1.0)):
    """FP64 vertex positions (m*m, 3) and normals (m*m, 3) of the heightfield:
    normals normalize(-dh/dx, 1, -dh/dz) by np.gradient (central inside)."""
    m = cells + 1
    xs = np.linspace(extent[0], extent[1], m)
    zs = np.linspace(extent[2], extent[3], m)
    X, Z = np.meshgrid(xs, zs, indexing="ij")
    Y = np.asarray(heights, dtype=np.float64).reshape(m, m)
    hx = np.gradient(Y, xs, axis=0)
    hz = np.gradient(Y, zs, axis=1)
    N = np.stack([-hx, np.ones_like(hx), -hz], -1)
    N /= np.linalg.norm(N, axis=-1, keepdims=True)
    return np.stack([X, Y, Z], -1).reshape(-1, 3), N.reshape(-1, 3)


class ItemSet:
    """The contributing items of a forward solve, with their geometry in FP64
    as functions of the heights."""

    def __init__(self, scene, paths, results, tri_vertex, cells, intensity, clamp):
        L = scene.lambdas.shape[0]
        contrib = results["status"].reshape(-1) == 2
        idx = np.nonzero(contrib)[0]
        self.item = idx
        self.pixel = paths["pixel"][idx // L].astype(np.int64)
        self.lam = (idx % L).astype(np.int64)
        self.event = paths["event"][idx // L].astype(np.int64)
        self.K = int(self.event.max() >> 24) if idx.size else 0
        assert np.all(self.event >> 24 == self.K), "one chain length per item set"
        self.tris = paths["tris"][idx // L][:, :self.K].astype(np.int64)
        self.x0 = results["params"].reshape(-1, 8)[idx][:, :2 * self.K].astype(np.float64)
        self.scene = scene
        self.tv = tri_vertex
        self.cells = cells
        self.n_points = scene.points.shape[0]
        lam = scene.lambdas.astype(np.float32)[self.lam]
        self.eta = (np.float32(scene.cauchy_a) + np.float32(scene.cauchy_b) / (lam * lam)).astype(np.float64)
        self.pb = np.tile(np.asarray(scene.p_light, dtype=np.float64), (idx.size, 1))
        self.pe = scene.points["p"][self.pixel].astype(np.float64)
        self.ng = scene.points["ng"][self.pixel].astype(np.float64)
        self.dielectric = scene.attribs["material"][self.tris] == 2  # PALZ_MATERIAL_DIELECTRIC
        self.intensity, self.clamp = intensity, clamp
        self.base_tris = scene.tris.reshape(-1, 3, 3).astype(np.float64)
        self.base_nrm = scene.attribs["shading_normals"].astype(np.float64)

    def keep(self, mask):
        """Restrict the set to the items where `mask` holds."""
        for name in ("item", "pixel", "lam", "event", "tris", "x0", "eta", "pb", "pe", "ng", "dielectric"):
            setattr(self, name, getattr(self, name)[mask])
        if hasattr(self, "_item_tv"):
            del self._item_tv

    def geometry(self, heights, sel):
        """FP64 triangle vertices and normals of items `sel` at `heights`."""
        P, N = scene_geometry(heights, self.cells)
        verts, nrms = [], []
        for k in range(self.K):
            t = self.tris[sel, k]
            tv = self.tv[t]                                   # (n, 3)
            hf = tv[:, 0] >= 0
            v = self.base_tris[t].copy()
            s = self.base_nrm[t].copy()
            v[hf] = P[tv[hf]]
            s[hf] = N[tv[hf]]
            verts.append(v)
            nrms.append(s)
        return verts, nrms

    def contributions(self, heights, sel=None):
        sel = np.arange(self.item.size) if sel is None else sel
        with np.errstate(divide="ignore", invalid="ignore"):
            return self._contributions(heights, sel)

    def _contributions(self, heights, sel):
        verts, nrms = self.geometry(heights, sel)
        x = solve(self.x0[sel], verts, nrms, self.pb[sel], self.pe[sel], self.eta[sel], self.event[sel], self.K)
        return contribution(x, verts, nrms, self.pb[sel], self.pe[sel], self.ng[sel], self.eta[sel],
                            self.event[sel], self.dielectric[sel], self.K, self.intensity, self.clamp)

    def contributions_many(self, heights_list, sel):
        """contributions(h, sel) for every h of heights_list in one vectorised
        solve: (len(heights_list), len(sel)) array."""
        with np.errstate(divide="ignore", invalid="ignore"):
            return self._contributions_many(heights_list, sel)

    # A chain that total internal reflection ends inside dA/dw is carried on
    # through zero directions (inf / nan) and then replaced by the reference's
    # FLT_MAX footprint (pk.h:844-930); those masked lanes are why the
    # floating-point warnings are silenced here.  A chain whose own value is
    # not finite is reported by loss_fd_ladder.
    def _contributions_many(self, heights_list, sel):
        m = len(heights_list)
        if sel.size == 0:
            return np.zeros((m, 0))
        geo = [self.geometry(h, sel) for h in heights_list]
        verts = [np.concatenate([g[0][k] for g in geo]) for k in range(self.K)]
        nrms = [np.concatenate([g[1][k] for g in geo]) for k in range(self.K)]

        def rep(a):
            return np.concatenate([a[sel]] * m)
        x = solve(rep(self.x0), verts, nrms, rep(self.pb), rep(self.pe), rep(self.eta), rep(self.event), self.K)
        c = contribution(x, verts, nrms, rep(self.pb), rep(self.pe), rep(self.ng), rep(self.eta), rep(self.event),
                         rep(self.dielectric), self.K, self.intensity, self.clamp)
        return c.reshape(m, sel.size)

    def image(self, heights):
        img = np.zeros(self.scene.lambdas.shape[0] * self.n_points)
        np.add.at(img, self.lam * self.n_points + self.pixel, self.contributions(heights))
        return img

    def depends(self, j, nb):
        """Items whose inputs depend on height j (their vertices' heights or the
        heights their normals use)."""
        users = [j] + [int(v) for v in np.nonzero(np.any(nb == j, axis=1))[0]]
        if not hasattr(self, "_item_tv"):
            # vertex id -> items whose chain triangles use that vertex (CSR)
            tv = np.concatenate([self.tv[self.tris[:, k]] for k in range(self.K)], axis=1)  # (n, 3K)
            items = np.repeat(np.arange(self.item.size), tv.shape[1])
            v = tv.reshape(-1)
            keep = v >= 0
            v, items = v[keep], items[keep]
            order = np.argsort(v, kind="stable")
            self._item_tv = (v[order], items[order])
        v_sorted, items_sorted = self._item_tv
        parts = [items_sorted[np.searchsorted(v_sorted, u):np.searchsorted(v_sorted, u, side="right")]
                 for u in users]
        return np.unique(np.concatenate(parts)) if parts else np.zeros(0, np.int64)

    def loss_fd(self, heights, target, j, nb, h=1e-4):
        """Central difference of L = sum (I - T)^2 in height j (FP64)."""
        base_img = self.image(heights)
        sel = self.depends(j, nb)
        c0 = self.contributions(heights, sel)
        out = []
        for s in (1.0, -1.0):
            hh = np.array(heights, dtype=np.float64)
            hh[j] += s * h
            img = base_img.copy()
            np.add.at(img, self.lam[sel] * self.n_points + self.pixel[sel], self.contributions(hh, sel) - c0)
            out.append(np.sum((img - target) ** 2))
        return (out[0] - out[1]) / (2.0 * h)


class Problem:
    """Items of every chain length of a forward solve (one ItemSet per K)."""

    def __init__(self, scene, paths, results, tri_vertex, normal_nb, cells, intensity, clamp):
        L = scene.lambdas.shape[0]
        K = paths["event"] >> 24
        self.sets = []
        for k in (1, 2, 3):
            m = K == k
            if not np.any(results["status"].reshape(-1, L)[m] == 2):
                continue
            self.sets.append(ItemSet(scene, paths[m], results.reshape(-1, L)[m].reshape(-1), tri_vertex, cells,
                                     intensity, clamp))
        self.nb = normal_nb
        self.size = scene.lambdas.shape[0] * scene.points.shape[0]

    def drop_nonfinite(self, heights):
        """Leave out the items whose FP64 contribution at `heights` is not
        finite (a square root of zero on their chain).  Returns how many, and
        the image of the items kept (self.image(heights) after the call)."""
        dropped = 0
        img = np.zeros(self.size)
        for s in self.sets:
            c = s.contributions(heights)
            finite = np.isfinite(c)
            dropped += int((~finite).sum())
            s.keep(finite)
            np.add.at(img, s.lam * s.n_points + s.pixel, c[finite])
        return dropped, img

    def image(self, heights):
        img = np.zeros(self.size)
        for s in self.sets:
            img += s.image(heights)
        return img

    def loss_fd_ladder(self, heights, target, j, steps, base):
        """Central differences of L = sum (I - T)^2 in height j for every step
        of `steps` (FP64; `base` = self.image(heights), or the image the
        gradient is taken around).  Returns (differences, finite): finite is
        False when a dependent item's FP64 chain is not finite at one of the
        stencil's heights (it reaches a square root of zero: grazing or
        total internal reflection within the step), in which case that
        height has no central difference at these steps."""
        sels = [s.depends(j, self.nb) for s in self.sets]
        hs = [np.array(heights, dtype=np.float64)]
        for h in steps:
            for sgn in (1.0, -1.0):
                hh = np.array(heights, dtype=np.float64)
                hh[j] += sgn * h
                hs.append(hh)
        cs = [s.contributions_many(hs, sel) for s, sel in zip(self.sets, sels)]
        finite = all(np.all(np.isfinite(c)) for c in cs)
        out = []
        for i in range(len(steps)):
            val = []
            for v in (1 + 2 * i, 2 + 2 * i):
                img = base.copy()
                for s, sel, c in zip(self.sets, sels, cs):
                    if sel.size:
                        np.add.at(img, s.lam[sel] * s.n_points + s.pixel[sel], c[v] - c[0])
                val.append(np.sum((img - target) ** 2))
            out.append((val[0] - val[1]) / (2.0 * steps[i]))
        return np.array(out), finite

    def loss_fd(self, heights, target, j, h=1e-4, base=None):
        """Central difference of L = sum (I - T)^2 in height j (FP64); `base`,
        if given, is self.image(heights), computed once for many j."""
        base = self.image(heights) if base is None else base
        out = []
        for sgn in (1.0, -1.0):
            hh = np.array(heights, dtype=np.float64)
            hh[j] += sgn * h
            img = base.copy()
            for s in self.sets:
                sel = s.depends(j, self.nb)
                if sel.size:
                    np.add.at(img, s.lam[sel] * s.n_points + s.pixel[sel],
                              s.contributions(hh, sel) - s.contributions(heights, sel))
            out.append(np.sum((img - target) ** 2))
        return (out[0] - out[1]) / (2.0 * h)
