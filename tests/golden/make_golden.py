"""Independent numpy brute-force restatement of the reference path, used to
write the golden fixtures under tests/golden/ that pin the CPU oracle.

It deliberately takes different routes from the oracle wherever the
mathematics allows, so a shared mistake is unlikely:
  * HEALPix NESTED -> (theta, phi) through nest2ring + the RING formulas
    (Gorski et al. 2005), with numpy arccos/sin/cos;
  * occlusion as a ray-sphere intersection |P + t s| <= R for some t > 0
    (per-antenna mode), or the reference's fixed cone cos(pi - asin(R/(R+h)))
    with numpy trigonometry (lunar mode);
  * orbit positions and visibilities with numpy sin/cos/exp;
  * the w-plane fit with numpy.linalg.lstsq, grouping with a stable argsort;
  * mt19937_64 as plain Python integers.
Reference semantics restated (file:line): orbit_online.cuh:39-231,317-333;
viss_recon_kernel.cuh:29-112, 200-287, 295-467, 470-701; common.hpp:255-277;
recon_mirror/viss_recon_kernel_mirror.cuh:462-904;
mirror_tile_vissfast/viss_recon_kernel_mirror_tile_vissfast.cuh:21-98, 867-875, 1039-1115;
main_recon_online.cu:83-139, 370-383, 497-512.

Tile cones only accelerate; the fixtures hold the pure per-point predicate
results, which equal the tile-cone form whenever the cone invariant holds
(checked separately).

Run:  python tests/golden/make_golden.py [case ...]   (writes tests/golden/*.npz)
"""
import math
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


# ---------------------------------------------------------------- mt19937_64
class MT64:
    def __init__(self, seed):
        self.mt = [0] * 312
        self.mt[0] = seed & 0xFFFFFFFFFFFFFFFF
        for i in range(1, 312):
            p = self.mt[i - 1]
            self.mt[i] = (6364136223846793005 * (p ^ (p >> 62)) + i) & 0xFFFFFFFFFFFFFFFF
        self.i = 312

    def __call__(self):
        M = 0xFFFFFFFFFFFFFFFF
        if self.i >= 312:
            for k in range(312):
                x = (self.mt[k] & 0xFFFFFFFF80000000) | (self.mt[(k + 1) % 312] & 0x7FFFFFFF)
                xa = x >> 1
                if x & 1:
                    xa ^= 0xB5026F5AA96619E9
                self.mt[k] = self.mt[(k + 156) % 312] ^ xa
            self.i = 0
        y = self.mt[self.i]
        self.i += 1
        y ^= (y >> 29) & 0x5555555555555555
        y ^= (y << 17) & 0x71D67FFFEDA60000 & M
        y ^= (y << 37) & 0xFFF7EEE000000000 & M
        y ^= y >> 43
        return y & M

    def u01(self):
        return (self() >> 11) * (1.0 / 9007199254740992.0)


# ---------------------------------------------------------------- HEALPix
def nest2ring(nside, ipix):
    ipix = np.asarray(ipix, np.int64)
    npface = nside * nside
    face = ipix // npface
    ipf = ipix % npface
    ix = np.zeros_like(ipix)
    iy = np.zeros_like(ipix)
    for b in range(30):
        ix |= ((ipf >> (2 * b)) & 1) << b
        iy |= ((ipf >> (2 * b + 1)) & 1) << b
    jrll = np.array([2, 2, 2, 2, 3, 3, 3, 3, 4, 4, 4, 4])[face]
    jpll = np.array([1, 3, 5, 7, 0, 2, 4, 6, 1, 3, 5, 7])[face]
    jr = jrll * nside - (ix + iy) - 1
    nl4 = 4 * nside
    ncap = 2 * nside * (nside - 1)
    npix = 12 * nside * nside
    ring = np.zeros_like(ipix)
    nr = np.where(jr < nside, jr, np.where(jr > 3 * nside, nl4 - jr, nside))
    kshift = np.where((jr >= nside) & (jr <= 3 * nside), (jr - nside) & 1, 0)
    jp = (jpll * nr + (ix - iy) + 1 + kshift) // 2
    jp = np.where(jp > nl4, jp - nl4, jp)
    jp = np.where(jp < 1, jp + nl4, jp)
    north = jr < nside
    south = jr > 3 * nside
    eq = ~north & ~south
    ring[north] = 2 * nr[north] * (nr[north] - 1) + jp[north] - 1
    ring[eq] = ncap + (jr[eq] - nside) * nl4 + jp[eq] - 1
    ring[south] = npix - 2 * nr[south] * (nr[south] + 1) + jp[south] - 1
    return ring


def ring2ang(nside, ip):
    ip = np.asarray(ip, np.int64)
    npix = 12 * nside * nside
    ncap = 2 * nside * (nside - 1)
    z = np.zeros(ip.shape)
    phi = np.zeros(ip.shape)
    for k, p in enumerate(ip.tolist()):
        if p < ncap:
            iring = (1 + math.isqrt(1 + 2 * p)) // 2
            iphi = p + 1 - 2 * iring * (iring - 1)
            z[k] = 1.0 - iring * iring / (3.0 * nside * nside)
            phi[k] = (iphi - 0.5) * math.pi / (2.0 * iring)
        elif p < npix - ncap:
            q = p - ncap
            iring = q // (4 * nside) + nside
            iphi = q % (4 * nside) + 1
            fodd = 1.0 if ((iring + nside) & 1) else 0.5
            z[k] = (2 * nside - iring) * 2.0 / (3.0 * nside)
            phi[k] = (iphi - fodd) * math.pi / (2.0 * nside)
        else:
            q = npix - p
            iring = (1 + math.isqrt(2 * q - 1)) // 2
            iphi = 4 * iring + 1 - (q - 2 * iring * (iring - 1))
            z[k] = -1.0 + iring * iring / (3.0 * nside * nside)
            phi[k] = (iphi - 0.5) * math.pi / (2.0 * iring)
    return np.arccos(z), phi


def lmn_nest(nside):
    npix = 12 * nside * nside
    th, ph = ring2ang(nside, nest2ring(nside, np.arange(npix)))
    st = np.sin(th)
    return st * np.cos(ph), -st * np.sin(ph), np.cos(th)


# ---------------------------------------------------------------- geometry
def circular_positions(seed, A, T, dt, rmin, rmax, gm):
    g = MT64(seed)
    pos = np.zeros((3, T, A))
    t = np.arange(T) * dt
    for a in range(A):
        r = rmin + (rmax - rmin) * g.u01()
        inc = math.radians(180.0 * g.u01())
        raan = math.radians(360.0 * g.u01())
        ph = math.radians(360.0 * g.u01())
        th = ph + math.sqrt(gm / r ** 3) * t
        pos[0, :, a] = r * (np.cos(th) * math.cos(raan) - np.sin(th) * math.cos(inc) * math.sin(raan))
        pos[1, :, a] = r * (np.cos(th) * math.sin(raan) + np.sin(th) * math.cos(inc) * math.cos(raan))
        pos[2, :, a] = r * np.sin(th) * math.sin(inc)
    return pos


def _round_away_f32(x):
    x = np.float32(x)
    return int(np.floor(x + np.float32(0.5))) if x >= 0 else int(np.ceil(x - np.float32(0.5)))


def lunar_positions(lam, day, seed, T_use):
    OR = math.ceil(2.0 * math.pi * (100e3 / lam))
    PC = _round_away_f32(np.float32(24.0) / np.float32(2.3))
    seg_len = OR // 3
    tw = _round_away_f32(np.float32(OR) * np.float32(7.0) * np.float32(24.0) / np.float32(2.3))
    mod14 = (day - 1) % 14
    doff = _round_away_f32(np.float32(mod14) * np.float32(OR) * np.float32(24.0) / np.float32(2.3))
    g = MT64(seed ^ ((day * 0x9E3779B97F4A7C15) & 0xFFFFFFFFFFFFFFFF))
    t1s = []
    for k in range(PC):
        os_, oe = k * OR, (k + 1) * OR - 1
        u = g.u01()
        t1 = os_ + _round_away_f32(np.float32(u * OR * (2.0 / 3.0)))
        if t1 + seg_len - 1 > oe:
            t1 = oe - seg_len + 1
        t1 = max(t1, os_)
        t1s.append(t1)

    def sat0(row):
        gg, k = divmod(row, OR)
        raan = math.radians(gg * 0.08 + (k + 1) / OR * 0.08 + (day - 1) * 0.8)
        M = 1.0 / OR + (2 * math.pi + math.pi / 200 - 1.0 / OR) * k / (OR - 1)
        i = math.radians(30.0)
        A = 2038.14
        return np.array([A * (math.cos(M) * math.cos(raan) - math.sin(M) * math.cos(i) * math.sin(raan)),
                         A * (math.cos(M) * math.sin(raan) + math.sin(M) * math.cos(i) * math.cos(raan)),
                         A * math.sin(M) * math.sin(i)])

    u0 = 100e3 / 23.0
    r1 = [0, u0, 4 * u0, 10 * u0, 16 * u0, 18 * u0, 21 * u0, 23 * u0]
    r2 = [0, 100.0, 0.4 * u0, u0, 1.6 * u0, 1.8 * u0, 2.1 * u0, 2.3 * u0]
    pos = np.zeros((3, T_use, 8))
    for t in range(T_use):
        row = t1s[t // seg_len] + t % seg_len
        p0 = sat0(row)
        pos[:, t, 0] = p0
        if row == 0:
            continue
        d = p0 - sat0(row - 1)
        mv = np.linalg.norm(d) + 1e-20
        p = (doff + row - 1) % (2 * tw)
        for s in range(1, 8):
            rt = r2[s] + (r1[s] - r2[s]) * p / (tw - 1) if p < tw else r1[s] + (r2[s] - r1[s]) * (p - tw) / (tw - 1)
            pos[:, t, s] = p0 - rt * d / mv / 1000.0
    return pos, OR, seg_len, PC * seg_len


# ---------------------------------------------------------------- pipeline
def turns(u, l, v, m, w=None, nm1=None):
    """u l + v m (+ w nm1) as outer products, reduced mod 1 in extended
    precision before the 2 pi scaling (|u| reaches 1e6 turns)."""
    L = np.longdouble
    x = np.outer(u.astype(L), l.astype(L)) + np.outer(v.astype(L), m.astype(L))
    if w is not None:
        x = x + np.outer(w.astype(L), nm1.astype(L))
    return (x - np.round(x)).astype(np.float64)


def pairs(A):
    return [(n, m) for m in range(A - 1) for n in range(m + 1, A)]  # (a = n, b = m)


def visible(s, P, R, h):
    """s: (npts, 3) directions, P: (3,) antenna.  True where visible."""
    if h is None:  # per-antenna sphere of radius R: ray P + t s misses for t > 0
        b = s @ P
        c = P @ P - R * R
        disc = b * b - c
        return ~((b < 0) & (disc >= 0))
    cosphi = math.cos(math.pi - math.asin(R / (R + h)))
    return (s @ P) / np.linalg.norm(P) >= cosphi


def run(cfg):
    nside, A = cfg["nside"], cfg["n_ant"]
    lam = 3e8 / cfg["freq_hz"]
    npix = 12 * nside * nside
    l, m, n = lmn_nest(nside)
    s = np.stack([l, m, n], 1)
    B = np.random.default_rng(cfg["sky_seed"]).lognormal(0.0, 1.0, npix) * (4 * math.pi / npix)
    segs = cfg["seg_tlen"]
    T_use = sum(segs)
    if cfg["geometry"] == "circular":
        pos = circular_positions(cfg["orbit_seed"], A, T_use, 10.0, 6900.0, 7100.0, 398600.4418)
        R, h = 6371.0, None
        # automatic bl_max: longest 3-D baseline (m) over the run
        blmax = max(np.linalg.norm(pos[:, t, a] - pos[:, t, b]) for t in range(T_use) for a, b in pairs(A)) * 1000.0
    else:
        pos, _, _, _ = lunar_positions(lam, cfg["day"], cfg["orbit_seed"], T_use)
        R, h = 1737.1, 300.0
        blmax = 100e3
    P = len(pairs(A))
    img_n = cfg["img_n"]
    jj, ii = np.meshgrid(np.arange(img_n), np.arange(img_n))
    gl = (2 * jj + 1 - img_n) / img_n
    gm = (2 * ii + 1 - img_n) / img_n
    ind = gl ** 2 + gm ** 2 < 1
    gn = np.sqrt(np.clip(1 - gl ** 2 - gm ** 2, 0, None))
    gs = np.stack([gl[ind], gm[ind], gn[ind]], 1)
    out = {"lmn": np.stack([l, m, n]), "pos": pos, "blmax": blmax}
    Cacc = np.zeros(npix)
    Wacc = np.zeros(npix, np.int64)
    Gacc = np.zeros(gs.shape[0])
    GW = np.zeros(gs.shape[0], np.int64)
    Wtot = 0
    t0 = 0
    full = cfg["recon_mode"] == 1
    for k, tl in enumerate(segs):
        rows = []
        for g in range(tl):
            for a, b in pairs(A):
                Pa, Pb = pos[:, t0 + g, a], pos[:, t0 + g, b]
                uvw = -(Pa - Pb) * 1e3 / lam
                rows.append((uvw, Pa, Pb))
        u = np.array([r[0][0] for r in rows])
        v = np.array([r[0][1] for r in rows])
        w = np.array([r[0][2] for r in rows])
        vis = np.array([visible(s, r[1], R, h) & visible(s, r[2], R, h) for r in rows])
        vraw = (vis * B[None, :] * np.exp(-2j * math.pi * turns(u, l, v, m, w, n - 1))).sum(1)
        vh = vraw * np.exp(-2j * math.pi * (w - np.round(w)))
        nh = len(rows)
        vfull = np.zeros(2 * nh, complex)
        for hh in range(nh):
            g, j = divmod(hh, P)
            vfull[2 * P * g + j] = vh[hh]
            vfull[2 * P * g + P + j] = np.conj(vh[hh])
        # w-plane fit over the full rows (both signs)
        U = np.concatenate([u, -u])
        V = np.concatenate([v, -v])
        Wf = np.concatenate([w, -w])
        (fa, fb), *_ = np.linalg.lstsq(np.stack([U, V], 1), Wf, rcond=None)
        lmmax = max(math.sqrt(1 + fa * fa), math.sqrt(1 + fb * fb))
        temp = math.ceil(2 * blmax / lam * 2 * lmmax)
        res = int(temp + 3 - (temp % 2))
        half = (res - 1) // 2
        du = 2 * blmax / lam / (res - 1)
        if full:
            ru = np.empty(2 * nh); rv = np.empty(2 * nh); rV = np.empty(2 * nh, complex)
            hidx = np.empty(2 * nh, np.int64)
            for i in range(2 * nh):
                g, j2 = divmod(i, 2 * P)
                neg = j2 >= P
                hh = g * P + (j2 - P if neg else j2)
                ru[i] = -u[hh] if neg else u[hh]
                rv[i] = -v[hh] if neg else v[hh]
                rV[i] = vfull[i]
                hidx[i] = hh
            rnd = lambda x: np.where(x >= 0, np.floor(x + 0.5), np.ceil(x - 0.5))
            Uc = rnd(ru / du) + half
            Vc = rnd(rv / du) + half
        else:
            ru, rv, rV, hidx = u, v, vh, np.arange(nh)
            Uc = np.floor(ru / du + 0.5) + half
            Vc = np.floor(rv / du + 0.5) + half
        inside = (Uc >= 0) & (Uc < res) & (Vc >= 0) & (Vc < res)
        key = np.where(inside, Uc.astype(np.int64) * res + Vc.astype(np.int64) + 1, 0)
        order = np.argsort(key, kind="stable")
        uk, start, cnt = np.unique(key[order], return_index=True, return_counts=True)
        offs = np.append(start, len(key))
        seg = {"vraw": vraw, "vh": vh, "vfull": vfull, "u": u, "v": v, "w": w, "vis": vis,
               "fa": fa, "fb": fb, "res": res, "du": du, "keys": uk, "offsets": offs, "idx_sorted": order}
        out[f"seg{k}"] = seg

        def recon(points):
            Cs = np.zeros(len(points))
            Ws = np.zeros(len(points), np.int64)
            lp = points[:, 0] + fa * points[:, 2]
            mp = points[:, 1] + fb * points[:, 2]
            pvis = np.array([visible(points, rows[hh][1], R, h) & visible(points, rows[hh][2], R, h)
                             for hh in range(nh)])
            for q, kk in enumerate(uk):
                if kk == 0:
                    continue
                Ucell, Vcell = divmod(kk - 1, res)
                ej = np.exp(2j * math.pi * turns(np.array([(Ucell - half) * du]), lp,
                                                 np.array([(Vcell - half) * du]), mp)[0])
                mem = order[offs[q]:offs[q + 1]]
                L = len(mem)
                if not cfg["blockage"]:
                    Cs += np.real(rV[mem].mean() * ej)
                    continue
                occ = cfg["occ_mode"]
                if occ == 0:
                    reps = list(range(L))
                elif occ == 1:
                    reps = [0]
                elif occ == 2:
                    reps = [L // 2]
                else:
                    reps = [0, L // 3, (2 * L) // 3, L - 1]
                acc = np.zeros(len(points), complex)
                cn = np.zeros(len(points), np.int64)
                for r in reps:
                    row = mem[r]
                    mk = pvis[hidx[row]]
                    acc += np.where(mk, rV[row], 0)
                    cn += mk
                ok = cn > 0
                Cs[ok] += np.real(acc[ok] / cn[ok] * ej[ok])
                Ws += ok
            return Cs, Ws

        cs, ws = recon(s)
        Cacc += cs
        Wacc += ws
        cg, wg = recon(gs)
        Gacc += cg
        GW += wg
        Wtot += len(uk)
        t0 += tl
    if cfg["blockage"]:
        Cm = np.where(Wacc > 0, Cacc / np.maximum(Wacc, 1), 0)
        Gm = np.where(GW > 0, Gacc / np.maximum(GW, 1), 0)
    else:
        Cm = Cacc / Wtot
        Gm = Gacc / Wtot
    img = np.zeros((img_n, img_n))
    Wimg = np.zeros((img_n, img_n), np.int64)
    img[ind] = Gm
    Wimg[ind] = GW
    out.update(C=Cm, W=Wacc, img=img, Wimg=Wimg)
    return out


CASES = {
    "circ_mirror_occ2": dict(geometry="circular", nside=8, n_ant=8, freq_hz=1e7, orbit_seed=42, seg_tlen=[8, 8],
                             img_n=32, blockage=1, occ_mode=2, recon_mode=0, sky_seed=11, day=1),
    "circ_full_occ3": dict(geometry="circular", nside=8, n_ant=6, freq_hz=3e7, orbit_seed=7, seg_tlen=[6, 5],
                           img_n=24, blockage=1, occ_mode=3, recon_mode=1, sky_seed=12, day=1),
    "circ_mirror_blk0": dict(geometry="circular", nside=4, n_ant=5, freq_hz=1e7, orbit_seed=3, seg_tlen=[10],
                             img_n=16, blockage=0, occ_mode=2, recon_mode=0, sky_seed=13, day=1),
    "lunar_mirror_occ0": dict(geometry="lunar8", nside=4, n_ant=8, freq_hz=1e6, orbit_seed=42, seg_tlen=[12, 12],
                              img_n=16, blockage=1, occ_mode=0, recon_mode=0, sky_seed=14, day=3),
    # the CPU-sized workload in full: Nside 32 (48 tiles of 256 pixels, so the
    # tile-cone classes are a mix of visible, hidden and mixed), 8 antennas,
    # 64 epochs in 4 segments of 16, 10 MHz, 256^2 image
    "circ_c1_mirror_occ2": dict(geometry="circular", nside=32, n_ant=8, freq_hz=1e7, orbit_seed=42,
                                seg_tlen=[16, 16, 16, 16], img_n=256, blockage=1, occ_mode=2, recon_mode=0,
                                sky_seed=15, day=1, compact=True),
    # first member of the sorted run (viss_recon_kernel.cuh:603-604)
    "circ_mirror_occ1": dict(geometry="circular", nside=8, n_ant=7, freq_hz=2e7, orbit_seed=5, seg_tlen=[7, 9],
                             img_n=24, blockage=1, occ_mode=1, recon_mode=0, sky_seed=16, day=1),
    # median member on the full uv grid (main_recon_online.cu form)
    "circ_full_occ2": dict(geometry="circular", nside=8, n_ant=6, freq_hz=1e7, orbit_seed=9, seg_tlen=[9, 6],
                           img_n=24, blockage=1, occ_mode=2, recon_mode=1, sky_seed=17, day=1),
    # any other occ_mode value takes the four-sample rule (the else branch,
    # viss_recon_kernel.cuh:605-612)
    "circ_mirror_occ4": dict(geometry="circular", nside=8, n_ant=6, freq_hz=1e7, orbit_seed=21, seg_tlen=[10],
                             img_n=16, blockage=1, occ_mode=4, recon_mode=0, sky_seed=18, day=1),
}


def compact(flat):
    """large cases keep the per-row occlusion masks as a SHA-256 of their
    packed bits (the test compares the oracle's predicate by the same
    digest), the image inside the unit disk only and every 8th pixel
    direction"""
    import hashlib

    out = {}
    for k, v in flat.items():
        if k == "lmn":
            out["lmn_stride8"] = v[:, ::8]
        elif k.endswith("/vis"):
            out[k + "_sha256"] = np.frombuffer(hashlib.sha256(np.packbits(v.astype(bool))).digest(), np.uint8)
        elif k in ("img", "Wimg"):
            n = v.shape[0]
            j = np.arange(n)
            disk = ((2 * j + 1 - n)[None, :] ** 2 + (2 * j + 1 - n)[:, None] ** 2) < n * n
            out[k + "_disk"] = v[disk]
            out["img_n"] = np.asarray(n)
        else:
            out[k] = v
    return out


def flatten(res):
    flat = {}
    for k, v in res.items():
        if isinstance(v, dict):
            for k2, v2 in v.items():
                flat[f"{k}/{k2}"] = np.asarray(v2)
        else:
            flat[k] = np.asarray(v)
    return flat


if __name__ == "__main__":
    import sys

    names = sys.argv[1:] or list(CASES)
    for name in names:
        cfg = CASES[name]
        flat = flatten(run(cfg))
        if cfg.get("compact"):
            flat = compact(flat)
        np.savez_compressed(os.path.join(HERE, f"{name}.npz"), **flat)
        print("wrote", name)
