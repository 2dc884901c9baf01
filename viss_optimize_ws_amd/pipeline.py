"""Host-side mirror of the reference's stage flow (main_recon_online.cu:274-512)
over the C-ABI of libvissws.so.

The reference is a CUDA program driven from ``main``; its per-segment flow is
    orbit_gen_make_segment -> viss_partial_all_halfsym[_tilecone] ->
    phase_correct + expand (conj fill) -> build_groups + recon -> normalise.
``Pipeline`` exposes exactly those stages.  Nothing here computes: every call
goes to the gfx950 library, and a missing library raises.
"""
import ctypes as C
import os

import numpy as np

from . import _abi as A

_LIB = None


class VissError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"[{code}] {msg}")
        self.code = code


def lib_path():
    return os.path.join(os.path.dirname(os.path.abspath(__file__)), "libvissws.so")


def load_library():
    """Load libvissws.so (built in-tree).  Raises if it is missing."""
    global _LIB
    if _LIB is None:
        p = lib_path()
        if not os.path.exists(p):
            raise RuntimeError(f"{p} is not built; run __graft_entry__.build() (hipcc, gfx950)")
        lib = C.CDLL(p)
        A.bind(lib, "vw_", A.STAGE_FUNCS)
        A.bind(lib, "vw_", A.DEVICE_ONLY)
        A.bind(lib, "vw_", A.UNET_FUNCS)
        _LIB = lib
    return _LIB


def default_config(geometry=A.VW_GEOM_CIRCULAR, **kw):
    """vw_config with the defaults of vw_config_init, then keyword overrides.

    Lunar defaults follow main_recon_online.cu:6-45 (10 MHz band, 10 segments,
    blockage=1, occ_mode=2, orbit_seed=42); circular defaults are the
    Earth-orbit workload (8 antennas, 64 epochs 10 s apart, 4 segments,
    R = 6371 km, radius in [6900, 7100] km, 256^2 image).
    """
    c = A.vw_config()
    c.geometry = geometry
    c.orbit_seed = 42
    c.day = 1
    c.blockage = 1
    c.occ_mode = 2
    c.recon_mode = A.VW_RECON_MIRROR
    c.tile_pix = 256
    c.precision = A.VW_PREC_PLAIN
    c.healpix_out = 1
    c.nsky = 1
    c.gm_km3_s2 = 398600.4418
    if geometry == A.VW_GEOM_LUNAR8:
        c.nside, c.freq_hz, c.n_ant, c.segs = 4096, 1e7, 8, 10
        c.occ_radius_km, c.occ_altitude_km = 1737.1, 300.0
        c.r_min_km = c.r_max_km = 2038.14
    else:
        c.nside, c.freq_hz, c.n_ant, c.n_epochs, c.dt_s, c.segs = 32, 1e7, 8, 64, 10.0, 4
        c.occ_radius_km, c.occ_altitude_km = 6371.0, -1.0
        c.r_min_km, c.r_max_km, c.img_n = 6900.0, 7100.0, 256
    seg_tlen = kw.pop("seg_tlen", None)
    for k, v in kw.items():
        if not hasattr(c, k):
            raise KeyError(k)
        setattr(c, k, v)
    if seg_tlen is not None:
        arr = (C.c_int32 * len(seg_tlen))(*[int(x) for x in seg_tlen])
        c.seg_tlen = C.cast(arr, C.POINTER(C.c_int32))
        c.n_seg_explicit = len(seg_tlen)
        c._seg_keepalive = arr
    return c


def _ptr(a, ctype=C.c_double):
    return a.ctypes.data_as(C.POINTER(ctype)) if a is not None else None


class Stages:
    """Stage driver over a library exporting the vissws stage ABI."""

    def __init__(self, lib, prefix, cfg):
        self._lib, self._p, self.cfg = lib, prefix, cfg
        self._ctx = C.c_void_p()
        rc = self._f("create")(C.byref(cfg), C.byref(self._ctx))
        if rc:
            raise VissError(rc, self._f("last_error")(None).decode())
        self.plan = A.vw_plan()
        self._check(self._f("get_plan")(self._ctx, C.byref(self.plan)))
        self.nh = 0
        # batched contexts stack per-sky outputs along a leading axis
        self.nsky = max(1, int(getattr(cfg, "nsky", 1)))

    def _stack(self, n, dtype=np.float64):
        return np.zeros((self.nsky, n) if self.nsky > 1 else n, dtype)

    def _f(self, name):
        return getattr(self._lib, self._p + name)

    def _check(self, rc):
        if rc:
            raise VissError(rc, self._f("last_error")(self._ctx).decode())

    def close(self):
        if self._ctx:
            self._f("destroy")(self._ctx)
            self._ctx = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- plan ----
    def schedule(self):
        n = self.plan.segs
        t0 = np.zeros(n, np.int32)
        tl = np.zeros(n, np.int32)
        self._check(self._f("get_schedule")(self._ctx, _ptr(t0, C.c_int32), _ptr(tl, C.c_int32)))
        return list(zip(t0.tolist(), tl.tolist()))

    # ---- stages ----
    def set_sky(self, B):
        """B: [npix] for one sky, [nsky][npix] for a batched context."""
        B = np.ascontiguousarray(B, dtype=np.float64)
        nsky = B.shape[0] if B.ndim == 2 else 1
        self._check(self._f("set_sky")(self._ctx, _ptr(B), B.shape[-1], nsky))

    def gen_segment(self, t0, tlen, geometry=False):
        self.nh = int(tlen) * self.plan.n_pairs
        if not geometry:
            self._check(self._f("gen_segment")(self._ctx, int(t0), int(tlen), None))
            return None
        nf = 2 * self.nh
        out = {k: np.zeros(nf) for k in ("u", "v", "w", "xa", "ya", "za", "xb", "yb", "zb")}
        g = A.vw_geom_host(*[_ptr(out[k]) for k in ("u", "v", "w", "xa", "ya", "za", "xb", "yb", "zb")])
        self._check(self._f("gen_segment")(self._ctx, int(t0), int(tlen), C.byref(g)))
        return out

    def conj_fill(self):
        vh = self._stack(self.nh, np.complex128)
        vf = self._stack(2 * self.nh, np.complex128)
        self._check(self._f("conj_fill")(self._ctx, _ptr(vh.view(np.float64)), _ptr(vf.view(np.float64))))
        return vh, vf

    def recon_segment(self):
        info = A.vw_seg_info()
        self._check(self._f("recon_segment")(self._ctx, C.byref(info)))
        return _info_dict(info)

    def finalize(self):
        npix = self.plan.npix
        Cm = self._stack(npix)
        per_sky = getattr(self, "per_sky_weights", False)
        W = np.zeros((self.nsky, npix) if per_sky else npix, np.uint32)
        nimg = self.cfg.img_n
        img = None
        if nimg:
            img = np.zeros((self.nsky, nimg, nimg) if self.nsky > 1 else (nimg, nimg))
        Wi = (np.zeros((self.nsky, nimg, nimg) if per_sky else (nimg, nimg), np.uint32)) if nimg else None
        self._check(self._f("finalize")(self._ctx, _ptr(Cm), _ptr(W, C.c_uint32), _ptr(img), _ptr(Wi, C.c_uint32)))
        return Cm, W, img, Wi

    def reset_accumulators(self):
        self._check(self._f("reset_accumulators")(self._ctx))

    # ---- introspection ----
    def lmn(self):
        n = self.plan.npix
        out = [np.zeros(n) for _ in range(4)]
        self._check(self._f("get_lmn")(self._ctx, *[_ptr(o) for o in out]))
        return out

    def tiles(self, img=False):
        nt = self.plan.n_img_tiles if img else self.plan.ntile
        out = np.zeros((5, nt))
        self._check(self._f("get_tiles")(self._ctx, int(img), _ptr(out)))
        return out

    def img_points(self):
        n = self.plan.n_img_pts
        idx = np.zeros(n, np.int64)
        l, m, nn = np.zeros(n), np.zeros(n), np.zeros(n)
        self._check(self._f("get_img_points")(self._ctx, _ptr(idx, C.c_int64), _ptr(l), _ptr(m), _ptr(nn)))
        return idx, l, m, nn

    def positions(self):
        out = np.zeros((3, self.plan.n_epochs, self.cfg.n_ant))
        self._check(self._f("get_positions")(self._ctx, _ptr(out)))
        return out

    def rows(self):
        u, v, w = np.zeros(self.nh), np.zeros(self.nh), np.zeros(self.nh)
        self._check(self._f("get_rows")(self._ctx, _ptr(u), _ptr(v), _ptr(w)))
        return u, v, w

    def tile_classes(self, img=False):
        nt = self.plan.n_img_tiles if img else self.plan.ntile
        out = np.zeros((self.nh, nt), np.uint8)
        self._check(self._f("get_tile_classes")(self._ctx, int(img), _ptr(out, C.c_uint8)))
        return out

    def masks(self, row0, nrows):
        out = np.zeros((nrows, self.plan.npix), np.uint8)
        self._check(self._f("get_masks")(self._ctx, int(row0), int(nrows), _ptr(out, C.c_uint8)))
        return out

    def groups(self):
        info = A.vw_seg_info()
        self._check(self._f("get_groups")(self._ctx, C.byref(info), None, None, None, None))
        nu = info.nuniq
        nr = self.nh * (2 if self.cfg.recon_mode == A.VW_RECON_FULL else 1)
        keys = np.zeros(nu, np.int64)
        off = np.zeros(nu + 1, np.int32)
        srt = np.zeros(nr, np.int32)
        rep = np.zeros(nu, np.int32)
        self._check(self._f("get_groups")(self._ctx, C.byref(info), _ptr(keys, C.c_int64), _ptr(off, C.c_int32),
                                          _ptr(srt, C.c_int32), _ptr(rep, C.c_int32)))
        return _info_dict(info), keys, off, srt, rep


def _info_dict(info):
    return {k: getattr(info, k) for k, _ in info._fields_}


class Pipeline(Stages):
    """The gfx950 path (libvissws.so)."""

    def __init__(self, cfg):
        super().__init__(load_library(), "vw_", cfg)

    def viss_half(self):
        out = self._stack(self.nh, np.complex128)
        self._check(self._lib.vw_viss_half(self._ctx, _ptr(out.view(np.float64))))
        return out

    def viss_adjoint(self, res):
        """adjoint of the current segment's half-visibility sum applied to
        residual rows res [n_half] (raw convention): [npix] pixel values"""
        res = np.ascontiguousarray(res, np.complex128)
        g = np.zeros(self.plan.npix)
        self._check(self._lib.vw_viss_adjoint(self._ctx, _ptr(res.view(np.float64)), _ptr(g)))
        return g

    def viss_counts(self):
        out = np.zeros(self.nh, np.int64)
        self._check(self._lib.vw_get_viss_counts(self._ctx, _ptr(out, C.c_int64)))
        return out

    def run_all(self):
        self._check(self._lib.vw_run_all(self._ctx))

    def run_all_graph(self):
        """one batched step (cleared accumulators, every segment) from a
        captured HIP graph"""
        self._check(self._lib.vw_run_all_graph(self._ctx))

    def set_day(self, day):
        self._check(self._lib.vw_set_day(self._ctx, int(day)))

    def accumulators(self):
        """Raw per-day accumulators: (C, W, G, GW, wtotal) before normalisation."""
        C_ = self._stack(self.plan.npix)
        W = np.empty(self.plan.npix, np.uint32)
        n = max(self.plan.n_img_pts, 1)
        G = self._stack(n)
        GW = np.empty(n, np.uint32)
        wt = C.c_int64(0)
        self._check(self._lib.vw_get_accumulators(self._ctx, _ptr(C_, C.c_double), _ptr(W, C.c_uint32),
                                                  _ptr(G, C.c_double), _ptr(GW, C.c_uint32), C.byref(wt)))
        k = self.plan.n_img_pts
        return C_, W, G[..., :k], GW[:k], wt.value

    def set_accumulators(self, C_, W, G, GW, wtotal):
        keep, ptrs = [], []
        for a, dt, ct in ((C_, np.float64, C.c_double), (W, np.uint32, C.c_uint32), (G, np.float64, C.c_double),
                          (GW, np.uint32, C.c_uint32)):
            if a is None:
                ptrs.append(None)
            else:
                keep.append(np.ascontiguousarray(a, dt))
                ptrs.append(_ptr(keep[-1], ct))
        self._check(self._lib.vw_set_accumulators(self._ctx, *ptrs, int(wtotal)))

    def set_batch_targets(self, viss_waves, recon_waves):
        self._check(self._lib.vw_set_batch_targets(self._ctx, int(viss_waves), int(recon_waves)))

    def batch_targets(self):
        v, r = C.c_int64(0), C.c_int64(0)
        self._check(self._lib.vw_get_batch_targets(self._ctx, C.byref(v), C.byref(r)))
        return v.value, r.value

    def autotune(self, reps=3):
        """Times every launch target of the batched kernels and keeps the
        fastest; clears the accumulators.  Returns (viss, recon) targets."""
        v, r = C.c_int64(0), C.c_int64(0)
        self._check(self._lib.vw_autotune(self._ctx, int(reps), C.byref(v), C.byref(r)))
        return v.value, r.value

    def set_sky_schedules(self, schedules):
        """per-sky segment lengths (a list of epoch lists, one per sky) for a
        batched context; None returns to the shared schedule.  vw_finalize
        then returns one weight map per sky."""
        if schedules is None:
            self._check(self._lib.vw_set_sky_schedules(self._ctx, None, None))
            self.per_sky_weights = False
            return
        nseg = np.array([len(x) for x in schedules], np.int32)
        tl = np.array([t for x in schedules for t in x], np.int32)
        self._check(self._lib.vw_set_sky_schedules(self._ctx, _ptr(nseg, C.c_int32), _ptr(tl, C.c_int32)))
        self.per_sky_weights = True

    def timings(self):
        t = A.vw_timings()
        self._check(self._lib.vw_get_timings(self._ctx, C.byref(t)))
        return {k: getattr(t, k) for k, _ in t._fields_}


def make_sky(seed, npix, kind=0):
    """Seeded synthetic sky (host utility of the library): kind 0 log-normal(0,1);
    kind 1 adds Poisson(100) (cap 200) Pareto(1.5) point sources."""
    lib = load_library()
    out = np.zeros(int(npix))
    rc = lib.vw_make_sky(int(seed), int(npix), int(kind), _ptr(out))
    if rc:
        raise VissError(rc, "vw_make_sky")
    return out
