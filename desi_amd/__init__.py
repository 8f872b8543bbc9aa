"""desi_amd -- MI355X (gfx950) pair / triplet histogram engine for 3PCFGPU.

Python mirror of the C ABI in include/pcf_abi.h (libpcf_hip.so), used by the
tests, the benchmark and the smoke check.  The compute path is the HIP library;
this module only marshals host arrays through ctypes and raises when the
native library is missing or a call fails.  There is no CPU fallback.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libpcf_hip.so")
CLI_PATH = os.path.join(HERE, "bin", "PCF")

# Histogram types (argv[1] of the reference, cuda/main.cu:59).
PCF_2ISO, PCF_2ANI, PCF_3ISO, PCF_3ANI = 0, 1, 2, 3
TYPES = {"2iso": PCF_2ISO, "2ani": PCF_2ANI, "3iso": PCF_3ISO, "3ani": PCF_3ANI}
# Which histogram.
PCF_XX, PCF_XY, PCF_XXX, PCF_XXY = 0, 1, 2, 3
# Arithmetic.
PCF_PREC_MIXED, PCF_PREC_FP64 = 0, 1
PCF_PAIR_AUTO, PCF_PAIR_DIRECT, PCF_PAIR_GRAM = 0, 1, 2
PCF_EST_NORMALIZE = 1
# Status codes (include/pcf_abi.h).
PCF_OK, PCF_ERR_ARG, PCF_ERR_HIP, PCF_ERR_NOMEM = 0, -1, -2, -3
PCF_ERR_RANGE, PCF_ERR_UNSUPPORTED, PCF_ERR_BUSY = -4, -5, -6

ERRORS = {0: "ok", -1: "invalid argument", -2: "HIP runtime error", -3: "out of device memory",
          -4: "coordinates out of range", -5: "unsupported", -6: "busy"}

EXPORTED_SYMBOLS = (
    "pcf_hist_size", "pcf_create", "pcf_destroy", "pcf_last_error", "pcf_create_error",
    "pcf_catalog_upload", "pcf_catalog_upload_device", "pcf_catalog_release",
    "pcf_catalog_size", "pcf_histogram", "pcf_last_stats", "pcf_analytic_rr",
    "pcf_analytic_3iso", "pcf_histogram_batch", "pcf_autotune", "pcf_set_max_batch",
    "pcf_set_hbm_budget", "pcf_submit", "pcf_poll", "pcf_wait", "pcf_estimator",
    "pcf_histogram_noalloc", "pcf_selftest_scan", "pcf_tune_save", "pcf_tune_load",
    "pcf_tune_lookup", "pcf_set_ref_bugs",
)
# Bytes of the static device array pcf_histogram_noalloc works in (per GPU).
PCF_NOALLOC_SCRATCH_BYTES = 128 * 1024 * 1024


class PcfJob(ctypes.Structure):
    _fields_ = [("type", ctypes.c_int32), ("hist", ctypes.c_int32), ("bpc", ctypes.c_int32),
                ("bins", ctypes.c_int32), ("dmax", ctypes.c_float), ("box", ctypes.c_float),
                ("precision", ctypes.c_int32), ("cat_x", ctypes.c_int32),
                ("cat_y", ctypes.c_int32), ("partitions", ctypes.c_int32),
                ("cells_per_rmax", ctypes.c_int32), ("rebuild_grid", ctypes.c_int32),
                ("lds_copies_log2", ctypes.c_int32), ("blocks_per_cu", ctypes.c_int32),
                ("pair_kernel", ctypes.c_int32), ("reserved", ctypes.c_int32)]


class PcfTuning(ctypes.Structure):
    _fields_ = [("cells_per_rmax", ctypes.c_int32), ("lds_copies_log2", ctypes.c_int32),
                ("blocks_per_cu", ctypes.c_int32), ("trials", ctypes.c_int32),
                ("best_ms", ctypes.c_double), ("default_ms", ctypes.c_double),
                ("pair_kernel", ctypes.c_int32), ("reserved", ctypes.c_int32)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class PcfStats(ctypes.Structure):
    _fields_ = [("evaluations", ctypes.c_uint64), ("rechecks", ctypes.c_uint64),
                ("work_items", ctypes.c_uint64), ("kernel_ms", ctypes.c_double),
                ("total_ms", ctypes.c_double), ("cells_per_rmax", ctypes.c_int32),
                ("graph_replay", ctypes.c_int32), ("chunks", ctypes.c_int32),
                ("pairs_listed", ctypes.c_int32), ("batched_jobs", ctypes.c_int32),
                ("direct_items", ctypes.c_int32), ("queue_ms", ctypes.c_double),
                ("latency_ms", ctypes.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class PcfError(RuntimeError):
    def __init__(self, rc, msg):
        super().__init__(f"{ERRORS.get(rc, rc)}: {msg}")
        self.rc = rc


_lib = None


def load_library(path=LIB_PATH):
    """Load libpcf_hip.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is missing: build it with `make -C desi_amd` "
                           "(or __graft_entry__.build())")
    L = ctypes.CDLL(path)
    c_i32, c_i64, vp = ctypes.c_int32, ctypes.c_int64, ctypes.c_void_p
    L.pcf_hist_size.argtypes = [c_i32, c_i32]
    L.pcf_hist_size.restype = c_i64
    L.pcf_create.argtypes = [ctypes.POINTER(vp), c_i32]
    L.pcf_destroy.argtypes = [vp]
    L.pcf_destroy.restype = None
    L.pcf_last_error.argtypes = [vp]
    L.pcf_last_error.restype = ctypes.c_char_p
    L.pcf_create_error.restype = ctypes.c_char_p
    L.pcf_catalog_upload.argtypes = [vp, vp, c_i64, ctypes.POINTER(c_i32)]
    L.pcf_catalog_upload_device.argtypes = [vp, vp, c_i64, ctypes.POINTER(c_i32)]
    L.pcf_catalog_release.argtypes = [vp, c_i32]
    L.pcf_catalog_size.argtypes = [vp, c_i32]
    L.pcf_catalog_size.restype = c_i64
    L.pcf_histogram.argtypes = [vp, ctypes.POINTER(PcfJob), vp, vp]
    L.pcf_histogram_noalloc.argtypes = [vp, ctypes.POINTER(PcfJob), vp, vp]
    L.pcf_selftest_scan.argtypes = [vp, vp, c_i64, ctypes.POINTER(c_i32)]
    L.pcf_last_stats.argtypes = [vp, ctypes.POINTER(PcfStats)]
    L.pcf_analytic_rr.argtypes = [vp, c_i32, c_i64, ctypes.c_float, ctypes.c_float, vp]
    L.pcf_histogram_batch.argtypes = [vp, ctypes.POINTER(PcfJob), c_i32, vp, vp]
    L.pcf_autotune.argtypes = [vp, ctypes.POINTER(PcfJob), c_i32, ctypes.POINTER(PcfTuning)]
    L.pcf_set_ref_bugs.argtypes = [vp, c_i32]
    L.pcf_tune_save.argtypes = [vp, ctypes.c_char_p]
    L.pcf_tune_load.argtypes = [vp, ctypes.c_char_p]
    L.pcf_tune_lookup.argtypes = [vp, ctypes.POINTER(PcfJob), c_i32, ctypes.POINTER(PcfTuning)]
    L.pcf_analytic_3iso.argtypes = [vp, c_i32, c_i32, ctypes.c_float, ctypes.c_float, c_i64,
                                    c_i32, c_i32, vp, vp, vp, vp]
    L.pcf_estimator.argtypes = [vp, c_i32, c_i64, vp, vp, vp, vp, ctypes.c_double,
                                ctypes.c_double, c_i32, vp]
    L.pcf_set_hbm_budget.argtypes = [vp, c_i64, c_i32]
    L.pcf_set_max_batch.argtypes = [vp, c_i32]
    L.pcf_submit.argtypes = [vp, ctypes.POINTER(PcfJob), ctypes.POINTER(c_i64)]
    L.pcf_poll.argtypes = [vp, c_i64, ctypes.POINTER(c_i32)]
    L.pcf_wait.argtypes = [vp, c_i64, vp, vp, ctypes.POINTER(PcfStats)]
    _lib = L
    return L


def hist_size(type_, bins):
    return {PCF_2ISO: bins, PCF_2ANI: bins ** 2, PCF_3ISO: bins ** 3,
            PCF_3ANI: bins ** 5}[type_]


def make_job(type_, hist, *, bins, dmax, cat_x, cat_y=-1, bpc=False, box=0.0,
             precision=PCF_PREC_MIXED, cells_per_rmax=0, partitions=35, rebuild_grid=False,
             lds_copies_log2=0, blocks_per_cu=0, pair_kernel=0):
    j = PcfJob()
    j.type = int(type_)
    j.hist = int(hist)
    j.bpc = int(bool(bpc))
    j.bins = int(bins)
    j.dmax = float(dmax)
    j.box = float(box)
    j.precision = int(precision)
    j.cat_x = int(cat_x)
    j.cat_y = int(cat_y)
    j.partitions = int(partitions)
    j.cells_per_rmax = int(cells_per_rmax)
    j.rebuild_grid = int(bool(rebuild_grid))
    j.lds_copies_log2 = int(lds_copies_log2)
    j.blocks_per_cu = int(blocks_per_cu)
    j.pair_kernel = int(pair_kernel)
    return j


class Engine:
    """One pcf_ctx on one HIP device."""

    def __init__(self, device=0):
        self.lib = load_library()
        self.ctx = ctypes.c_void_p()
        rc = self.lib.pcf_create(ctypes.byref(self.ctx), int(device))
        if rc != 0:
            raise PcfError(rc, self.lib.pcf_create_error().decode())

    def close(self):
        if self.ctx:
            self.lib.pcf_destroy(self.ctx)
            self.ctx = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _check(self, rc):
        if rc != 0:
            raise PcfError(rc, self.lib.pcf_last_error(self.ctx).decode())

    def upload(self, xyzw):
        a = np.ascontiguousarray(xyzw, dtype=np.float32)
        if a.ndim != 2 or a.shape[1] != 4:
            raise ValueError("catalog must be (n, 4) float32: x, y, z, w")
        cid = ctypes.c_int32()
        self._check(self.lib.pcf_catalog_upload(self.ctx, a.ctypes.data, len(a),
                                                ctypes.byref(cid)))
        return cid.value

    def upload_device(self, dev_ptr, n):
        cid = ctypes.c_int32()
        self._check(self.lib.pcf_catalog_upload_device(self.ctx, ctypes.c_void_p(dev_ptr), n,
                                                       ctypes.byref(cid)))
        return cid.value

    def release(self, cid):
        self._check(self.lib.pcf_catalog_release(self.ctx, cid))

    def histogram(self, job):
        n = hist_size(job.type, job.bins)
        c = np.zeros(n, np.uint64)
        w = np.zeros(n, np.float64)
        self._check(self.lib.pcf_histogram(self.ctx, ctypes.byref(job), c.ctypes.data,
                                           w.ctypes.data))
        return c, w

    def histogram_noalloc(self, job):
        """pcf_histogram_noalloc: the same pair histogram with no device
        allocation (buffers in the library's static scratch array)."""
        n = hist_size(job.type, job.bins)
        c = np.zeros(n, np.uint64)
        w = np.zeros(n, np.float64)
        self._check(self.lib.pcf_histogram_noalloc(self.ctx, ctypes.byref(job), c.ctypes.data,
                                                   w.ctypes.data))
        return c, w

    def selftest_scan(self, counts):
        """The library's device prefix sum over `counts`: its total, or -1
        when the total is 2^31 or more (pcf_selftest_scan)."""
        a = np.ascontiguousarray(counts, dtype=np.int32)
        tot = ctypes.c_int32()
        self._check(self.lib.pcf_selftest_scan(self.ctx, a.ctypes.data, len(a),
                                               ctypes.byref(tot)))
        return tot.value

    def stats(self):
        s = PcfStats()
        self._check(self.lib.pcf_last_stats(self.ctx, ctypes.byref(s)))
        return s.as_dict()

    def analytic_rr(self, bins, np_, dmax, box):
        rr = np.zeros(bins, np.float64)
        self._check(self.lib.pcf_analytic_rr(self.ctx, bins, np_, dmax, box, rr.ctypes.data))
        return rr

    def estimator(self, hists, n_data=0.0, n_rand=0.0, normalize=False):
        """Landy-Szalay estimator on the device (pcf_estimator): hists is
        (DD, DR, RR) or (DDD, DDR, DRR, RRR), equal-length float64 arrays."""
        hs = [np.ascontiguousarray(h, dtype=np.float64).ravel() for h in hists]
        if len(hs) not in (3, 4) or any(len(h) != len(hs[0]) for h in hs):
            raise ValueError("estimator takes 3 (2PCF) or 4 (3PCF) equal-length histograms")
        order = 2 if len(hs) == 3 else 3
        xi = np.zeros(len(hs[0]), np.float64)
        h3 = hs[3].ctypes.data if order == 3 else None
        self._check(self.lib.pcf_estimator(self.ctx, order, len(xi), hs[0].ctypes.data,
                                           hs[1].ctypes.data, hs[2].ctypes.data, h3,
                                           float(n_data), float(n_rand),
                                           PCF_EST_NORMALIZE if normalize else 0,
                                           xi.ctypes.data))
        return xi

    def histogram_batch(self, jobs):
        """Several pair histograms in one launch (pcf_histogram_batch): a list
        of (counts, weights) per job."""
        n = len(jobs)
        arr = (PcfJob * n)(*jobs)
        off = np.cumsum([0] + [hist_size(j.type, j.bins) for j in jobs])
        c = np.zeros(off[-1], np.uint64)
        w = np.zeros(off[-1], np.float64)
        self._check(self.lib.pcf_histogram_batch(self.ctx, arr, n, c.ctypes.data, w.ctypes.data))
        return [(c[off[k]:off[k + 1]], w[off[k]:off[k + 1]]) for k in range(n)]

    def autotune(self, jobs):
        """Tune the pair kernels for one job or a batch (pcf_autotune); the
        choice is cached on the context and used by later calls."""
        if isinstance(jobs, PcfJob):
            jobs = [jobs]
        arr = (PcfJob * len(jobs))(*jobs)
        t = PcfTuning()
        self._check(self.lib.pcf_autotune(self.ctx, arr, len(jobs), ctypes.byref(t)))
        return t.as_dict()

    def set_ref_bugs(self, on=True):
        """Reproduce the reference's defects (pcf_set_ref_bugs)."""
        self._check(self.lib.pcf_set_ref_bugs(self.ctx, 1 if on else 0))

    def tune_save(self, path):
        """Write the context's tuning cache to a file (pcf_tune_save)."""
        self._check(self.lib.pcf_tune_save(self.ctx, os.fsencode(path)))

    def tune_load(self, path):
        """Merge a tuning cache file into the context (pcf_tune_load)."""
        self._check(self.lib.pcf_tune_load(self.ctx, os.fsencode(path)))

    def tune_lookup(self, jobs):
        """The cached tuning of one job or a batch, or None (pcf_tune_lookup)."""
        if isinstance(jobs, PcfJob):
            jobs = [jobs]
        arr = (PcfJob * len(jobs))(*jobs)
        t = PcfTuning()
        rc = self.lib.pcf_tune_lookup(self.ctx, arr, len(jobs), ctypes.byref(t))
        if rc < 0:
            self._check(rc)
        return t.as_dict() if rc == 1 else None

    def analytic_3iso(self, cat, bins, dmax, box, np_=0, ptt=0, bn_ref=0):
        """Analytic DDR (= DRR) and RRR of the periodic 3iso case plus the
        intermediate ff_av and ff_av_ref (pcf_analytic_3iso)."""
        ptt_e = ptt if ptt > 0 else 100
        ref_e = bn_ref if bn_ref > 0 else 200
        ddr = np.zeros(bins ** 3, np.float64)
        rrr = np.zeros(bins ** 3, np.float64)
        ff = np.zeros(bins, np.float64)
        ffr = np.zeros(bins * ref_e, np.float64)
        self._check(self.lib.pcf_analytic_3iso(self.ctx, cat, bins, dmax, box, int(np_), ptt_e,
                                               ref_e, ddr.ctypes.data, rrr.ctypes.data,
                                               ff.ctypes.data, ffr.ctypes.data))
        return ddr, rrr, ff, ffr

    # ---- asynchronous (serving) ------------------------------------------
    def set_hbm_budget(self, nbytes, max_inflight=4):
        self._check(self.lib.pcf_set_hbm_budget(self.ctx, int(nbytes), int(max_inflight)))

    def set_max_batch(self, n):
        self._check(self.lib.pcf_set_max_batch(self.ctx, int(n)))

    def submit(self, job):
        t = ctypes.c_int64()
        self._check(self.lib.pcf_submit(self.ctx, ctypes.byref(job), ctypes.byref(t)))
        return t.value

    def poll(self, ticket):
        d = ctypes.c_int32()
        self._check(self.lib.pcf_poll(self.ctx, ticket, ctypes.byref(d)))
        return bool(d.value)

    def wait(self, ticket, type_, bins):
        n = hist_size(type_, bins)
        c = np.zeros(n, np.uint64)
        w = np.zeros(n, np.float64)
        s = PcfStats()
        self._check(self.lib.pcf_wait(self.ctx, ticket, c.ctypes.data, w.ctypes.data,
                                      ctypes.byref(s)))
        return c, w, s.as_dict()


def read_catalog(path, n=None):
    """The reference's 4-column text catalog (readme.md:36-40)."""
    a = np.loadtxt(path, dtype=np.float32, ndmin=2, max_rows=n)
    if a.shape[1] != 4:
        raise ValueError(f"{path}: expected 4 columns, got {a.shape[1]}")
    return a
